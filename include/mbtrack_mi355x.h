/*
 * mbtrack-mi355x: C ABI of the MI355X multi-bunch tracking turn.
 *
 * One process drives one GPU.  Every bunch of the ring lives in HBM as
 * structure-of-arrays FP64 coordinates; the per-bunch MPI ranks of the
 * reference become thread blocks, and its per-turn MPI_Bcast "allgathers"
 * become device arrays read by the next kernel.
 *
 * All functions return MBT_OK (0) or a negative MBT_E* code; mbt_last_error()
 * returns a human-readable message for the calling thread's last failure.
 * No torch or HIP types appear in these signatures.
 *
 * Reference interfaces replaced (fracul/mtrack, file:line):
 *   mbt_config_read        read_input + setup_ring_parameters +
 *                          macrop_model_setup_parameters + e_beam_setup +
 *                          setup_tracking_parameters   src/input.c:136,1553,1775,1289,1698
 *   mbt_generate_bunch     weak_generate_bunch_distribution src/bunch_weak.c:172-292
 *   mbt_ring_create        worker_weakweak set-up   src/worker_weakweak.c:58-255
 *                          (Green functions src/transform_weak.c:289-471,
 *                          LR histograms + wake_phasor_init :839-952,
 *                          RW history src/worker_weakweak.c:203-219)
 *   mbt_ring_track         the worker turn loop src/worker_weakweak.c:259-394 for
 *                          all bunches at once: transform_weak_bunch_selffield
 *                          src/transform_weak.c:553-771, RF feedback
 *                          src/worker_weakweak.c:268-278 + src/feedback_rf.c:5-34,
 *                          transform_weak_bunch_optic src/transform_weak.c:21-192,
 *                          weak_bunch_calc_statistics src/statistics.c:131-157,
 *                          weak_bunch_calc_ampinv src/statistics.c:163-195,
 *                          transform_weak_bunch_RW_longrange_cyclic
 *                          src/transform_weak.c:780-833, fnp_ring_update +
 *                          construct_wake_phasor src/transform_weak.c:839-1048
 *   mbt_manager_weakweak   manager_weakweak + worker_weakweak as one process
 *                          src/manager_weakweak.c:14-350, src/worker_weakweak.c:19-471
 *   mbt_wofz               complex_errorfunc_wofz_algo680 src/physic.c:75-301
 */
#ifndef MBTRACK_MI355X_H
#define MBTRACK_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBT_OK 0
#define MBT_EINVAL -1  /* bad argument or configuration */
#define MBT_EIO -2     /* file could not be read or written */
#define MBT_EDEVICE -3 /* HIP runtime failure */
#define MBT_ENOMEM -4
#define MBT_EUNSUPPORTED -5

#define MBT_LON 0
#define MBT_HOR 1
#define MBT_VER 2

#define MBT_MAX_RES 16
#define MBT_MAX_LR 8
#define MBT_MAX_AHC 8
#define MBT_MAX_WAKEFILE 8
#define MBT_PATH 1024

/* Filling patterns: numerical values of the reference's filling_t, src/def.h:31-47 */
#define MBT_FILL_FROMFILE -1
#define MBT_FILL_UNIFORM 0

/* Stats layout: per bunch per turn 12 doubles in the order a worker sends them
 * (src/worker_weakweak.c:321-324): pos[3] pos_sigma[3] slope[3] slope_sigma[3],
 * each indexed LON, HOR, VER.  Ampinv: ampinv[3] ampinv_cm[3]. */
#define MBT_NSTAT 12
#define MBT_NAMP 6

typedef struct mbt_resonator {
  double Rs, wr, Qfactor;
  int plane;
} mbt_resonator;

typedef struct mbt_lr_resonator {
  double Rs, wr, Qfactor, detune;
  int m;
  unsigned Nturn, Nbu;
} mbt_lr_resonator;

typedef struct mbt_active_hc {
  double Vpeak, nHC, phi_aHC;
} mbt_active_hc;

/* Every input of the .conf file and every quantity derived from it.  Field names
 * are the reference's physics names (src/types.h:60-209, src/bunch.h:23-49). */
typedef struct mbt_config {
  /* [tracking] */
  char jobtitle[256];
  char input_filename[MBT_PATH];
  char work_path[MBT_PATH];
  char fill_filename[MBT_PATH];
  int filling, bunch_model;
  int TrackPlane[3];
  long NrevTot, NrevMon, NrevPotentialsOut;
  int EnableRW_short, EnableRW_short_LON, EnableRW_long, EnableFBII, EnableQuantum;
  int EnableIdealHC, EnableAmpinv_out, EnableDiffCurr, EnableResonators;
  double current_ratio;
  int Nmlt, NmultiT, triggerRW;
  int scan;
  long Nscan, NrevScan;
  double scan_start, scan_step;
  double s0, mult;
  /* [ring] inputs */
  double E0, Lc, rho0, ac, Je, Vrf0, frf, Iring, ApertureH, ApertureV;
  double QH0, QV0, CHH, CHV, CVH, CVV, Gzix, Gziz, emittanceH, couplbeta;
  double rhorw, beffL[2], beffH[2], beffV[2], fYokoyaV, fYokoyaH;
  double beta1[3], alpha1[3], gamma1[3], dispH1, disppH1;
  double U0, taue;
  /* [ring] derived (src/input.c:1553-1695) */
  unsigned Nharm;
  double h, R, Gamma, Gamma2, Gamma3, wrf, T0, w0, Urad, De, tauC, aexpe;
  double q, Fq, epsMax, Vrfp, wso, wso2, fso, Qso, wgziH, wgziV, sn0, phai0, fc1, fc12;
  double Jx, Jz, taux, tauz;
  int AmpDependdQH, AmpDependdQV;
  int m_aHC;
  double phi_n, HC_k;
  /* impedances */
  int n_res;
  mbt_resonator res[MBT_MAX_RES];
  int n_wakefile;
  char wakefile[MBT_MAX_WAKEFILE][MBT_PATH];
  int wakefile_plane[MBT_MAX_WAKEFILE];
  int n_lr[3];
  mbt_lr_resonator lr[3][MBT_MAX_LR];
  unsigned Nbumax;
  int has_rf_feedback, rffb_resonator, rffb_len_average;
  double rffb_vrf_design, rffb_phi0_design;
  int n_ahc;
  mbt_active_hc ahc[MBT_MAX_AHC];
  /* [selffield] */
  int Ncell;
  double Nsigma, dT, sigma_tau;
  double RsisL, aindL, RsisV, aindV, RsisH, aindH;
  /* [distribution] */
  int Nslc, Np;
  double fNp;
  int modeHT, mode_excitation;
  int nGen[3], iseed[3];
  char gendst_datafile[3][MBT_PATH];
  double xtauCM_offset, xepsCM_offset, sgm_xtau, sgm_xeps, corr_epstau;
  double xxCM_offset, xpCM_offset, sgm_xx, sgm_xp, corr_xpx;
  double zzCM_offset, zpCM_offset, sgm_zz, sgm_zp, corr_zpz;
  double sgmatau;
  double pos_offset[3], slope_offset[3], pos_sgm[3], slope_sgm[3], correlate[3];
  /* filling (src/input.c:1289-1512) and currents (src/manager_weakweak.c:58-93) */
  int Nbunch;
  int *nfFill;      /* [Nharm] */
  double *Ib_frac;  /* [Nbunch], by filled-bunch ordinal */
  double *Ibunch;   /* [Nharm] bunch current [mA] */
  double *bunch_Ib; /* [Nbunch] bunch current [A] */
  int *bucket;      /* [Nbunch] bucket of the k-th filled bunch */
  int Nbunch_out;
  int *bunch_out;
} mbt_config;

const char *mbt_last_error(void);

/* ---- host configuration (C) ------------------------------------------------ */
/* Reads a .conf file (read_conf_file, src/input.c:278-850) or a legacy .inp
 * file (read_input_file, src/input.c:198-276), chosen by the extension as the
 * reference does, then derives every machine parameter. */
int mbt_config_read(const char *conf_path, const char *work_path, mbt_config *cfg);
void mbt_config_free(mbt_config *cfg);
/* Heap-allocated configuration for bindings; mbt_config_release frees both. */
mbt_config *mbt_config_alloc(void);
void mbt_config_release(mbt_config *cfg);
/* Scalar fields by name (e.g. "Nbunch", "T0", "TrackPlane_HOR"). */
int mbt_config_get(const mbt_config *cfg, const char *name, double *value);
int mbt_config_set(mbt_config *cfg, const char *name, double value);
/* Bucket of the k-th filled bunch, -1 if out of range. */
int mbt_config_bucket(const mbt_config *cfg, int k);
/* One-key lookups with the reader's grammar; 1 found, 0 absent or unparsable, <0 error.
 * Replace config_get_double/_int/_str and config_have_section, src/input.c:972-1100. */
int mbt_conf_get_double(const char *path, const char *section, const char *key, double *value);
int mbt_conf_get_int(const char *path, const char *section, const char *key, int *value);
int mbt_conf_get_str(const char *path, const char *section, const char *key, char *value, int len);
int mbt_conf_have_section(const char *path, const char *section);
/* Writes every scalar of the configuration as "name value" (%.17g), one per line. */
int mbt_config_dump(const mbt_config *cfg, const char *path);
/* Green functions Gl/Gv/Gh[Ncell] for the configuration (src/transform_weak.c:289-471);
 * counts of contributing impedances per plane in planes[3] (PlaneL, PlaneH, PlaneV order LON,HOR,VER). */
int mbt_greens_functions(const mbt_config *cfg, double *Gl, double *Gv, double *Gh, int planes[3]);
/* Initial coordinates of bucket kb, SoA: out[0..Np) tau, [Np..2Np) delta, x, xp, z, zp. */
int mbt_generate_bunch(const mbt_config *cfg, int kb, double *out);
/* All filled bunches into out[Nbunch][6][Np] (host threads over bunches). */
int mbt_generate_all(const mbt_config *cfg, double *out);

/* ---- the ring on the GPU ---------------------------------------------------- */
typedef struct mbt_ring mbt_ring;

/* flags */
#define MBT_RING_NO_INIT_DIST 1 /* leave coordinates zero; caller uploads them */
#define MBT_RING_NO_GRAPHS 2    /* launch every kernel of every turn eagerly instead of replaying
                                   HIP graphs of whole blocks of turns */
#define MBT_RING_NO_RW_GEMM 4   /* ensembles: keep the per-bunch resistive-wall sum even where the
                                   rings share one tap table (otherwise one FP64 MFMA GEMM) */
#define MBT_RING_MIXED 8        /* mixed precision: wake (selffield) convolutions in FP32 with
                                   hi/lo compensation (double-float sums) and the
                                   resistive-wall GEMM of ensembles (where the FP64 build uses
                                   one) in FP32 MFMA with hi/lo operand splitting and FP64 sums
                                   of FP32 chunks; a lone ring keeps its per-bunch FP64 RW sum,
                                   which is latency-bound, unless MBT_RING_RW_GEMM is set;
                                   coordinates, phase advance and the
                                   LR / RF-feedback recurrence stay FP64 */
#define MBT_RING_RW_OVERLAP 16  /* resistive-wall sum without the GEMM: always split it into the
                                   taps of earlier turns' rows, summed by the leading blocks of
                                   the turn's map, and the taps of the row the turn writes,
                                   summed after the reduction.  Without the flag only rings of
                                   at least 4,194,304 macroparticles split it; below that one
                                   fused pass after the reduction is faster */
#define MBT_RING_RW_GEMM 32     /* the GEMM form of the resistive-wall sum (FP64 MFMA, or FP32
                                   MFMA with hi/lo splitting under MBT_RING_MIXED) from one ring
                                   on, where by default it needs 8 rings sharing one tap table;
                                   overrides MBT_RING_RW_OVERLAP */
#define MBT_RING_HOST_INIT_DIST 64 /* make the initial distributions on the host and upload them;
                                   by default rings whose distribution is the reference's Gaussian
                                   generator with pairwise different plane seeds get theirs from a
                                   device kernel, bit for bit the same */

int mbt_ring_create(const mbt_config *cfg, int device, int flags, mbt_ring **out);
void mbt_ring_destroy(mbt_ring *r);
/* Replace / read the coordinates of filled bunch ordinal ib (SoA layout as
 * mbt_generate_bunch), as the bunch stands at the end of a turn: the device
 * applies a turn's resistive-wall kick (src/transform_weak.c:780-833) at the
 * start of the next pass, so _get_bunch adds the kick still pending and
 * _set_bunch drops the bunch's pending kick.  Between tracked turns the next
 * turn's wake tables come from the slice histograms of the coordinates that
 * were tracked; replace bunches before mbt_ring_prepare. */
int mbt_ring_set_bunch(mbt_ring *r, int ib, const double *soa6);
int mbt_ring_get_bunch(const mbt_ring *r, int ib, double *soa6);
/* Finishes set-up after coordinates are final: initial statistics, LR-resonator
 * phasor initialisation, RF-feedback design solve and RW history fill.
 * init_stats receives Nbunch*12 doubles (may be NULL). */
int mbt_ring_prepare(mbt_ring *r, double *init_stats);
/* Tracks nturns turns.  stats: nturns*Nbunch*12 doubles, ampinv: nturns*Nbunch*6
 * (written on monitor turns only), rf: nturns*2 (phai0, Vrf0 used by the turn's map).
 * Any output pointer may be NULL. */
int mbt_ring_track(mbt_ring *r, long nturns, double *stats, double *ampinv, double *rf);
/* mbt_ring_track split in two, so that a caller can do its per-turn host work
 * (the reference manager's files) while the next turns run: _async enqueues
 * nturns turns (1 .. mbt_ring_async_max(r), one block of the device's
 * statistics ring) and returns; _wait waits for them and copies their outputs
 * as mbt_ring_track does.  One block at a time: other tracking calls on the
 * ring fail with MBT_EINVAL while one is pending.  No counterpart in the
 * reference, whose manager receives each turn's statistics over MPI
 * (src/manager_weakweak.c:179-299). */
long mbt_ring_async_max(const mbt_ring *r);
int mbt_ring_track_async(mbt_ring *r, long nturns);
int mbt_ring_track_wait(mbt_ring *r, double *stats, double *ampinv, double *rf);
/* Statistics after the last pending RW kick (src/worker_weakweak.c:397-405). */
int mbt_ring_final_stats(mbt_ring *r, double *stats);
/* Selffield histogram of the last completed turn (clamped rule,
 * src/transform_weak.c:608-631) as counts[Nbunch*Ncell] plus trash counters [2*Nbunch]. */
int mbt_ring_histograms(const mbt_ring *r, double *counts, int *trash);
/* Selffield tables the next turn will use for bunch ordinal ib, 7 rows of Ncell:
 * factG*GL1, LR wake, slice counts, -factG*GlambdaV, dipoleV, -factG*GlambdaH,
 * dipoleH (the columns of potentials_bunch_<kb>.dat, src/transform_weak.c:735-768);
 * extra[3]: phase and voltage before the next turn's RF-feedback solve, scan value. */
int mbt_ring_potentials(const mbt_ring *r, int ib, double *out, double *extra);
/* RF state of ring `ring` (0 for a single ring) after the last tracked turn:
 * out[3] = phai0, Vrf0 [MV] the last turn's map used (before a turn has run:
 * the configured ones) and whether the RF feedback is still on -- the
 * worker's ring.phai0, ring.Vrf0 and ring.has_rf_feedback printed at the end of
 * the run (src/worker_weakweak.c:446-448). */
int mbt_ring_rf_state(const mbt_ring *r, int ring, double *out);
/* RW dipole history [(Nmlt+2)*h][2] in the reference's cyclic layout
 * (src/worker_weakweak.c:203-219). */
int mbt_ring_rw_history(const mbt_ring *r, double *out);
/* Kernel timing of the last mbt_ring_track call, milliseconds per kernel class:
 * [0] turn map pass, [1] per-bunch reduction, [2] LR recurrence, [3] RW sum. */
int mbt_ring_kernel_times(const mbt_ring *r, double *ms, long *launches);
/* Bytes of HBM the turn map pass must move per turn (algorithmic). */
double mbt_ring_bytes_per_turn(const mbt_ring *r);
/* No counterpart in the reference (a measurement aid): how many allocations of
 * the coordinates the set-up tried (rings of 16 GB of coordinates and more try
 * up to 3 and keep the fastest for the turn pass's traffic) and the in-place
 * stream rate in GB/s of the one kept (0 when the ring is too small to search). */
int mbt_ring_placement(const mbt_ring *r, int *candidates, double *gbps);
/* No counterpart in the reference (a measurement aid): how many turn graphs
 * (one per block of turns, reused for later blocks whose first monitor turn
 * sits at the same offset) the ring has captured and instantiated so far. */
int mbt_ring_graph_builds(const mbt_ring *r, long *count);
/* Per-kernel timing with HIP events around every launch group (0 = off). */
int mbt_ring_set_timing(mbt_ring *r, int on);
long mbt_ring_turn(const mbt_ring *r);
/* Work decomposition of the per-macroparticle pass: chunks of 256 * items
 * macroparticles of one bunch per work-group, items in [1, 64] for rings (all
 * rings of an ensemble together) of fewer than 4,194,304 macroparticles and in
 * [4, 64] otherwise.  The default is 8, halved while that leaves fewer than
 * 2,048 work-groups.  No counterpart in the reference (a tuning knob of this
 * build); changes results only at the rounding level (grouping of the moment
 * sums). */
int mbt_ring_set_chunk(mbt_ring *r, int items);
int mbt_ring_get_chunk(const mbt_ring *r);
/* Times the ring's own per-macroparticle turn kernel for items = 1, 2, 4, 8,
 * 16, 32, 64 (1 and 2 only below 4,194,304 macroparticles) -- the map, kicks,
 * histogram and moment sums computed, the coordinates written back unchanged,
 * so the ring's state is untouched -- keeps the fastest and returns it in
 * *best_items; ms[7] (may be NULL) receives each candidate's time, 0 for those
 * not tried.  Call after mbt_ring_prepare. */
int mbt_ring_autotune(mbt_ring *r, int *best_items, double *ms);
/* No counterpart in the reference (a measurement aid): the rate in GB/s of a
 * plain in-place update of the ring's own coordinates (six FP64 values, 96
 * bytes per slot, every bit written back unchanged, so the ring's state is
 * untouched), averaged over reps passes after one untimed pass -- the ceiling
 * their layout and placement in device memory set for the per-macroparticle
 * pass's traffic. */
int mbt_ring_stream_ceiling(mbt_ring *r, int reps, double *gbps);

/* ---- ensembles of independent rings on one GPU ------------------------------ */
/* No counterpart in the reference, which runs one ring per MPI job: nrings
 * independent configurations (their own bunch counts, macroparticle counts,
 * currents, optics, wakes and RF feedback) tracked by the same kernels, every
 * launch covering all rings.  A single ring is an ensemble of one, and every
 * mbt_ring_* call above accepts an ensemble: bunch ordinals, stats and ampinv rows
 * run over the rings' bunches in ring order (ring 0 first), rf rows hold
 * nrings (phai0, Vrf0) pairs.  The rings must share Ncell and NrevMon. */
int mbt_ensemble_create(const mbt_config *const *cfgs, int nrings, int device, int flags, mbt_ring **out);
/* Coordinates of global bunch ordinal g (the ring's own Np, SoA as mbt_generate_bunch). */
int mbt_ensemble_set_bunch(mbt_ring *r, int g, const double *soa6);
int mbt_ensemble_get_bunch(const mbt_ring *r, int g, double *soa6);
/* Number of rings, total bunches and total macroparticles. */
int mbt_ensemble_size(const mbt_ring *r, int *nrings, int *nbunch_total, long long *nparticles);

/* ---- weak-strong turn with fast beam-ion interaction (BunchModel = STRONG) --- */
/* Slice-centroid bunches (2 Nslc + 1 slices each) with the FBII ion train and the
 * resistive-wall kick on the centroid history.  Replaces manager_weakstrong /
 * worker_weakstrong (src/manager_weakstrong.c:7-263, src/worker_weakstrong.c:14-247):
 * strong_generate_eCMs + generate_initial_general_distribution2
 * (src/bunch_strong.c:10-216), transformSbunch_shortrange / _optic / _longrange
 * (src/rw_weakstrong.c:10-253), fbi_weakstrong_bunch / _train
 * (src/fbi_weakstrong.c:49-451), bunch_strong_distribution_statistics
 * (src/bunch.c:516-550). */
typedef struct mbt_ws mbt_ws;
/* flags: bunch currents of the filling (the weak-weak manager's rule,
 * src/manager_weakweak.c:58-80) instead of the zeros the reference leaves in
 * ring.Ibunch in this mode (which make the ions chargeless) */
#define MBT_WS_BUNCH_CURRENTS 1
/* flags: run the ion train as the reference's relay, one launch per bunch in
 * train order, instead of one lane per ion over all later bunches (the default;
 * the relay is also taken when the default's per-bunch partials would exceed
 * 8 GiB).  Both give the same sums up to their summation order. */
#define MBT_WS_SERIAL_TRAIN 2
/* flags: mixed precision -- the Faddeeva function of the Bassetti-Erskine field
 * tables in FP32; centroids, ions and the field arithmetic around it stay FP64 */
#define MBT_WS_MIXED 4
int mbt_ws_create(const mbt_config *cfg, int device, int flags, mbt_ws **out);
void mbt_ws_destroy(mbt_ws *w);
/* nturns turns; per turn (any output may be NULL):
 *   cm[t][ib][4]    x, x', z, z' centroid each bunch reports after its synchrotron step
 *   mon[t][ib][18]  on monitor turns (turn % NrevMon == 0): pos[3] pos_sigma[3]
 *                   slope[3] slope_sigma[3] ampinv[3] ampinv_cm[3], else 0
 *   ion[t][ib][3]   ion cloud after bunch ib's passage: sum Nri x, sum Nri y, sum Nri
 *   zcm[t][kb]      slice-averaged z of every bucket at the end of the turn */
int mbt_ws_track(mbt_ws *w, long nturns, double *cm, double *mon, double *ion, double *zcm);
/* mbt_ws_track split in two, so that a caller can write the previous turns'
 * files while the next turns run: _async enqueues nturns turns
 * (1 .. mbt_ws_async_max(w), one block of the device's output buffers) and
 * returns; _wait waits for them and copies their outputs as mbt_ws_track does.
 * One block at a time: mbt_ws_track and _async fail with MBT_EINVAL while one
 * is pending, and so does _async with train timing on; mbt_ws_get_slices
 * waits for the pending turns.  No counterpart in the reference, whose manager
 * receives each turn's slices over MPI (src/manager_weakstrong.c:77-248). */
long mbt_ws_async_max(const mbt_ws *w);
int mbt_ws_track_async(mbt_ws *w, long nturns);
int mbt_ws_track_wait(mbt_ws *w, double *cm, double *mon, double *ion, double *zcm);
/* slice arrays of filled bunch ordinal ib: out[6][S] = tau, delta, x, x', z, z' */
int mbt_ws_get_slices(const mbt_ws *w, int ib, double *out);
/* macro-electrons per slice npop[S]; *nslices = S */
int mbt_ws_population(const mbt_ws *w, int *npop, int *nslices);
/* the generated macro-electrons dist[6][Np] */
int mbt_ws_distribution(const mbt_ws *w, double *dist);
long mbt_ws_turn(const mbt_ws *w);
/* host set-up pieces (no device): FBII local grid spacing; slice centroids of
 * every bucket slices[h][6][S], npop[S], grid counts nofe[S][5][5], dist[6][Np] */
void mbt_ws_grid(const mbt_config *cfg, double *agLOCx, double *agLOCz);
int mbt_ws_generate(const mbt_config *cfg, double *slices, int *npop, int *nofe, double *dist);
/* No counterpart in the reference (measurement aids): HIP events on the object's
 * stream around the ion train of every turn (k_ws_ions, or the relay's launches)
 * while on; mbt_ws_train_time returns the summed milliseconds and the number of
 * trains of the last mbt_ws_track call. */
int mbt_ws_set_timing(mbt_ws *w, int on);
int mbt_ws_train_time(const mbt_ws *w, double *ms, long *launches);
/* manager_weakstrong + its workers on one GPU: the reference's files in cfg->work_path */
int mbt_manager_weakstrong(const mbt_config *cfg, int device, int flags);

/* ---- end-to-end drop-in ----------------------------------------------------- */
/* Runs manager_weakweak + all workers of the reference on one GPU and writes the
 * reference's output files into cfg->work_path. */
int mbt_manager_weakweak(const mbt_config *cfg, int device);

/* ---- Faddeeva function on the GPU ------------------------------------------- */
/* w(z) for n points (x[i] + i y[i]); FP64. Returns status per point in flag (0 ok,
 * 1 overflow, as complex_errorfunc_wofz_algo680 src/physic.c:75-301). */
int mbt_wofz(int device, long n, const double *x, const double *y, double *wr, double *wi, int *flag);
/* The same in FP32 (the mixed-precision beam-ion field); no reference counterpart. */
int mbt_wofz_f32(int device, long n, const float *x, const float *y, float *wr, float *wi, int *flag);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mbt_wofz with its device staging in a fixed array of the
 * library (9 MiB per device, held until the process ends) instead of hipMalloc'ed
 * buffers; at most 262,144 points per call (more: MBT_EINVAL, nothing computed).
 * Calls that overlap on one device share that array and race. */
int mbt_wofz_noalloc(int device, long n, const double *x, const double *y, double *wr, double *wi, int *flag);

#ifdef __cplusplus
}
#endif
#endif
