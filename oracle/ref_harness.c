/*
 * Reference harness: test infrastructure only.
 *
 * Links the reference's own translation units (built in place from
 * /root/reference/src by oracle/Makefile, never copied) and replays the
 * weak-weak worker turn of worker_weakweak.c:259-394 for every filled bucket
 * inside one process, in the phase order the MPI program imposes:
 *
 *   phase 1  per bunch: selffield kick, RF-feedback solve, one-turn map,
 *            statistics                        (worker_weakweak.c:264-303)
 *   phase 2  per bunch: RW dipole history write (worker_weakweak.c:306-318)
 *   phase 3  per bunch: ampinv on monitor turns, RW long-range kicks
 *                                              (worker_weakweak.c:327-354)
 *   phase 4  per bunch: LR-resonator histogram update (:356-386)
 *
 * Every MPI_Bcast of the reference is an allgather of identical data, so one
 * shared history array and one shared fnp_ring per plane hold exactly what each
 * rank would hold.  Per-rank state that differs (ring_t copy with its RF
 * feedback, phasors, selffield model, RW trigger counter) is kept per bunch.
 *
 * Output (all %.17g):
 *   stats.txt   : turn kb pos[3] pos_sigma[3] slope[3] slope_sigma[3]
 *                 (turn -1 = initial distribution), one line per bunch per turn
 *   ampinv.txt  : turn kb ampinv[3] ampinv_cm[3] on monitor turns
 *   rf.txt      : turn phai0 Vrf0 (rank of the first filled bucket)
 *   hist.txt    : turn kb n_trash_low n_trash_high counts[Ncell] (drop rule of
 *                 fnp_ring_update, transform_weak.c:839-865) when dump_hist
 *   parts_<kb>.txt : final particle coordinates tau delta x xp z zp
 *
 * Usage: refharness input.conf outdir [nturns] [dump_hist] [dump_parts_kb]
 * Known single-process caveat: glibc rand() state and the static box_muller
 * cache (bunch.c:407-436) are process-wide; with an even Np the cache is empty
 * at every bunch boundary and each bunch reseeds, so the replay equals the MPI
 * run.  Configurations with odd Np and quantum excitation are not replayed.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <time.h>
#include <omp.h>
#include "types.h"
#include "bunch.h"
#include "input.h"
#include "transform_weak.h"
#include "statistics.h"
#include "feedback_rf.h"
#include "cyclic_array.h"
#include "tracking.h"
#include "fbi.h"

ring_t ring;
tracking_t track;
grid_t fbii_grid;

typedef struct {
  int kb;
  ring_t ring;
  rf_feedback_t rffb;
  int fb_off;
  weak_bunch_t bunch;
  selffield_model_t sf;
  double phasor_end[32], phasor_end_HOR[32], phasor_end_VER[32];
  int m;
  double scan_val;
  FILE *trafo_fp;
} worker_t;

static void dump_stats(FILE *fp, long turn, int kb, const bunch_stats_t *s)
{
  fprintf(fp, "%ld %d", turn, kb);
  for (int p = 0; p < 3; p++) fprintf(fp, " %.17g", s->pos.v[p]);
  for (int p = 0; p < 3; p++) fprintf(fp, " %.17g", s->pos_sigma.v[p]);
  for (int p = 0; p < 3; p++) fprintf(fp, " %.17g", s->slope.v[p]);
  for (int p = 0; p < 3; p++) fprintf(fp, " %.17g", s->slope_sigma.v[p]);
  fprintf(fp, "\n");
}

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: refharness input.conf outdir [nturns] [dump_hist] [dump_parts_kb]\n");
    return 2;
  }
  long nturns_override = argc > 3 ? atol(argv[3]) : -1;
  int dump_hist = argc > 4 ? atoi(argv[4]) : 0;
  int dump_parts_kb = argc > 5 ? atoi(argv[5]) : -1;
  /* timing: the first serial_turns turns on one thread, the rest on all */
  long serial_turns = argc > 6 ? atol(argv[6]) : 0;
  char conf[FILENAME_MAX];
  snprintf(conf, sizeof conf, "%s", argv[1]);
  snprintf(track.work_path, FILENAME_MAX, "%s", argv[2]);

  bunch_macroparticle_model_t mp;
  selffield_model_t sf0;
  e_beam_t ebeam;
  memset(&mp, 0, sizeof mp);
  memset(&sf0, 0, sizeof sf0);
  memset(&ebeam, 0, sizeof ebeam);
  if (!read_input(conf, &ring, &track, &ebeam, &mp, &sf0)) return 1;
  /* Q1 of the reference: with RW long-range alone beffH3/beffV3 are never read
     (src/input.c:465-478) and the kick divides by zero.  Read them with the
     reference's own accessor, as the project's reader does. */
  if (track.EnableRW_long && !track.EnableRW_short) {
    FILE *cf = fopen(conf, "r");
    config_get_fprint = false;
    if (track.TrackPlane[HOR]) config_get_double(cf, "ring", "beffH3", &ring.beffH[0]);
    if (track.TrackPlane[VER]) config_get_double(cf, "ring", "beffV3", &ring.beffV[0]);
    fclose(cf);
  }
  setup_ring_parameters(&ring);
  macrop_model_setup_parameters(ring, track, &mp);
  fbii_grid_setup_parameters(&fbii_grid, &ebeam, mp);
  if (!e_beam_setup(&track, &ring, &ebeam)) return 1;
  if (!setup_tracking_parameters(&ring, &track, &sf0, &mp)) return 1;
  if (nturns_override > 0) track.NrevTot = nturns_override;
  if (track.scan == 0) track.NrevScan = track.NrevTot + 1;

  {
    /* derived parameters, same names as the project's mbt_config_dump */
    char pp[FILENAME_MAX + 64];
    snprintf(pp, sizeof pp, "%s/params.txt", track.work_path);
    FILE *fpp = fopen(pp, "w");
#define PD(name, v) fprintf(fpp, "%s %.17g\n", name, (double)(v))
    PD("Nharm", ring.Nharm); PD("h", ring.h); PD("T0", ring.T0); PD("w0", ring.w0); PD("wrf", ring.wrf);
    PD("R", ring.R); PD("Gamma", ring.Gamma); PD("Gamma2", ring.Gamma2); PD("Gamma3", ring.Gamma3);
    PD("U0", ring.U0); PD("Urad", ring.Urad); PD("De", ring.De); PD("taue", ring.taue); PD("tauC", ring.tauC);
    PD("aexpe", ring.aexpe); PD("Je", ring.Je); PD("rho0", ring.rho0); PD("q", ring.q); PD("Fq", ring.Fq);
    PD("epsMax", ring.epsMax); PD("Vrfp", ring.Vrfp); PD("wso", ring.wso); PD("wso2", ring.wso2);
    PD("fso", ring.fso); PD("Qso", ring.Qso); PD("wgziH", ring.wgziH); PD("wgziV", ring.wgziV);
    PD("sn0", ring.sn0); PD("phai0", ring.phai0); PD("Vrf0", ring.Vrf0); PD("fc1", ring.fc1); PD("fc12", ring.fc12);
    PD("Jx", ring.Jx); PD("Jz", ring.Jz); PD("taux", ring.taux); PD("tauz", ring.tauz);
    PD("gamma1[1]", ring.gamma1[1]); PD("gamma1[2]", ring.gamma1[2]); PD("Gzix", ring.Gzix); PD("Gziz", ring.Gziz);
    PD("CHH", ring.CHH); PD("CHV", ring.CHV); PD("CVH", ring.CVH); PD("CVV", ring.CVV);
    PD("AmpDependdQH", ring.AmpDependdQH); PD("AmpDependdQV", ring.AmpDependdQV);
    PD("m_aHC", ring.m_aHC); PD("phi_n", ring.phi_n); PD("HC_k", ring.HC_k); PD("Nbumax", ring.Nbumax);
    PD("rhorw", ring.rhorw); PD("beffH[0]", ring.beffH[0]); PD("beffV[0]", ring.beffV[0]);
    PD("NrevTot", track.NrevTot); PD("NrevMon", track.NrevMon); PD("NrevScan", track.NrevScan);
    PD("Nscan", track.Nscan); PD("scan", track.scan); PD("scan_start", track.scan_start);
    PD("scan_step", track.scan_step); PD("Nmlt", track.Nmlt); PD("NmultiT", track.NmultiT);
    PD("triggerRW", track.triggerRW); PD("EnableRW_long", track.EnableRW_long);
    PD("EnableQuantum", track.EnableQuantum); PD("EnableIdealHC", track.EnableIdealHC);
    PD("Np", mp.Np); PD("iseed[0]", mp.iseed[0]); PD("iseed[1]", mp.iseed[1]); PD("iseed[2]", mp.iseed[2]);
    for (int i = 0; i < 3; i++) {
      char nm[32];
      snprintf(nm, sizeof nm, "pos_offset[%d]", i); PD(nm, mp.pos_offset.v[i]);
      snprintf(nm, sizeof nm, "slope_offset[%d]", i); PD(nm, mp.slope_offset.v[i]);
      snprintf(nm, sizeof nm, "pos_sgm[%d]", i); PD(nm, mp.pos_sgm.v[i]);
      snprintf(nm, sizeof nm, "slope_sgm[%d]", i); PD(nm, mp.slope_sgm.v[i]);
      snprintf(nm, sizeof nm, "correlate[%d]", i); PD(nm, mp.correlate.v[i]);
    }
    PD("Ncell", sf0.Ncell); PD("Nsigma", sf0.Nsigma); PD("Nbunch", ebeam.Nbunch); PD("Nbunch_out", track.Nbunch_out);
    PD("has_rf_feedback", ring.has_rf_feedback);
    if (ring.has_rf_feedback) {
      PD("rffb_vrf_design", ring.rf_feedback->vrf_design);
      PD("rffb_phi0_design", ring.rf_feedback->phi0_design);
    }
    for (int i = 0; i < track.Nbunch_out; i++) fprintf(fpp, "bunch_out %d\n", track.bunch_out[i]);
#undef PD
    fclose(fpp);
  }
  const int h = ring.Nharm;
  const int Nb = ebeam.Nbunch;
  /* manager: currents (manager_weakweak.c:58-93) */
  int *bnumbers = calloc(Nb, sizeof(int));
  double *bunch_Ib = calloc(Nb, sizeof(double));
  for (int i = 0, n = 0; i < h; i++)
    if (ebeam.nfFill[i]) bnumbers[n++] = i;
  for (int k = 0; k < Nb; k++) {
    bunch_Ib[k] = ebeam.Ib_frac[k] * ring.Iring / (((double)Nb) * FKILO);
    ring.Ibunch[bnumbers[k]] = bunch_Ib[k] * FKILO;
  }
  mp.fNp = (double)mp.Np;
  {
    char pp[FILENAME_MAX + 64];
    snprintf(pp, sizeof pp, "%s/params.txt", track.work_path);
    FILE *fpp = fopen(pp, "a");
    for (int i = 0; i < h; i++)
      if (ebeam.nfFill[i]) fprintf(fpp, "Ibunch_%d %.17g\n", i, ring.Ibunch[i]);
    fclose(fpp);
  }

  char path[FILENAME_MAX + 64];
  snprintf(path, sizeof path, "%s/stats.txt", track.work_path);
  FILE *fst = fopen(path, "w");
  snprintf(path, sizeof path, "%s/ampinv.txt", track.work_path);
  FILE *famp = fopen(path, "w");
  snprintf(path, sizeof path, "%s/rf.txt", track.work_path);
  FILE *frf = fopen(path, "w");
  FILE *fh = NULL;
  if (dump_hist) {
    snprintf(path, sizeof path, "%s/hist.txt", track.work_path);
    fh = fopen(path, "w");
  }
  if (!fst || !famp || !frf) return 1;

  worker_t *W = calloc(Nb, sizeof(worker_t));
  for (int k = 0; k < Nb; k++) {
    worker_t *w = &W[k];
    w->kb = bnumbers[k];
    w->ring = ring;
    w->ring.Ibunch = malloc(h * sizeof(double));
    memcpy(w->ring.Ibunch, ring.Ibunch, h * sizeof(double));
    if (ring.rf_feedback) {
      w->rffb = *ring.rf_feedback;
      w->ring.rf_feedback = &w->rffb;
    }
    w->sf = sf0;
    w->scan_val = track.scan_start;
    if (!weak_bunch_create(&w->bunch, &track, mp.Np, bunch_Ib[k], w->kb, true)) return 1;
    if (weak_generate_bunch_distribution(&w->bunch, mp, track.TrackPlane) < 0) return 1;
    if (!selffield_model_init(&w->sf, mp.pos_sgm.xtau, &w->bunch)) return 1;
    if (w->ring.resonators_size > 0) construct_greensfunc_resonator(&w->sf, &w->ring);
    if (w->ring.wakefiles_size > 0) construct_greensfunc_file(&w->sf, &w->ring);
    if (track.EnableRW_short || track.EnableRW_short_LON) construct_greensfunc_RW(&w->sf, &w->ring);
    if (k == 0 && w->sf.Ncell > 0) {
      /* the reference's Green functions (identical for every bunch) */
      snprintf(path, sizeof path, "%s/greens.txt", track.work_path);
      FILE *fg = fopen(path, "w");
      if (!fg) return 1;
      for (int i = 0; i < w->sf.Ncell; i++)
        fprintf(fg, "%d %.17g %.17g %.17g\n", i, w->sf.Gl[i], w->sf.Gv[i], w->sf.Gh[i]);
      fprintf(fg, "planes %d %d %d\n", w->sf.PlaneL, w->sf.PlaneH, w->sf.PlaneV);
      fclose(fg);
    }
    weak_bunch_calc_statistics(&w->bunch, &track);
    dump_stats(fst, -1, w->kb, &w->bunch.stats);
    if (w->bunch.kb_out == 1 && track.NrevPotentialsOut <= track.NrevTot) {
      snprintf(path, sizeof path, "%s/potentials_bunch_%d.dat", track.work_path, w->kb);
      w->trafo_fp = fopen(path, "w");
    }
    /* statistics accumulators of kb_out ranks restart after the header
       (worker_weakweak.c:233-234); others keep accumulating, unused */
    if (w->bunch.kb_out == 1) weak_bunch_reset_statistics(&w->bunch.stats);
  }

  const int Nbin = sf0.Ncell;
  double *fnp_ring = NULL, *fnp_HOR = NULL, *fnp_VER = NULL;
  int nlrL = ring.longrange_resonators_size[LON];
  int nlrH = ring.longrange_resonators_size[HOR];
  int nlrV = ring.longrange_resonators_size[VER];
  if (nlrL > 0) {
    fnp_ring = calloc((size_t)Nbin * h, sizeof(double));
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      if (w->ring.has_rf_feedback) {
        if (w->ring.rf_feedback->len_average < 0) {
          w->ring.rf_feedback->len_average = -w->ring.rf_feedback->len_average;
          w->fb_off = 1;
        }
        rffb_init(w->ring.rf_feedback);
      }
    }
    for (int k = 0; k < Nb; k++)
      fnp_ring_update(&W[k].ring, fnp_ring, &W[k].sf, &W[k].bunch, W[k].kb, LON);
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      wake_phasor_init(&w->ring, fnp_ring, &w->sf, &w->bunch, w->kb, w->phasor_end, &ebeam, LON);
      if (w->ring.has_rf_feedback) {
        double vmbar, phibar;
        rffb_calc_mean_voltage_phase(w->ring.rf_feedback, &vmbar, &phibar);
        vmbar = vmbar * w->ring.E0 * FKILO;
        double num = w->ring.rf_feedback->vrf_design * sin(w->ring.rf_feedback->phi0_design) - vmbar * sin(phibar);
        w->ring.phai0 = atan2(num, w->ring.rf_feedback->vrf_design * cos(w->ring.rf_feedback->phi0_design) - vmbar * cos(phibar));
        w->ring.Vrf0 = num / sin(w->ring.phai0);
        if (w->fb_off) w->ring.has_rf_feedback = 0;
      }
    }
  }
  if (track.TrackPlane[HOR] && nlrH > 0) {
    fnp_HOR = calloc((size_t)Nbin * h, sizeof(double));
    for (int k = 0; k < Nb; k++)
      fnp_ring_update(&W[k].ring, fnp_HOR, &W[k].sf, &W[k].bunch, W[k].kb, HOR);
    for (int k = 0; k < Nb; k++)
      wake_phasor_init(&W[k].ring, fnp_HOR, &W[k].sf, &W[k].bunch, W[k].kb, W[k].phasor_end_HOR, &ebeam, HOR);
  }
  if (track.TrackPlane[VER] && nlrV > 0) {
    fnp_VER = calloc((size_t)Nbin * h, sizeof(double));
    for (int k = 0; k < Nb; k++)
      fnp_ring_update(&W[k].ring, fnp_VER, &W[k].sf, &W[k].bunch, W[k].kb, VER);
    for (int k = 0; k < Nb; k++)
      wake_phasor_init(&W[k].ring, fnp_VER, &W[k].sf, &W[k].bunch, W[k].kb, W[k].phasor_end_VER, &ebeam, VER);
  }

  cyclic_array_t dipole_RW;
  if (track.EnableRW_long > 0) {
    cyclic_array_create(&dipole_RW, (track.Nmlt + 2) * h, 2);
    for (int k = 0; k < Nb; k++) {
      double d[2] = {W[k].bunch.stats.pos.x, W[k].bunch.stats.pos.z};
      cyclic_array_set(W[k].kb, d, &dipole_RW);
    }
    for (int k = 0; k < Nb; k++) {
      double *mom = cyclic_array_access(W[k].kb, &dipole_RW);
      for (int j = 1; j < track.Nmlt; j++) cyclic_array_set(j * h + W[k].kb, mom, &dipole_RW);
    }
  }

  int lr_tot = nlrL + nlrH + nlrV;
  int *Nstat = calloc(Nb, sizeof(int));
  struct timespec t_a, t_b;
  clock_gettime(CLOCK_MONOTONIC, &t_a);
  /* The reference runs one MPI rank per bunch; here the per-bunch work of a turn
     runs on the host's cores (OpenMP threads over bunches, each bunch's own
     state only) wherever the reference's code touches no shared state: not with
     quantum excitation, whose glibc rand stream and Box-Muller cache are
     process-wide.  Files are written after each parallel loop, in bunch order. */
  const int qe_off = !track.EnableQuantum;
  struct timespec t_s;
  t_s = t_a;
  for (long rev = 0; rev < track.NrevTot; rev++) {
    const int par = qe_off && rev >= serial_turns;
    if (rev == serial_turns && serial_turns > 0) clock_gettime(CLOCK_MONOTONIC, &t_s);
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      if (w->sf.PlaneL + w->sf.PlaneV + w->sf.PlaneH + lr_tot > 0)
        transform_weak_bunch_selffield(&w->bunch, w->sf, NULL, &w->ring, rev, w->trafo_fp, w->scan_val,
                                       w->kb, &ebeam, w->phasor_end, w->phasor_end_HOR,
                                       w->phasor_end_VER, fnp_ring, fnp_HOR, fnp_VER);
      if (w->ring.has_rf_feedback) {
        double vmbar, phibar;
        rffb_calc_mean_voltage_phase(w->ring.rf_feedback, &vmbar, &phibar);
        vmbar = vmbar * w->ring.E0 * FKILO;
        double num = w->ring.rf_feedback->vrf_design * sin(w->ring.rf_feedback->phi0_design) - vmbar * sin(phibar);
        w->ring.phai0 = atan2(num, w->ring.rf_feedback->vrf_design * cos(w->ring.rf_feedback->phi0_design) - vmbar * cos(phibar));
        w->ring.Vrf0 = num / sin(w->ring.phai0);
      }
      int iseed = mp.iseed[LON] + w->bunch.kb;
      transform_weak_bunch_optic(&w->bunch, &mp, iseed + rev, w->kb, &w->ring);
      weak_bunch_calc_statistics(&w->bunch, &track);
    }
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      dump_stats(fst, rev, w->kb, &w->bunch.stats);
      if (w->bunch.kb_out == 1) {
        Nstat[k]++;
        if (Nstat[k] % track.NrevMon == 0 || (rev + 1) % track.NrevScan == 0) {
          weak_bunch_update_statistics(&w->bunch.stats, Nstat[k]);
          weak_bunch_reset_statistics(&w->bunch.stats);
          Nstat[k] = 0;
        }
      }
    }
    fprintf(frf, "%ld %.17g %.17g\n", rev, W[0].ring.phai0, W[0].ring.Vrf0);
    if (track.EnableRW_long > 0)
      for (int k = 0; k < Nb; k++) {
        double *mom = cyclic_array_access(-rev * h + W[k].kb, &dipole_RW);
        mom[0] = W[k].bunch.stats.pos.x;
        mom[1] = W[k].bunch.stats.pos.z;
      }
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      if ((rev + 1) % track.NrevMon == 0) weak_bunch_calc_ampinv(&w->bunch, &track);
      if (track.EnableRW_long) {
        const int bpos = -rev * h + w->kb;
        if (track.TrackPlane[HOR])
          transform_weak_bunch_RW_longrange_cyclic(w->m, bpos, &dipole_RW, HOR, &mp, &ebeam, &w->bunch, &w->ring, &track);
        if (track.TrackPlane[VER])
          transform_weak_bunch_RW_longrange_cyclic(w->m, bpos, &dipole_RW, VER, &mp, &ebeam, &w->bunch, &w->ring, &track);
        w->m++;
      }
    }
    if ((rev + 1) % track.NrevMon == 0)
      for (int k = 0; k < Nb; k++) {
        const bunch_stats_t *s = &W[k].bunch.stats;
        fprintf(famp, "%ld %d %.17g %.17g %.17g %.17g %.17g %.17g\n", rev, W[k].kb,
                s->ampinv.v[0], s->ampinv.v[1], s->ampinv.v[2],
                s->ampinv_cm.v[0], s->ampinv_cm.v[1], s->ampinv_cm.v[2]);
      }
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int k = 0; k < Nb; k++) {
      worker_t *w = &W[k];
      if (nlrL > 0) fnp_ring_update(&w->ring, fnp_ring, &w->sf, &w->bunch, w->kb, LON);
      if (nlrH > 0 && track.TrackPlane[HOR]) fnp_ring_update(&w->ring, fnp_HOR, &w->sf, &w->bunch, w->kb, HOR);
      if (nlrV > 0 && track.TrackPlane[VER]) fnp_ring_update(&w->ring, fnp_VER, &w->sf, &w->bunch, w->kb, VER);
    }
    if (fh && Nbin > 0) {
      double *tmp = calloc((size_t)Nbin * h, sizeof(double));
      for (int k = 0; k < Nb; k++) {
        worker_t *w = &W[k];
        fnp_ring_update(&w->ring, tmp, &w->sf, &w->bunch, w->kb, LON);
        fprintf(fh, "%ld %d %d %d", rev, w->kb, w->bunch.N_trash_low, w->bunch.N_trash_high);
        for (int j = 0; j < Nbin; j++) fprintf(fh, " %.17g", tmp[w->kb * Nbin + j]);
        fprintf(fh, "\n");
      }
      free(tmp);
    }
    if ((rev + 1) % track.NrevScan == 0)
      for (int k = 0; k < Nb; k++) {
        tracking_scan_step_worker(rev, &W[k].scan_val, &track, &W[k].ring, &ebeam, &W[k].bunch);
        W[k].m = 0;
      }
  }
  clock_gettime(CLOCK_MONOTONIC, &t_b);
  fprintf(stderr, "turn_loop_seconds %.6f\n", (t_b.tv_sec - t_a.tv_sec) + 1e-9 * (t_b.tv_nsec - t_a.tv_nsec));
  if (serial_turns > 0 && serial_turns < track.NrevTot) {
    fprintf(stderr, "serial_turns %ld seconds %.6f\n", serial_turns,
            (t_s.tv_sec - t_a.tv_sec) + 1e-9 * (t_s.tv_nsec - t_a.tv_nsec));
    fprintf(stderr, "parallel_turns %ld seconds %.6f threads %d\n", track.NrevTot - serial_turns,
            (t_b.tv_sec - t_s.tv_sec) + 1e-9 * (t_b.tv_nsec - t_s.tv_nsec), qe_off ? omp_get_max_threads() : 1);
  }
  if (track.scan == 0)
    for (int k = 0; k < Nb; k++) {
      weak_bunch_calc_statistics(&W[k].bunch, &track);
      dump_stats(fst, track.NrevTot, W[k].kb, &W[k].bunch.stats);
    }
  for (int k = 0; k < Nb; k++) {
    if (W[k].kb != dump_parts_kb) continue;
    snprintf(path, sizeof path, "%s/parts_%d.txt", track.work_path, W[k].kb);
    FILE *fp = fopen(path, "w");
    for (unsigned j = 0; j < W[k].bunch.Np; j++) {
      particle_t *p = &W[k].bunch.particles[j];
      fprintf(fp, "%.17g %.17g %.17g %.17g %.17g %.17g\n", p->pos.xtau, p->slope.xtau,
              p->pos.x, p->slope.x, p->pos.z, p->slope.z);
    }
    fclose(fp);
  }
  for (int k = 0; k < Nb; k++)
    if (W[k].trafo_fp) fclose(W[k].trafo_fp);
  fclose(fst);
  fclose(famp);
  fclose(frf);
  if (fh) fclose(fh);
  return 0;
}
