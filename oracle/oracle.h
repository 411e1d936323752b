/*
 * CPU oracle for the multi-bunch tracking turn -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's weak-weak turn, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  The
 * product (libmbtrack_mi355x) never links, loads or calls it.
 *
 * Every function cites the reference code it follows.  How each is pinned is
 * listed in tests/test_oracle_*.py: bit-identical agreement with the compiled
 * reference (oracle/_ref, built from /root/reference where present) through the
 * committed golden fixtures, glibc's own rand(), the reference's Faddeeva
 * known answers, and closed forms.
 */
#ifndef MBT_ORACLE_H
#define MBT_ORACLE_H
#include "mbtrack_mi355x.h"

#define ORC_RNG_REFERENCE 0 /* glibc rand + float polar Box-Muller, per-rank state */
#define ORC_RNG_PHILOX 1    /* counter-based stream of the GPU build */

typedef struct orc_ring orc_ring;

orc_ring *orc_ring_create(const mbt_config *c, int rng_mode);
void orc_ring_destroy(orc_ring *r);
int orc_ring_set_bunch(orc_ring *r, int ib, const double *soa6);
int orc_ring_get_bunch(const orc_ring *r, int ib, double *soa6);
int orc_ring_prepare(orc_ring *r, double *init_stats);
int orc_ring_track(orc_ring *r, long nturns, double *stats, double *ampinv, double *rf);
int orc_ring_final_stats(orc_ring *r, double *stats);
int orc_ring_histograms(const orc_ring *r, double *counts, int *trash);
/* the Green functions Gl, Gv, Gh [Ncell] and the plane counters PlaneL/V/H */
int orc_ring_greens(const orc_ring *r, double *Gl, double *Gv, double *Gh, int *planes);
int orc_generate_bunch(const mbt_config *c, int kb, double *out);

/* glibc rand() stream after srand(seed) */
void orc_glibc_rand_seq(unsigned seed, int n, int *out);
/* Philox4x32-10 block */
void orc_philox4x32(const unsigned ctr[4], const unsigned key[2], unsigned out[4]);
/* the QE normals n[0..4] of one particle-turn (counter-based stream) */
void orc_qe_normals(unsigned seed, unsigned bucket, unsigned long long rev, unsigned j, double n[5]);
float orc_logf(float u);
void orc_sincos2pif(float t, float *s, float *c);
/* Faddeeva w(z), ACM TOMS 680 restated (src/physic.c:75-301); returns 0 or 1 (overflow) */
int orc_wofz(double x, double y, double *u, double *v);

/* ---- weak-strong turn with fast beam-ion interaction (oracle_ws.c) */
#define ORC_WS_IBUNCH_REFERENCE 0 /* ring.Ibunch left zero, as the reference's STRONG mode does */
#define ORC_WS_IBUNCH_FILL 1      /* bunch currents of the filling (the weak-weak manager's rule) */
typedef struct orc_ws orc_ws;
/* dist (may be NULL) receives the generated macro-electrons, 6 x Np (tau, delta, x, x', z, z') */
orc_ws *orc_ws_create(const mbt_config *c, int ibunch_mode, double *dist);
void orc_ws_destroy(orc_ws *w);
int orc_ws_track(orc_ws *w, long nturns, double *cm, double *mon, double *ion, double *zcm);
int orc_ws_get_slices(const orc_ws *w, int ib, double *out);
int orc_ws_population(const orc_ws *w, int *npop, int *nofe);

#endif
