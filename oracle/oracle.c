/*
 * CPU oracle: FP64 restatement of the reference's weak-weak multi-bunch turn.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -ffp-contract=off so
 * each multiply and add rounds separately, as the reference's x86-64 build does.
 *
 * The MPI program runs one rank per bunch; its per-turn exchanges are
 * allgathers of identical data, and the quantities every rank computes
 * redundantly (long-range resonator phasors, RF-feedback solve) are identical
 * on all ranks.  The oracle therefore holds one copy of the shared state and
 * one set of per-bunch state, and runs the turn in the phase order the MPI
 * program imposes (src/worker_weakweak.c:259-394):
 *   1. LR-resonator phasor recurrence (construct_wake_phasor, run once for all
 *      buckets instead of once per rank), RF-feedback solve;
 *   2. per bunch: selffield kick, one-turn map, statistics;
 *   3. RW dipole history write for all bunches;
 *   4. per bunch: ampinv on monitor turns, RW long-range kicks;
 *   5. per bunch: LR histogram update; scan step.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define LON 0
#define HOR 1
#define VER 2
#define FKILO 1.0e+03
#define FGIGA 1.0e+09
#define FMILLI 1.0e-03
#define FTERA 1.0e+12
#define FNANO 1.0e-09
#define C_LIGHT 2.997925e+08
#define Z_0 376.73
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------ glibc rand() */

/* glibc random(3) TYPE_3, in the closed form r[i] = r[i-3] + r[i-31] (mod 2^32)
 * over a generated sequence: r[0] = seed, r[1..30] by the minimal standard LCG
 * (16807, 2^31-1) computed with glibc's signed 32-bit Schrage step, r[31..33] =
 * r[0..2], output k = r[k+344] >> 1.  Kept as a 34-entry sliding window. */
typedef struct orc_rand {
  uint32_t w[34];
  long n; /* index of the next r[] to produce */
} orc_rand;

static void orc_srand(orc_rand *g, unsigned seed)
{
  int32_t r[34];
  if (seed == 0) seed = 1;
  r[0] = (int32_t)seed;
  int32_t word = r[0];
  for (int i = 1; i < 31; i++) {
    long hi = word / 127773, lo = word % 127773;
    long t = 16807 * lo - 2836 * hi;
    if (t < 0) t += 2147483647;
    word = (int32_t)t;
    r[i] = word;
  }
  for (int i = 31; i < 34; i++) r[i] = r[i - 31];
  for (int i = 0; i < 34; i++) g->w[i] = (uint32_t)r[i];
  g->n = 34;
  for (int k = 0; k < 310; k++) {
    uint32_t v = g->w[(g->n - 3) % 34] + g->w[(g->n - 31) % 34];
    g->w[g->n % 34] = v;
    g->n++;
  }
}

static int orc_rand_next(orc_rand *g)
{
  uint32_t v = g->w[(g->n - 3) % 34] + g->w[(g->n - 31) % 34];
  g->w[g->n % 34] = v;
  g->n++;
  return (int)(v >> 1);
}

void orc_glibc_rand_seq(unsigned seed, int n, int *out)
{
  orc_rand g;
  orc_srand(&g, seed);
  for (int i = 0; i < n; i++) out[i] = orc_rand_next(&g);
}

/* One MPI rank's random state: glibc stream, bunch.c's "last seed" (starts at 0,
 * src/bunch.h:17) and box_muller's static pair cache (src/bunch.c:409-411). */
typedef struct rank_rng {
  orc_rand g;
  int iseed0;
  int seeded;
  float y2;
  int use_last;
} rank_rng;

static void rank_seed(rank_rng *s, int iseed)
{
  if (iseed != s->iseed0) {
    orc_srand(&s->g, (unsigned)iseed);
    s->iseed0 = iseed;
    s->seeded = 1;
  }
  if (!s->seeded) {
    orc_srand(&s->g, 1u);
    s->seeded = 1;
  }
}

/* c_fnorm, src/bunch.c:379-405 */
static double orc_c_fnorm(rank_rng *s, int iseed)
{
  rank_seed(s, iseed);
  double sum = 0.0;
  for (int i = 0; i < 12; i++) {
    double u = ((double)orc_rand_next(&s->g)) / ((double)2147483647);
    double v = u - 0.5;
    sum = sum + v;
  }
  return sum;
}

/* one rank's c_fnorm stream for the other oracle units (oracle_ws.c) */
void *orc_rank_rng_new(void) { return calloc(1, sizeof(rank_rng)); }
double orc_rank_c_fnorm(void *s, int iseed) { return orc_c_fnorm((rank_rng *)s, iseed); }

/* box_muller, src/bunch.c:407-436: Marsaglia polar method in float */
static double orc_box_muller(rank_rng *s, int iseed)
{
  float x1, x2, w, y1;
  rank_seed(s, iseed);
  if (s->use_last) {
    y1 = s->y2;
    s->use_last = 0;
  } else {
    do {
      x1 = 2.0 * ((double)orc_rand_next(&s->g)) / ((double)2147483647) - 1.0;
      x2 = 2.0 * ((double)orc_rand_next(&s->g)) / ((double)2147483647) - 1.0;
      w = x1 * x1 + x2 * x2;
    } while (w >= 1.0);
    w = sqrt((-2.0 * log(w)) / w);
    y1 = x1 * w;
    s->y2 = x2 * w;
    s->use_last = 1;
  }
  return y1;
}

/* -------------------------------------------------- counter-based QE stream */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) */
void orc_philox4x32(const unsigned ctr[4], const unsigned key[2], unsigned out[4])
{
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* float natural log on (0, 1]: u = m 2^e with m in [sqrt(1/2), sqrt(2)),
 * ln m = 2 atanh(s), s = (m-1)/(m+1), odd series to s^9; explicit fmaf only. */
float orc_logf(float u)
{
  union {
    float f;
    uint32_t i;
  } b = {u};
  int e = (int)((b.i >> 23) & 0xff) - 127;
  b.i = (b.i & 0x007fffffu) | 0x3f800000u;
  float m = b.f;
  if (m > 1.41421356f) {
    m = m * 0.5f;
    e += 1;
  }
  float s = (m - 1.0f) / (m + 1.0f);
  float s2 = s * s;
  float p = fmaf(s2, 0.11111111f, 0.14285715f);
  p = fmaf(s2, p, 0.2f);
  p = fmaf(s2, p, 0.33333334f);
  p = fmaf(s2, p, 1.0f);
  float lm = 2.0f * s * p;
  return fmaf((float)e, 0.69314718f, lm);
}

/* sin and cos of 2 pi t for t in [0, 1]: quadrant from rint(4t), Taylor
 * polynomials on [-pi/4, pi/4]; explicit fmaf only. */
void orc_sincos2pif(float t, float *s, float *c)
{
  if (t > 0.5f) t = t - 1.0f;
  float q = rintf(4.0f * t);
  float r = fmaf(q, -0.25f, t);
  float x = r * 6.2831855f;
  float x2 = x * x;
  float sp = fmaf(x2, 2.7557319e-6f, -1.9841270e-4f);
  sp = fmaf(x2, sp, 8.3333333e-3f);
  sp = fmaf(x2, sp, -0.16666667f);
  float sn = fmaf(x * x2, sp, x);
  float cp = fmaf(x2, 2.4801587e-5f, -1.3888889e-3f);
  cp = fmaf(x2, cp, 4.1666667e-2f);
  cp = fmaf(x2, cp, -0.5f);
  float cs = fmaf(x2, cp, 1.0f);
  int iq = ((int)q) & 3;
  switch (iq) {
  case 0: *s = sn; *c = cs; break;
  case 1: *s = cs; *c = -sn; break;
  case 2: *s = -sn; *c = -cs; break;
  default: *s = -cs; *c = sn; break;
  }
}

static inline float u01(uint32_t x) { return ((float)((x >> 8) + 1u)) * 5.9604645e-8f; }

static inline void box_muller_pair(uint32_t a, uint32_t b, double *n0, double *n1)
{
  float r = sqrtf(-2.0f * orc_logf(u01(a)));
  float s, c;
  orc_sincos2pif(u01(b), &s, &c);
  *n0 = (double)(r * c);
  *n1 = (double)(r * s);
}

/* Normals of particle j of the bunch in `bucket` at turn `rev`: key (seed,
 * bucket), counters (j, rev lo, rev hi, 0|1); n[0] LON, n[1..2] VER, n[3..4] HOR. */
void orc_qe_normals(unsigned seed, unsigned bucket, unsigned long long rev, unsigned j, double n[5])
{
  unsigned key[2] = {seed, bucket};
  unsigned ctr[4] = {j, (unsigned)rev, (unsigned)(rev >> 32), 0u};
  unsigned o[4];
  orc_philox4x32(ctr, key, o);
  double t;
  box_muller_pair(o[0], o[1], &n[0], &n[1]);
  box_muller_pair(o[2], o[3], &n[2], &n[3]);
  ctr[3] = 1u;
  orc_philox4x32(ctr, key, o);
  box_muller_pair(o[0], o[1], &n[4], &t);
}

/* -------------------------------------------------- initial distribution */

/* weak_generate_bunch_distribution, src/bunch_weak.c:85-105,172-292 (normal
 * planes; file input and the vertical excitation mode are exercised through
 * the product's generator against the reference harness) */
static void gen_plane(const mbt_config *c, rank_rng *s, int kb, int plane, int Np, double *pos, double *slope)
{
  const int iseed = c->iseed[plane] + kb;
  for (int j = 0; j < Np; j++) {
    double tmp_pos = c->pos_sgm[plane] * orc_c_fnorm(s, iseed);
    pos[j] = c->pos_offset[plane] + tmp_pos;
    slope[j] = c->slope_offset[plane] + c->slope_sgm[plane] * orc_c_fnorm(s, iseed) + c->correlate[plane] * tmp_pos;
  }
}

static int gen_bunch(const mbt_config *c, int kb, double *out, rank_rng *s)
{
  int Np = c->Np;
  memset(out, 0, sizeof(double) * 6 * (size_t)Np);
  if (Np == 1) {
    out[0] = c->pos_offset[0];
    out[1] = c->slope_offset[0];
    if (c->TrackPlane[1]) {
      out[2] = c->pos_offset[1];
      out[3] = c->slope_offset[1];
    }
    if (c->TrackPlane[2]) {
      out[4] = c->pos_offset[2];
      out[5] = c->slope_offset[2];
    }
    return 0;
  }
  if (!c->nGen[0] || (c->TrackPlane[1] && !c->nGen[1]) || (c->TrackPlane[2] && (!c->nGen[2] || c->mode_excitation)))
    return -1;
  gen_plane(c, s, kb, 0, Np, out, out + Np);
  if (c->TrackPlane[1]) gen_plane(c, s, kb, 1, Np, out + 2 * (size_t)Np, out + 3 * (size_t)Np);
  if (c->TrackPlane[2]) gen_plane(c, s, kb, 2, Np, out + 4 * (size_t)Np, out + 5 * (size_t)Np);
  return 0;
}

int orc_generate_bunch(const mbt_config *c, int kb, double *out)
{
  rank_rng s;
  memset(&s, 0, sizeof s);
  return gen_bunch(c, kb, out, &s);
}

/* ------------------------------------------------------------------ ring */

typedef struct { double pos[3], slope[3]; } particle;

typedef struct obunch {
  int kb;
  unsigned Np;
  double Ib;
  particle *p;
  double pos[3], slope[3], pos_sigma[3], slope_sigma[3];
  double ampinv[3], ampinv_cm[3];
  int N_trash_low, N_trash_high;
  int m; /* RW trigger counter */
  rank_rng rng;
  int *mapcell;
  double *fnp, *dipV, *dipH;
} obunch;

struct orc_ring {
  mbt_config c; /* shallow copy: arrays still owned by the caller */
  int rng_mode;
  int Nb, h, Nc;
  obunch *b;
  double *Gl, *Gv, *Gh;
  int PlaneL, PlaneH, PlaneV;
  double phai0, Vrf0;
  int has_fb, fb_off, fb_count;
  double *fb_hist; /* [len_average][2] */
  double phasor_end[3][2 * MBT_MAX_LR];
  double *fnp_ring[3]; /* [h*Nc] */
  double *lr_wake[3];  /* [h*Nc] */
  double *hist;        /* RW dipole history [(Nmlt+2)*h][2] */
  long n_hist;
  long rev;
  double scan_val;
  double Iring;
  double *Ibunch;
};

void orc_ring_destroy(orc_ring *r)
{
  if (!r) return;
  for (int k = 0; k < r->Nb; k++) {
    free(r->b[k].p);
    free(r->b[k].mapcell);
    free(r->b[k].fnp);
    free(r->b[k].dipV);
    free(r->b[k].dipH);
  }
  free(r->b);
  free(r->Gl);
  free(r->Gv);
  free(r->Gh);
  free(r->fb_hist);
  for (int p = 0; p < 3; p++) {
    free(r->fnp_ring[p]);
    free(r->lr_wake[p]);
  }
  free(r->hist);
  free(r->Ibunch);
  free(r);
}

/* Green functions: construct_greensfunc_resonator src/transform_weak.c:289-403 */
static void greens(orc_ring *r)
{
  const mbt_config *c = &r->c;
  int Nc = r->Nc;
  for (int j = 0; j < c->n_res; j++) {
    const mbt_resonator *R = &c->res[j];
    if (!c->TrackPlane[R->plane]) continue;
    double Q = R->Qfactor, w = R->wr;
    double resol = w * c->sigma_tau * c->dT;
    if (R->plane == LON) {
      double sq = sqrt(1.0 - 1.0 / (4.0 * Q * Q));
      double ar = 0.5 * resol / Q, arq = sq * resol, fac = 0.5 / (Q * sq), ampl = w * R->Rs / Q;
      r->Gl[0] += 0.5 * ampl;
      for (int n = 1; n < Nc; n++) {
        double fn = (double)n;
        double arg = fn * arq;
        double argE = fn * ar;
        r->Gl[n] += ampl * exp(-argE) * (cos(arg) - fac * sin(arg));
      }
      r->PlaneL++;
    } else {
      double *G = R->plane == VER ? r->Gv : r->Gh;
      if (Q > 0.5) {
        double sq = sqrt(4.0 * Q * Q - 1.0), ar = 0.5 * resol / Q, arq = sq * resol / (2 * Q);
        double ampl = 2 * R->Rs * w / sq;
        for (int n = 0; n < Nc; n++) {
          double fn = (double)n;
          double arg = fn * arq;
          double argE = fn * ar;
          G[n] += ampl * exp(-argE) * sin(arg);
        }
      }
      if (Q < 0.5) {
        double sq = sqrt(1.0 - 4.0 * Q * Q), ar = 0.5 * resol / Q, arq = sq * resol / (2 * Q);
        double ampl = 2 * R->Rs * w / sq;
        for (int n = 0; n < Nc; n++) {
          double fn = (double)n;
          double arg = fn * arq;
          double argE = fn * ar;
          G[n] += ampl * exp(-argE) * sinh(arg);
        }
      }
      if (R->plane == VER) r->PlaneV++;
      else r->PlaneH++;
    }
  }
}

/* construct_greensfunc_file, src/transform_weak.c:406-471.  Each file's
 * samples (t [ns], w) are read whole first; the reader then starts at the first
 * sample with t >= 0 and walks the bin edges k dTau as the reference does: for
 * bin i it advances to the first sample beyond i dTau (remembering the last one
 * at or before it), skips the bins that sample also lies beyond, and fills the
 * bins since the last filled one by the linear interpolation between the two
 * samples.  After the last sample the reference's reader keeps the last pair
 * for both ends (0 / 0 in the interpolation); bins beyond Ncell that its walk
 * would write are not written here.  Longitudinal functions get their first
 * bin halved. */
static int greens_file(orc_ring *r)
{
  const mbt_config *c = &r->c;
  const int Nc = r->Nc;
  const double dTau = c->sigma_tau * c->dT;
  for (int j = 0; j < c->n_wakefile; j++) {
    FILE *fp = fopen(c->wakefile[j], "r");
    if (!fp) return -1;
    size_t n = 0, cap = 1024;
    double *ts = malloc(cap * sizeof(double)), *ws = malloc(cap * sizeof(double)), a, b;
    while (fscanf(fp, "%lf %lf", &a, &b) == 2) {
      if (n == cap) {
        cap *= 2;
        ts = realloc(ts, cap * sizeof(double));
        ws = realloc(ws, cap * sizeof(double));
      }
      ts[n] = a;
      ws[n] = b;
      n++;
    }
    fclose(fp);
    double *G = calloc(Nc, sizeof(double));
    size_t p = 0;
    double t = 0.0, w = 0.0;
    while (p < n) {
      t = ts[p];
      w = ws[p];
      p++;
      if (!(t < 0)) break;
    }
    double lt = 0.0, lw = 0.0;
    long last = -1;
    int eof = 0;
    for (long i = 0; i < Nc; i++) {
      if (eof) G[i] = G[i - 1];
      while (t <= i * dTau) {
        lt = t;
        lw = w;
        if (p >= n) {
          eof = 1;
          break;
        }
        t = ts[p];
        w = ws[p];
        p++;
      }
      while (t > (i + 1) * dTau) i++;
      for (long k = last + 1; k <= i && k < Nc; k++) G[k] = (w * (k * dTau - lt) + lw * (t - k * dTau)) / (t - lt);
      if (i > last + 1) i--;
      last = i;
    }
    double *dst = c->wakefile_plane[j] == LON ? r->Gl : c->wakefile_plane[j] == HOR ? r->Gh : r->Gv;
    if (c->wakefile_plane[j] == LON) G[0] = G[0] / 2.;
    for (int i = 0; i < Nc; i++) dst[i] += G[i];
    if (c->wakefile_plane[j] == LON) r->PlaneL++;
    else if (c->wakefile_plane[j] == HOR) r->PlaneH++;
    else r->PlaneV++;
    free(G);
    free(ts);
    free(ws);
  }
  return 0;
}

/* RW_table_LON / RW_table_TRANS, src/rw_table.c:311-395, on the tabulated
 * first-bin values (data shared with the product: the reference gives no
 * closed form for them) */
#include "../mtrack_amd/csrc/rw_first_bin_data.h"
static int rw_table(const double *w, double uval, double *wert)
{
  *wert = 0.0;
  if (uval > 2000.) *wert = w[RWFB_N - 1];
  int n = 0;
  while (*wert == 0. && n < RWFB_N - 1) {
    double u_diff = rwfb_u[n] - uval, u_next_diff = rwfb_u[n + 1] - uval;
    if (u_diff <= 0. && u_next_diff > 0.) {
      double du = rwfb_u[n + 1] - rwfb_u[n], dw = w[n + 1] - w[n];
      *wert = w[n] - dw / du * u_diff;
    }
    n++;
  }
  return *wert == 0.0 ? -1 : 1;
}

/* construct_greensfunc_RW, src/transform_weak.c:474-547 */
static int greens_rw(orc_ring *r)
{
  const mbt_config *c = &r->c;
  const int Nc = r->Nc;
  const double sgmatau = c->sigma_tau, dTau = c->dT, pi = M_PI, Z0 = 376.73, CL = 2.997925e+08;
  const double beffL1 = c->beffL[0], beffL2 = pow(c->beffL[1], 2);
  const double beffV3 = pow(c->beffV[0], 3), beffH3 = pow(c->beffH[0], 3);
  const double beffV4 = pow(c->beffV[1], 4), beffH4 = pow(c->beffH[1], 4);
  const double sigmarw = 1.0 / c->rhorw, t_const = sgmatau * dTau;
  double rw0;
  if (c->TrackPlane[LON] && c->EnableRW_short_LON) {
    double ARW = sqrt(Z0 / (sigmarw * CL * pi)) / (4 * beffL1 * pi) * c->Lc;
    for (int nc = 1; nc < Nc; nc++) {
      double tRW = (nc + 0.5) * t_const;
      r->Gl[nc] -= ARW / (tRW * sqrt(tRW));
    }
    double s0 = pow((2 * beffL2 * c->rhorw / Z0), 1.0 / 3.0);
    if (rw_table(rwfb_lon, sgmatau * dTau * CL / s0, &rw0) < 0) return -1;
    r->Gl[0] += rw0 * 4 * Z0 * CL / (pi * beffL2) * c->Lc;
    r->PlaneL++;
    if (!c->EnableRW_short) return 0;
  }
  if (c->TrackPlane[VER]) {
    double ARW = sqrt(Z0 * CL / sigmarw / pi) / (pi * beffV3) * c->Lc;
    double s0 = pow((2 * c->beffV[1] * c->beffV[1] * c->rhorw / Z0), 1.0 / 3.0);
    if (rw_table(rwfb_trans, sgmatau * dTau * CL / s0, &rw0) < 0) return -1;
    r->Gv[0] += rw0 * 8 * Z0 * CL / (pi * beffV4) * s0 * c->Lc;
    for (int nc = 1; nc < Nc; nc++) {
      double tRW = (nc + 0.5) * t_const;
      r->Gv[nc] += ARW / sqrt(tRW);
    }
    r->PlaneV++;
  }
  if (c->TrackPlane[HOR]) {
    double ARW = sqrt(Z0 * CL / sigmarw / pi) / (pi * beffH3) * c->Lc;
    double s0 = pow((2 * c->beffH[1] * c->beffH[1] * c->rhorw / Z0), 1.0 / 3.0);
    if (rw_table(rwfb_trans, sgmatau * dTau * CL / s0, &rw0) < 0) return -1;
    r->Gh[0] += rw0 * 8 * Z0 * CL / (pi * beffH4) * s0 * c->Lc;
    for (int nc = 1; nc < Nc; nc++) {
      double tRW = (nc + 0.5) * t_const;
      r->Gh[nc] += ARW / sqrt(tRW);
    }
    r->PlaneH++;
  }
  return 0;
}

int orc_ring_greens(const orc_ring *r, double *Gl, double *Gv, double *Gh, int *planes)
{
  for (int i = 0; i < r->Nc; i++) {
    Gl[i] = r->Gl[i];
    Gv[i] = r->Gv[i];
    Gh[i] = r->Gh[i];
  }
  planes[0] = r->PlaneL;
  planes[1] = r->PlaneH;
  planes[2] = r->PlaneV;
  return 0;
}

orc_ring *orc_ring_create(const mbt_config *c, int rng_mode)
{
  orc_ring *r = calloc(1, sizeof *r);
  r->c = *c;
  r->rng_mode = rng_mode;
  r->Nb = c->Nbunch;
  r->h = (int)c->Nharm;
  r->Nc = c->Ncell;
  r->phai0 = c->phai0;
  r->Vrf0 = c->Vrf0;
  r->Iring = c->Iring;
  r->scan_val = c->scan_start;
  r->Ibunch = malloc(r->h * sizeof(double));
  memcpy(r->Ibunch, c->Ibunch, r->h * sizeof(double));
  r->b = calloc(r->Nb, sizeof(obunch));
  int Nc = r->Nc > 0 ? r->Nc : 1;
  for (int k = 0; k < r->Nb; k++) {
    obunch *b = &r->b[k];
    b->kb = c->bucket[k];
    b->Np = (unsigned)c->Np;
    b->Ib = c->bunch_Ib[k];
    b->p = calloc(b->Np, sizeof(particle));
    b->mapcell = calloc(b->Np, sizeof(int));
    b->fnp = calloc(Nc, sizeof(double));
    b->dipV = calloc(Nc, sizeof(double));
    b->dipH = calloc(Nc, sizeof(double));
    double *tmp = malloc(sizeof(double) * 6 * b->Np);
    if (gen_bunch(c, b->kb, tmp, &b->rng) != 0) {
      free(tmp);
      orc_ring_destroy(r);
      return NULL;
    }
    orc_ring_set_bunch(r, k, tmp);
    free(tmp);
  }
  r->Gl = calloc(Nc, sizeof(double));
  r->Gv = calloc(Nc, sizeof(double));
  r->Gh = calloc(Nc, sizeof(double));
  if (r->Nc > 0) {
    greens(r);
    if (greens_file(r) < 0 || ((c->EnableRW_short || c->EnableRW_short_LON) && greens_rw(r) < 0)) {
      orc_ring_destroy(r);
      return NULL;
    }
  }
  for (int p = 0; p < 3; p++) {
    r->fnp_ring[p] = calloc((size_t)r->h * Nc, sizeof(double));
    r->lr_wake[p] = calloc((size_t)r->h * Nc, sizeof(double));
  }
  return r;
}

int orc_ring_set_bunch(orc_ring *r, int ib, const double *s)
{
  obunch *b = &r->b[ib];
  unsigned Np = b->Np;
  for (unsigned j = 0; j < Np; j++) {
    b->p[j].pos[0] = s[j];
    b->p[j].slope[0] = s[Np + j];
    b->p[j].pos[1] = s[2 * Np + j];
    b->p[j].slope[1] = s[3 * Np + j];
    b->p[j].pos[2] = s[4 * Np + j];
    b->p[j].slope[2] = s[5 * Np + j];
  }
  return 0;
}

int orc_ring_get_bunch(const orc_ring *r, int ib, double *s)
{
  const obunch *b = &r->b[ib];
  unsigned Np = b->Np;
  for (unsigned j = 0; j < Np; j++) {
    s[j] = b->p[j].pos[0];
    s[Np + j] = b->p[j].slope[0];
    s[2 * Np + j] = b->p[j].pos[1];
    s[3 * Np + j] = b->p[j].slope[1];
    s[4 * Np + j] = b->p[j].pos[2];
    s[5 * Np + j] = b->p[j].slope[2];
  }
  return 0;
}

/* weak_bunch_mean_rms + weak_bunch_calc_statistics, src/statistics.c:91-157 */
static void calc_stats(orc_ring *r, obunch *b)
{
  for (int pl = 0; pl < 3; pl++) {
    if (!r->c.TrackPlane[pl]) continue;
    double ps = 0.0, ss = 0.0;
    for (unsigned j = 0; j < b->Np; j++) {
      ps += b->p[j].pos[pl];
      ss += b->p[j].slope[pl];
    }
    double pa = ps / ((double)b->Np), sa = ss / ((double)b->Np);
    b->pos[pl] = pa;
    b->slope[pl] = sa;
    ps = 0.0;
    ss = 0.0;
    for (unsigned j = 0; j < b->Np; j++) {
      ps += pow(pa - b->p[j].pos[pl], 2);
      ss += pow(sa - b->p[j].slope[pl], 2);
    }
    b->pos_sigma[pl] = sqrt(ps / ((double)b->Np));
    b->slope_sigma[pl] = sqrt(ss / ((double)b->Np));
  }
}

/* weak_bunch_calc_ampinv, src/statistics.c:163-195 (optics of the global ring:
 * the LON plane's are zero) */
static void calc_ampinv(orc_ring *r, obunch *b)
{
  const mbt_config *c = &r->c;
  for (int pl = 0; pl < 3; pl++) {
    if (!c->TrackPlane[pl]) continue;
    double sum = 0.0;
    for (unsigned j = 0; j < b->Np; j++) {
      double x = b->p[j].pos[pl], xp = b->p[j].slope[pl];
      sum += c->gamma1[pl] * pow(x, 2) + 2.0 * c->alpha1[pl] * x * xp + c->beta1[pl] * pow(xp, 2);
    }
    b->ampinv[pl] = sum / ((double)b->Np);
    b->ampinv_cm[pl] = c->gamma1[pl] * pow(b->pos[pl], 2) + 2.0 * c->alpha1[pl] * b->pos[pl] * b->slope[pl] +
                       c->beta1[pl] * pow(b->slope[pl], 2);
  }
}

static void put_stats(const obunch *b, double *o)
{
  for (int p = 0; p < 3; p++) {
    o[p] = b->pos[p];
    o[3 + p] = b->pos_sigma[p];
    o[6 + p] = b->slope[p];
    o[9 + p] = b->slope_sigma[p];
  }
}

/* rffb_get_vrf_phi, src/feedback_rf.c:5-12 (atan2 argument order as there) */
static void fb_sample(orc_ring *r, double vreal, double vimag)
{
  double phi = atan2(vreal, vimag);
  double vrf = sqrt(vreal * vreal + vimag * vimag);
  int n = r->c.rffb_len_average;
  unsigned key = ((unsigned)r->fb_count % (unsigned)n) * 2;
  r->fb_hist[key] = vrf;
  r->fb_hist[key + 1] = phi;
  r->fb_count++;
}

/* rffb_calc_mean_voltage_phase src/feedback_rf.c:21-34 + the solve of
 * src/worker_weakweak.c:268-278 */
static void fb_solve(orc_ring *r)
{
  double vs = 0, ps = 0;
  int n = r->c.rffb_len_average;
  for (int m = 0; m < n; m++) {
    vs += r->fb_hist[2 * m];
    ps += r->fb_hist[2 * m + 1];
  }
  double vmbar = vs / n, phibar = ps / n;
  vmbar = vmbar * r->c.E0 * FKILO;
  double num = r->c.rffb_vrf_design * sin(r->c.rffb_phi0_design) - vmbar * sin(phibar);
  r->phai0 = atan2(num, r->c.rffb_vrf_design * cos(r->c.rffb_phi0_design) - vmbar * cos(phibar));
  r->Vrf0 = num / sin(r->phai0);
}

/* The recurrence of construct_wake_phasor src/transform_weak.c:961-1048 (and
 * of wake_phasor_init :869-952 when init_turns > 0), over all buckets once;
 * lr_wake of every filled bucket is collected in the same pass. */
static void phasor(orc_ring *r, int plane, int init_turns)
{
  const mbt_config *c = &r->c;
  int Nbin = r->Nc, h = r->h;
  double *fnp = r->fnp_ring[plane];
  double *lw = r->lr_wake[plane];
  if (!init_turns) memset(lw, 0, sizeof(double) * (size_t)h * Nbin);
  for (int l = 0; l < c->n_lr[plane]; l++) {
    const mbt_lr_resonator *res = &c->lr[plane][l];
    double alpha = res->wr * 0.5 / res->Qfactor;
    double Amp = 2 * alpha * res->Rs * c->T0 / c->E0 / FGIGA;
    double dTau = c->dT * c->sigma_tau;
    double tbucket = (c->T0 / c->h - Nbin * dTau);
    double fac = Amp / r->b[0].Np * FMILLI;
    double C0 = -alpha, C1 = res->wr;
    double pr0 = exp(C0 * dTau) * cos(C1 * dTau), pr1 = exp(C0 * dTau) * sin(C1 * dTau);
    double Vo0, Vo1, Vn0, Vn1;
    if (init_turns) {
      Vo0 = 0.0;
      Vo1 = 0.0;
    } else {
      Vo0 = r->phasor_end[plane][2 * l];
      Vo1 = r->phasor_end[plane][2 * l + 1];
    }
    Vn0 = r->phasor_end[plane][2 * l];
    Vn1 = r->phasor_end[plane][2 * l + 1];
    int fb = r->has_fb && plane == LON && c->rffb_resonator == l + 1;
    double taubeam = 0, taub2beam = 0;
    if (fb) {
      taubeam = c->Nsigma * c->sigma_tau + dTau * 1.5;
      taub2beam = -c->Nsigma * c->sigma_tau + dTau * 0.5 - tbucket;
    }
    double pb0 = exp(C0 * taubeam) * cos(C1 * taubeam), pb1 = exp(C0 * taubeam) * sin(C1 * taubeam);
    double pa0 = exp(C0 * taub2beam) * cos(C1 * taub2beam), pa1 = exp(C0 * taub2beam) * sin(C1 * taub2beam);
    int nturn = init_turns ? init_turns : 1;
    for (int k = 0; k < nturn; k++)
      for (int m = 0; m < h; m++) {
        int i = h - m - 1;
        if (fb) fb_sample(r, Vn0 * pb0 - Vn1 * pb1, Vn0 * pb1 + Vn1 * pb0);
        if (c->nfFill[i] == 1) {
          double Ii = r->Ibunch[i];
          for (int j = 0; j < Nbin; j++) {
            Vn0 = (Vo0 * pr0 - Vo1 * pr1) - Ii * fnp[i * Nbin + j] * fac;
            Vn1 = (Vo0 * pr1 + Vo1 * pr0);
            if (!init_turns) lw[i * Nbin + j] += Vn0 + 0.5 * Ii * fnp[i * Nbin + j] * fac;
            Vo0 = Vn0;
            Vo1 = Vn1;
          }
          Vn0 = (Vo0 * exp(C0 * tbucket) * cos(C1 * tbucket) - Vo1 * exp(C0 * tbucket) * sin(C1 * tbucket));
          Vn1 = (Vo0 * exp(C0 * tbucket) * sin(C1 * tbucket) + Vo1 * exp(C0 * tbucket) * cos(C1 * tbucket));
          Vo0 = Vn0;
          Vo1 = Vn1;
        } else {
          double tb = c->T0 / c->h;
          Vn0 = (Vo0 * exp(C0 * tb) * cos(C1 * tb) - Vo1 * exp(C0 * tb) * sin(C1 * tb));
          Vn1 = (Vo0 * exp(C0 * tb) * sin(C1 * tb) + Vo1 * exp(C0 * tb) * cos(C1 * tb));
          Vo0 = Vn0;
          Vo1 = Vn1;
        }
        if (fb) fb_sample(r, Vn0 * pa0 - Vn1 * pa1, Vn0 * pa1 + Vn1 * pa0);
      }
    r->phasor_end[plane][2 * l] = Vn0;
    r->phasor_end[plane][2 * l + 1] = Vn1;
  }
}

/* fnp_ring_update, src/transform_weak.c:839-865 (out-of-range particles dropped) */
static void fnp_update(orc_ring *r, obunch *b, int plane)
{
  int Nbin = r->Nc;
  double *f = r->fnp_ring[plane] + (size_t)b->kb * Nbin;
  for (int i = 0; i < Nbin; i++) f[i] = 0.0;
  for (unsigned j = 0; j < b->Np; j++) {
    double Tau = -b->p[j].pos[0] / r->c.sigma_tau;
    int icell = (int)((Tau + r->c.Nsigma) / r->c.dT + 1);
    if (icell >= 0 && icell < Nbin) {
      if (plane == LON) f[icell] += 1.0;
      else if (plane == HOR) f[icell] += b->p[j].pos[1];
      else f[icell] += b->p[j].pos[2];
    }
  }
}

/* transform_weak_bunch_selffield src/transform_weak.c:553-731.  Reads of
 * fnp[] outside [0, Ncell) by the resistive/inductive smoothing (:673,:680-681)
 * and the GL1[Ncell] iteration of :658 are replaced by zero / omitted. */
static void selffield(orc_ring *r, obunch *b)
{
  const mbt_config *c = &r->c;
  int Nc = r->Nc;
  double sgm = c->sigma_tau;
  double sgm2 = sgm * sgm, dTau2 = c->dT * c->dT;
  double factG = -c->T0 * b->Ib / (b->Np * c->E0 * FGIGA);
  double *fnp = b->fnp, *dV = b->dipV, *dH = b->dipH;
  double GL1[Nc], GV[Nc], GH[Nc];
  memset(fnp, 0, Nc * sizeof(double));
  memset(dV, 0, Nc * sizeof(double));
  memset(dH, 0, Nc * sizeof(double));
  memset(GL1, 0, sizeof GL1);
  memset(GV, 0, sizeof GV);
  memset(GH, 0, sizeof GH);
  for (unsigned j = 0; j < b->Np; j++) {
    double Tau = -b->p[j].pos[0] / sgm;
    int icell = (int)((Tau + c->Nsigma) / c->dT + 1);
    if (icell < 0) {
      icell = 0;
      b->N_trash_low++;
    } else if (icell >= Nc) {
      icell = Nc - 1;
      b->N_trash_high++;
    }
    b->mapcell[j] = icell;
    fnp[icell] += 1.0;
    dV[icell] += b->p[j].pos[2];
    dH[icell] += b->p[j].pos[1];
  }
  int icmin = Nc - 1, icmax = 0;
  for (int i = 0; i < icmin; i++)
    if (fnp[i] > 0) {
      icmin = i;
      break;
    }
  for (int i = Nc - 1; i > icmax; i--)
    if (fnp[i] > 0) {
      icmax = i;
      break;
    }
#define FNP(i) (((i) >= 0 && (i) < Nc) ? fnp[i] : 0.0)
  if (r->PlaneL > 0) {
    for (int i = icmin; i < Nc; i++) {
      double Gn = 0.0;
      for (int k = i; k >= icmin; k--) Gn += fnp[k] * r->Gl[i - k];
      GL1[i] = 1 * Gn;
    }
    if (c->RsisL > 0)
      for (int i = icmin + 1; i <= icmax; i++) {
        double nt = (0.5 * FNP(i - 1) + FNP(i) + 0.5 * FNP(i + 1)) / 2.0;
        GL1[i] += c->RsisL * nt / (sgm * c->dT);
      }
    if (c->aindL > 0)
      for (int i = icmin + 1; i <= icmax; i++) {
        double n0 = (0.25 * FNP(i - 3) + 0.5 * FNP(i - 2) + FNP(i - 1) + 0.5 * FNP(i) + 0.25 * FNP(i + 1)) / 2.5;
        double n1 = (0.25 * FNP(i - 1) + 0.5 * FNP(i) + FNP(i + 1) + 0.5 * FNP(i + 2) + 0.25 * FNP(i + 3)) / 2.5;
        GL1[i] += (n1 - n0) / (2.0 * sgm2 * dTau2) * c->aindL;
      }
  }
#undef FNP
  const double *lrL = r->lr_wake[LON] + (size_t)b->kb * Nc;
  const double *lrH = r->lr_wake[HOR] + (size_t)b->kb * Nc;
  const double *lrV = r->lr_wake[VER] + (size_t)b->kb * Nc;
  if (c->n_lr[LON] + r->PlaneL > 0)
    for (unsigned j = 0; j < b->Np; j++) b->p[j].slope[0] += factG * GL1[b->mapcell[j]] + lrL[b->mapcell[j]];
  if (r->PlaneV > 0) {
    for (int i = 0; i < Nc; i++) {
      double Gn = 0.0;
      for (int k = i; k >= 0; k--) Gn += dV[k] * r->Gv[i - k];
      GV[i] = 1 * Gn;
    }
    for (unsigned j = 0; j < b->Np; j++) b->p[j].slope[2] += -factG * GV[b->mapcell[j]] + lrV[b->mapcell[j]];
  }
  if (r->PlaneH > 0) {
    for (int i = 1; i < Nc; i++) {
      double Gn = 0.0;
      for (int k = i; k >= 0; k--) Gn += dH[k] * r->Gh[i - k];
      GH[i] = 1 * Gn;
    }
    for (unsigned j = 0; j < b->Np; j++) b->p[j].slope[1] += -factG * GH[b->mapcell[j]] + lrH[b->mapcell[j]];
  }
}

/* transform_weak_bunch_optic, src/transform_weak.c:21-192 */
static void optic(orc_ring *r, obunch *b, long rev)
{
  const mbt_config *c = &r->c;
  const unsigned Np = b->Np;
  const int kb = b->kb;
  const int iseed = c->iseed[LON] + kb + (int)rev;
  const int qe = c->EnableQuantum;
  const double damp = 1.0 - c->De;
  const double excite = 2.0 * c->slope_sgm[LON] * sqrt(c->T0 / c->taue);
  const double qVp = c->pos_sgm[VER] * sqrt(2.0 * c->T0 / c->tauz);
  const double qVs = c->slope_sgm[VER] * sqrt(2.0 * c->T0 / c->tauz);
  const double qHp = c->pos_sgm[HOR] * sqrt(2.0 * c->T0 / c->taux);
  const double qHs = c->slope_sgm[HOR] * sqrt(2.0 * c->T0 / c->taux);
  const double eps_const = r->Vrf0 / (c->E0 * FKILO);
  particle *p0 = malloc(Np * sizeof(particle));
  double *rff = malloc(Np * sizeof(double));
  double *nrm = NULL;
  if (qe && r->rng_mode == ORC_RNG_PHILOX) {
    nrm = malloc(sizeof(double) * 5 * Np);
    for (unsigned j = 0; j < Np; j++)
      orc_qe_normals((unsigned)c->iseed[LON], (unsigned)kb, (unsigned long long)rev, j, nrm + 5 * (size_t)j);
  }
#define NORMAL(j, slot) (r->rng_mode == ORC_RNG_PHILOX ? nrm[5 * (size_t)(j) + (slot)] : orc_box_muller(&b->rng, iseed))
  for (unsigned j = 0; j < Np; j++) {
    p0[j] = b->p[j];
    particle *p = &b->p[j];
    double g;
    if (c->EnableIdealHC == 0) {
      g = eps_const * sin(c->wrf * p0[j].pos[0] + r->phai0);
      for (int a = 0; a < c->n_ahc; a++) {
        const mbt_active_hc *A = &c->ahc[a];
        g += A->Vpeak / (c->E0 * FGIGA) * sin(c->wrf * A->nHC * p0[j].pos[0] + A->phi_aHC + A->nHC * kb * 2 * M_PI);
      }
    } else {
      g = eps_const * (sin(c->wrf * p0[j].pos[0] + r->phai0) +
                       c->HC_k * sin(c->m_aHC * c->wrf * p0[j].pos[0] + c->m_aHC * c->phi_n));
    }
    p->slope[0] = p0[j].slope[0] + g - c->Urad;
    if (qe) {
      p->slope[0] = p->slope[0] * damp + excite * NORMAL(j, 0);
      rff[j] = (1.0 + p0[j].slope[0]) / (1.0 + p0[j].slope[0] + g);
    }
    p->pos[0] = p0[j].pos[0] - p->slope[0] * c->T0 * c->ac;
  }
  if (c->TrackPlane[VER]) {
    double PsiV0 = 2.0 * M_PI * c->QV0;
    for (unsigned j = 0; j < Np; j++) {
      particle *p = &b->p[j];
      double dQ = 0.0;
      if (c->AmpDependdQV) {
        double EV = c->gamma1[VER] * p0[j].pos[2] * p0[j].pos[2] + 2 * c->alpha1[VER] * p0[j].pos[2] * p0[j].slope[2] +
                    c->beta1[VER] * p0[j].slope[2] * p0[j].slope[2];
        double EH = c->gamma1[VER] * p0[j].pos[1] * p0[j].pos[1] + 2 * c->alpha1[VER] * p0[j].pos[1] * p0[j].slope[1] +
                    c->beta1[VER] * p0[j].slope[1] * p0[j].slope[1];
        dQ = c->CVH * EH + c->CVV * EV;
      }
      double Psi = PsiV0 * (1.0 + c->Gziz * p0[j].slope[0] + dQ);
      double cs = cos(Psi), sn = sin(Psi);
      double a11 = cs + c->alpha1[VER] * sn, a12 = c->beta1[VER] * sn;
      double a21 = -c->gamma1[VER] * sn, a22 = cs - c->alpha1[VER] * sn;
      p->pos[2] = a11 * p0[j].pos[2] + a12 * p0[j].slope[2];
      p->slope[2] = a21 * p0[j].pos[2] + a22 * p0[j].slope[2];
      if (qe) {
        p->pos[2] = p->pos[2] + qVp * NORMAL(j, 1);
        p->slope[2] = rff[j] * p->slope[2] + qVs * NORMAL(j, 2);
      }
    }
  }
  if (c->TrackPlane[HOR]) {
    double PsiH0 = 2.0 * M_PI * c->QH0;
    for (unsigned j = 0; j < Np; j++) {
      particle *p = &b->p[j];
      double dQ = 0;
      if (c->AmpDependdQV) { /* the HOR tune shift is gated by the VER flag and uses VER optics */
        double EV = c->gamma1[VER] * p0[j].pos[2] * p0[j].pos[2] + 2 * c->alpha1[VER] * p0[j].pos[2] * p0[j].slope[2] +
                    c->beta1[VER] * p0[j].slope[2] * p0[j].slope[2];
        double EH = c->gamma1[VER] * p0[j].pos[1] * p0[j].pos[1] + 2 * c->alpha1[VER] * p0[j].pos[1] * p0[j].slope[1] +
                    c->beta1[VER] * p0[j].slope[1] * p0[j].slope[1];
        dQ = c->CHH * EH + c->CHV * EV;
      }
      double Psi = PsiH0 * (1.0 + c->Gzix * p0[j].slope[0] + dQ);
      double cs = cos(Psi), sn = sin(Psi);
      double a11 = cs + c->alpha1[HOR] * sn;
      double a12 = c->beta1[HOR] * sn;
      double a13 = (1.0 - a11) * c->dispH1 - a12 * c->disppH1;
      double a21 = -c->gamma1[HOR] * sn;
      double a22 = cs - c->alpha1[HOR] * sn;
      double a23 = -a21 * c->dispH1 + (1.0 - a22) * c->disppH1;
      p->pos[1] = a11 * p0[j].pos[1] + a12 * p0[j].slope[1] + a13 * p->slope[0];
      p->slope[1] = a21 * p0[j].pos[1] + a22 * p0[j].slope[1] + a23 * p->slope[0];
      if (qe) {
        p->pos[1] = p->pos[1] + qHp * NORMAL(j, 3);
        p->slope[1] = rff[j] * p->slope[1] + qHs * NORMAL(j, 4);
      }
    }
  }
#undef NORMAL
  free(p0);
  free(rff);
  free(nrm);
}

static inline double *hist_at(orc_ring *r, long k)
{
  long n = r->n_hist;
  long m = k < 0 ? (k % n + n) % n : k % n;
  return r->hist + 2 * m;
}

/* transform_weak_bunch_RW_longrange_cyclic, src/transform_weak.c:780-833 */
static void rw_kick(orc_ring *r, obunch *b, int plane, long bpos)
{
  const mbt_config *c = &r->c;
  int h = (int)c->Nharm;
  double beff3 = plane == HOR ? c->beffH[0] : c->beffV[0];
  if (c->triggerRW && b->m < 50) beff3 /= 2;
  double wv = 0.;
  double deltab = c->T0 / c->h;
  double RWconst = c->T0 / c->E0 / c->fNp / FTERA * c->Lc / (M_PI * pow(beff3, 3)) * sqrt(Z_0 * C_LIGHT * c->rhorw / M_PI);
  for (long m = 1; m < ((long)(c->Nmlt + 1) * h + 1); m++) {
    long k = bpos + m;
    long in_mod = k < 0 ? (k % h + h) % h : k % h;
    if (c->nfFill[in_mod]) {
      const double *mom = hist_at(r, k);
      double tt = m * deltab;
      wv += c->fNp * mom[plane - 1] * r->Ibunch[in_mod] / sqrt(tt);
    }
  }
  wv *= RWconst;
  for (unsigned j = 0; j < b->Np; j++) b->p[j].slope[plane] += wv;
}

int orc_ring_prepare(orc_ring *r, double *init_stats)
{
  const mbt_config *c = &r->c;
  for (int k = 0; k < r->Nb; k++) {
    calc_stats(r, &r->b[k]);
    if (init_stats) put_stats(&r->b[k], init_stats + MBT_NSTAT * (size_t)k);
  }
  if (c->n_lr[LON] > 0) {
    if (c->has_rf_feedback) {
      r->has_fb = 1;
      r->c.rffb_len_average = abs(c->rffb_len_average);
      r->fb_off = c->rffb_len_average < 0;
      r->fb_hist = calloc(2 * (size_t)r->c.rffb_len_average, sizeof(double));
      r->fb_count = 0;
    }
    for (int k = 0; k < r->Nb; k++) fnp_update(r, &r->b[k], LON);
    phasor(r, LON, (int)(c->Nbumax / c->h));
    if (r->has_fb) {
      fb_solve(r);
      if (r->fb_off) r->has_fb = 0;
    }
  }
  if (c->TrackPlane[HOR] && c->n_lr[HOR] > 0) {
    for (int k = 0; k < r->Nb; k++) fnp_update(r, &r->b[k], HOR);
    phasor(r, HOR, (int)(c->Nbumax / c->h));
  }
  if (c->TrackPlane[VER] && c->n_lr[VER] > 0) {
    for (int k = 0; k < r->Nb; k++) fnp_update(r, &r->b[k], VER);
    phasor(r, VER, (int)(c->Nbumax / c->h));
  }
  if (c->EnableRW_long > 0) {
    int h = r->h;
    r->n_hist = (long)(c->Nmlt + 2) * h;
    r->hist = calloc(2 * (size_t)r->n_hist, sizeof(double));
    for (int k = 0; k < r->Nb; k++) {
      double *m = hist_at(r, r->b[k].kb);
      m[0] = r->b[k].pos[HOR];
      m[1] = r->b[k].pos[VER];
    }
    for (int k = 0; k < r->Nb; k++) {
      double *m = hist_at(r, r->b[k].kb);
      for (int j = 1; j < c->Nmlt; j++) {
        double *d = hist_at(r, (long)j * h + r->b[k].kb);
        d[0] = m[0];
        d[1] = m[1];
      }
    }
  }
  r->rev = 0;
  return 0;
}

int orc_ring_track(orc_ring *r, long nturns, double *stats, double *ampinv, double *rf)
{
  const mbt_config *c = &r->c;
  int lr_tot = c->n_lr[0] + c->n_lr[1] + c->n_lr[2];
  for (long t = 0; t < nturns; t++) {
    long rev = r->rev;
    if (r->PlaneL + r->PlaneV + r->PlaneH + lr_tot > 0) {
      if (c->n_lr[LON] > 0) phasor(r, LON, 0);
      if (c->n_lr[HOR] > 0) phasor(r, HOR, 0);
      if (c->n_lr[VER] > 0) phasor(r, VER, 0);
    }
    if (r->has_fb) fb_solve(r);
    if (rf) {
      rf[2 * t] = r->phai0;
      rf[2 * t + 1] = r->Vrf0;
    }
    for (int k = 0; k < r->Nb; k++) {
      obunch *b = &r->b[k];
      if (r->PlaneL + r->PlaneV + r->PlaneH + lr_tot > 0) selffield(r, b);
      optic(r, b, rev);
      calc_stats(r, b);
      if (stats) put_stats(b, stats + ((size_t)t * r->Nb + k) * MBT_NSTAT);
    }
    if (c->EnableRW_long > 0)
      for (int k = 0; k < r->Nb; k++) {
        double *m = hist_at(r, -rev * r->h + r->b[k].kb);
        m[0] = r->b[k].pos[HOR];
        m[1] = r->b[k].pos[VER];
      }
    for (int k = 0; k < r->Nb; k++) {
      obunch *b = &r->b[k];
      if ((rev + 1) % c->NrevMon == 0) {
        calc_ampinv(r, b);
        if (ampinv) {
          double *o = ampinv + ((size_t)t * r->Nb + k) * MBT_NAMP;
          for (int p = 0; p < 3; p++) {
            o[p] = b->ampinv[p];
            o[3 + p] = b->ampinv_cm[p];
          }
        }
      }
      if (c->EnableRW_long) {
        long bpos = -rev * r->h + b->kb;
        if (c->TrackPlane[HOR]) rw_kick(r, b, HOR, bpos);
        if (c->TrackPlane[VER]) rw_kick(r, b, VER, bpos);
        b->m++;
      }
    }
    for (int k = 0; k < r->Nb; k++) {
      if (c->n_lr[LON] > 0) fnp_update(r, &r->b[k], LON);
      if (c->n_lr[HOR] > 0 && c->TrackPlane[HOR]) fnp_update(r, &r->b[k], HOR);
      if (c->n_lr[VER] > 0 && c->TrackPlane[VER]) fnp_update(r, &r->b[k], VER);
    }
    if ((rev + 1) % c->NrevScan == 0) {
      /* tracking_scan_step_worker, src/tracking.c:43-59 */
      if (c->scan == 1) {
        r->Iring += c->scan_step;
        r->scan_val = r->Iring;
        for (int k = 0; k < r->Nb; k++) r->b[k].Ib += c->scan_step / (((double)r->Nb) * FKILO);
        for (int i = 0; i < r->h; i++)
          if (c->nfFill[i]) r->Ibunch[i] += c->scan_step / ((double)r->Nb);
      }
      for (int k = 0; k < r->Nb; k++) r->b[k].m = 0;
    }
    r->rev++;
  }
  return 0;
}

int orc_ring_final_stats(orc_ring *r, double *stats)
{
  for (int k = 0; k < r->Nb; k++) {
    calc_stats(r, &r->b[k]);
    put_stats(&r->b[k], stats + MBT_NSTAT * (size_t)k);
  }
  return 0;
}

/* Clamped histogram of the current coordinates (the binning rule of
 * src/transform_weak.c:608-631) and the cumulative trash counters. */
int orc_ring_histograms(const orc_ring *r, double *counts, int *trash)
{
  int Nc = r->Nc;
  for (int k = 0; k < r->Nb; k++) {
    const obunch *b = &r->b[k];
    double *f = counts + (size_t)k * Nc;
    memset(f, 0, Nc * sizeof(double));
    for (unsigned j = 0; j < b->Np; j++) {
      double Tau = -b->p[j].pos[0] / r->c.sigma_tau;
      int icell = (int)((Tau + r->c.Nsigma) / r->c.dT + 1);
      if (icell < 0) icell = 0;
      else if (icell >= Nc) icell = Nc - 1;
      f[icell] += 1.0;
    }
    trash[2 * k] = b->N_trash_low;
    trash[2 * k + 1] = b->N_trash_high;
  }
  return 0;
}
