/*
 * Set-up time host code: glibc rand() stream, initial bunch distributions and
 * the selffield Green functions.
 *
 * Initial distributions are generated bit-identically to the reference
 * (weak_generate_bunch_distribution src/bunch_weak.c:172-292, c_fnorm
 * src/bunch.c:379-405): each bunch owns an independent generator (the reference
 * runs one MPI rank per bunch, each with its own rand() state and its own
 * "last seed" memory initialised to 0), so bunches are generated in parallel.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mbt_host.h"

/* glibc random(3), TYPE_3 (x^31 + x^3 + 1 additive feedback): the seed fills a
 * 31-word table through the Park-Miller minimal standard generator (Schrage
 * decomposition with a signed 32-bit word, as glibc's srandom_r computes it);
 * a front index starting 3 words ahead of the rear index walks the table,
 * each step adds the rear word into the front word (mod 2^32) and returns it
 * shifted right by one; the first 310 outputs are discarded. */
void mbt_glibc_srand(mbt_glibc_rand *g, unsigned seed)
{
  if (seed == 0) seed = 1;
  int word = (int)seed;
  g->r[0] = word;
  for (int i = 1; i < 31; i++) {
    long hi = word / 127773;
    long lo = word % 127773;
    long w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    word = (int)w;
    g->r[i] = word;
  }
  g->f = 3;
  g->b = 0;
  for (int k = 0; k < 310; k++) (void)mbt_glibc_rand_next(g);
}

int mbt_glibc_rand_next(mbt_glibc_rand *g)
{
  unsigned v = (unsigned)g->r[g->f] + (unsigned)g->r[g->b];
  g->r[g->f] = (int)v;
  g->f = g->f == 30 ? 0 : g->f + 1;
  g->b = g->b == 30 ? 0 : g->b + 1;
  return (int)(v >> 1);
}

/* Per-bunch state mirroring one reference rank: generator + last seed (the
 * header-static iseed0 of src/bunch.h:17 starts at 0 in every process). */
typedef struct gen_state {
  mbt_glibc_rand g;
  int iseed0;
  int seeded;
} gen_state;

static void gen_seed(gen_state *s, int iseed)
{
  if (iseed != s->iseed0) {
    mbt_glibc_srand(&s->g, (unsigned)iseed);
    s->iseed0 = iseed;
    s->seeded = 1;
  }
  if (!s->seeded) { /* glibc's unseeded state equals srand(1) */
    mbt_glibc_srand(&s->g, 1u);
    s->seeded = 1;
  }
}

static double gen_fnorm(gen_state *s, int iseed)
{
  gen_seed(s, iseed);
  double sum = 0.0;
  for (int i = 0; i < 12; i++) {
    double u = ((double)mbt_glibc_rand_next(&s->g)) / ((double)2147483647);
    double v = u - 0.5;
    sum = sum + v;
  }
  return sum;
}

/* src/bunch_weak.c:85-105 */
static void normal_plane(const mbt_config *c, gen_state *s, int kb, int plane, int Np, double *pos, double *slope)
{
  const double po = c->pos_offset[plane], ps = c->pos_sgm[plane];
  const double so = c->slope_offset[plane], ss = c->slope_sgm[plane], cr = c->correlate[plane];
  const int iseed = c->iseed[plane] + kb;
  for (int j = 0; j < Np; j++) {
    double tp = ps * gen_fnorm(s, iseed);
    pos[j] = po + tp;
    slope[j] = so + ss * gen_fnorm(s, iseed) + cr * tp;
  }
}

/* src/bunch_weak.c:108-170 (vertical head-tail mode excitation) */
static void excitation_plane(const mbt_config *c, gen_state *s, int Np, const double *tau, const double *delta,
                             double *z, double *zp)
{
  double psum = 0.0, ssum = 0.0;
  for (int j = 0; j < Np; j++) {
    psum += tau[j];
    ssum += delta[j];
  }
  double tau_off = psum / ((double)Np);
  double del_off = ssum / ((double)Np);
  gen_seed(s, c->iseed[MBT_VER]);
  (void)mbt_glibc_rand_next(&s->g); /* phi_beta = 0*rand()... */
  double phi_beta = 0.0;
  double fmodeHT = (double)c->modeHT;
  double norm_constant = 0.0;
  if (c->EnableIdealHC == 1)
    norm_constant = 1.85407 * c->ac * c->T0 / (MBT_PI * MBT_PI) / c->Nharm / c->wso /
                    sqrt((c->m_aHC * c->m_aHC - 1) / 6.0);
  for (int j = 0; j < Np; j++) {
    double sx = delta[j] - del_off;
    double px = tau[j] - tau_off;
    double px2 = px * px;
    double nf;
    if (c->EnableIdealHC == 1)
      nf = norm_constant / sqrt((px2 + sqrt(px2 * px2 + 4 * sx * sx * norm_constant * norm_constant)) / 2.0);
    else nf = c->fc1;
    double psi = atan(nf * sx / (tau[j] - tau_off));
    if (psi < 0.0 && sx > 0.0) psi = psi + MBT_PI;
    if (psi > 0.0 && sx < 0.0) psi = psi + MBT_PI;
    double argA = 0 * c->wgziV * tau[j];
    double argB = fmodeHT * psi + phi_beta;
    double cA = cos(argA), sA = sin(argA), cB = cos(argB), sB = sin(argB);
    z[j] = (c->zzCM_offset + sqrt(2) * c->sgm_zz * (cA * cB + sA * sB)) / MBT_FKILO;
    zp[j] = (c->zpCM_offset - sqrt(2) * c->sgm_zz / c->beta1[MBT_VER] * (sA * cB - cA * sB) +
             sqrt(2) * c->sgm_zz * c->alpha1[MBT_VER] / c->beta1[MBT_VER] * (cA * cB + sA * sB)) /
            MBT_FKILO;
  }
}

static int read_plane_file(const char *path, int Np, double spos, double sslope, double *pos, double *slope)
{
  FILE *fp = fopen(path, "r");
  if (!fp) {
    fprintf(stderr, "Error ! Cannot open %s\n", path);
    mbt_set_error("cannot open distribution file %s", path);
    return MBT_EIO;
  }
  for (int j = 0; j < Np; j++) {
    int idx;
    double a, b;
    if (fscanf(fp, "%d %lf %lf\n", &idx, &a, &b) != 3) {
      fclose(fp);
      mbt_set_error("short distribution file %s", path);
      return MBT_EIO;
    }
    pos[j] = a * spos;
    slope[j] = b * sslope;
  }
  fclose(fp);
  return MBT_OK;
}

int mbt_generate_bunch(const mbt_config *c, int kb, double *out)
{
  const int Np = c->Np;
  memset(out, 0, sizeof(double) * 6 * (size_t)(Np > 0 ? Np : 0));
  if (Np <= 0) return MBT_OK;
  double *tau = out, *del = out + Np, *x = out + 2 * (size_t)Np, *xp = out + 3 * (size_t)Np;
  double *z = out + 4 * (size_t)Np, *zp = out + 5 * (size_t)Np;
  if (Np == 1) {
    tau[0] = c->pos_offset[0];
    del[0] = c->slope_offset[0];
    if (c->TrackPlane[1]) {
      x[0] = c->pos_offset[1];
      xp[0] = c->slope_offset[1];
    }
    if (c->TrackPlane[2]) {
      z[0] = c->pos_offset[2];
      zp[0] = c->slope_offset[2];
    }
    return MBT_OK;
  }
  gen_state s;
  memset(&s, 0, sizeof s);
  int rc;
  if (c->nGen[0]) normal_plane(c, &s, kb, 0, Np, tau, del);
  else if ((rc = read_plane_file(c->gendst_datafile[0], Np, MBT_FNANO, 1.0, tau, del)) != MBT_OK) return rc;
  if (c->TrackPlane[1]) {
    if (c->nGen[1]) normal_plane(c, &s, kb, 1, Np, x, xp);
    else if ((rc = read_plane_file(c->gendst_datafile[1], Np, MBT_FMILLI, MBT_FMILLI, x, xp)) != MBT_OK) return rc;
  }
  if (c->TrackPlane[2]) {
    if (c->nGen[2]) {
      if (c->mode_excitation) excitation_plane(c, &s, Np, tau, del, z, zp);
      else normal_plane(c, &s, kb, 2, Np, z, zp);
    } else if ((rc = read_plane_file(c->gendst_datafile[2], Np, MBT_FMILLI, MBT_FMILLI, z, zp)) != MBT_OK)
      return rc;
  }
  return MBT_OK;
}

/* ------------------------------------------------------------ Green functions */

/* src/transform_weak.c:289-403 (broadband resonators) and :406-471 (wake files) */
#include "rw_first_bin_data.h"

/* First-bin value of the short-range resistive wall at u (RW_table_LON /
 * RW_table_TRANS, src/rw_table.c:311-395): linear between the tabulated points
 * u_k <= u < u_k+1, the last value beyond the table, 0 (an error) below it or at
 * its last point. */
static double rw_first_bin(const double *w, double u)
{
  if (u > 2000.) return w[RWFB_N - 1];
  for (int k = 0; k < RWFB_N - 1; k++) {
    const double d = rwfb_u[k] - u, dn = rwfb_u[k + 1] - u;
    if (d <= 0. && dn > 0.) {
      const double du = rwfb_u[k + 1] - rwfb_u[k], dw = w[k + 1] - w[k];
      return w[k] - dw / du * d;
    }
  }
  return 0.0;
}

/* construct_greensfunc_RW, src/transform_weak.c:474-547: the short-range
 * resistive-wall wake added to the Green functions (1/t^(3/2) longitudinal and
 * 1/sqrt(t) transverse tails from bin 1 on, tabulated first bins); the
 * reference's operation order. */
static int greens_rw(const mbt_config *c, double *Gl, double *Gv, double *Gh, int planes[3])
{
  const int Nc = c->Ncell;
  const double sgmatau = c->sigma_tau, dTau = c->dT, pi = M_PI;
  const double beffL1 = c->beffL[0], beffL2 = pow(c->beffL[1], 2);
  const double beffV3 = pow(c->beffV[0], 3), beffH3 = pow(c->beffH[0], 3);
  const double beffV4 = pow(c->beffV[1], 4), beffH4 = pow(c->beffH[1], 4);
  const double sigmarw = 1.0 / c->rhorw, t_const = sgmatau * dTau;
  if (c->TrackPlane[MBT_LON] && c->EnableRW_short_LON) {
    const double ARW = sqrt(MBT_Z0 / (sigmarw * MBT_C_LIGHT * pi)) / (4 * beffL1 * pi) * c->Lc;
    for (int nc = 1; nc < Nc; nc++) {
      const double tRW = (nc + 0.5) * t_const;
      Gl[nc] -= ARW / (tRW * sqrt(tRW));
    }
    const double s0 = pow((2 * beffL2 * c->rhorw / MBT_Z0), 1.0 / 3.0);
    const double v = rw_first_bin(rwfb_lon, sgmatau * dTau * MBT_C_LIGHT / s0);
    if (v == 0.0) {
      mbt_set_error("short-range RW: bin width outside the first-bin table (longitudinal)");
      return MBT_EINVAL;
    }
    Gl[0] += v * 4 * MBT_Z0 * MBT_C_LIGHT / (pi * beffL2) * c->Lc;
    planes[MBT_LON]++;
    if (!c->EnableRW_short) return MBT_OK;
  }
  for (int pl = MBT_HOR; pl <= MBT_VER; pl++) {
    const int plane = pl == MBT_HOR ? MBT_VER : MBT_HOR; /* the reference does VER first */
    if (!c->TrackPlane[plane]) continue;
    const double b3 = plane == MBT_VER ? beffV3 : beffH3, b4 = plane == MBT_VER ? beffV4 : beffH4;
    const double b1 = plane == MBT_VER ? c->beffV[1] : c->beffH[1];
    double *G = plane == MBT_VER ? Gv : Gh;
    const double ARW = sqrt(MBT_Z0 * MBT_C_LIGHT / sigmarw / pi) / (pi * b3) * c->Lc;
    const double s0 = pow((2 * b1 * b1 * c->rhorw / MBT_Z0), 1.0 / 3.0);
    const double v = rw_first_bin(rwfb_trans, sgmatau * dTau * MBT_C_LIGHT / s0);
    if (v == 0.0) {
      mbt_set_error("short-range RW: bin width outside the first-bin table (transverse)");
      return MBT_EINVAL;
    }
    G[0] += v * 8 * MBT_Z0 * MBT_C_LIGHT / (pi * b4) * s0 * c->Lc;
    for (int nc = 1; nc < Nc; nc++) {
      const double tRW = (nc + 0.5) * t_const;
      G[nc] += ARW / sqrt(tRW);
    }
    planes[plane]++;
  }
  return MBT_OK;
}

int mbt_greens_functions(const mbt_config *c, double *Gl, double *Gv, double *Gh, int planes[3])
{
  const int Nc = c->Ncell;
  planes[0] = planes[1] = planes[2] = 0;
  if (Nc <= 0) return MBT_OK;
  memset(Gl, 0, Nc * sizeof(double));
  memset(Gv, 0, Nc * sizeof(double));
  memset(Gh, 0, Nc * sizeof(double));
  const double sgm = c->sigma_tau;
  for (int j = 0; j < c->n_res; j++) {
    const mbt_resonator *r = &c->res[j];
    if (!c->TrackPlane[r->plane]) continue;
    double R = r->Rs, Q = r->Qfactor, w = r->wr;
    double resol = w * sgm * c->dT;
    if (r->plane == MBT_LON) {
      double sq = sqrt(1.0 - 1.0 / (4.0 * Q * Q));
      double ar = 0.5 * resol / Q;
      double arq = sq * resol;
      double fac = 0.5 / (Q * sq);
      double ampl = w * R / Q;
      Gl[0] += 0.5 * ampl;
      for (int n = 1; n < Nc; n++) {
        double fn = (double)n;
        double arg = fn * arq, argE = fn * ar;
        Gl[n] += ampl * exp(-argE) * (cos(arg) - fac * sin(arg));
      }
      planes[MBT_LON]++;
    } else {
      double *G = r->plane == MBT_VER ? Gv : Gh;
      if (Q > 0.5) {
        double sq = sqrt(4.0 * Q * Q - 1.0);
        double ar = 0.5 * resol / Q;
        double arq = sq * resol / (2 * Q);
        double ampl = 2 * R * w / sq;
        for (int n = 0; n < Nc; n++) {
          double fn = (double)n;
          G[n] += ampl * exp(-(fn * ar)) * sin(fn * arq);
        }
      }
      if (Q < 0.5) {
        double sq = sqrt(1.0 - 4.0 * Q * Q);
        double ar = 0.5 * resol / Q;
        double arq = sq * resol / (2 * Q);
        double ampl = 2 * R * w / sq;
        for (int n = 0; n < Nc; n++) {
          double fn = (double)n;
          G[n] += ampl * exp(-(fn * ar)) * sinh(fn * arq);
        }
      }
      planes[r->plane]++;
    }
  }
  const double dTau = sgm * c->dT;
  for (int j = 0; j < c->n_wakefile; j++) {
    FILE *wp = fopen(c->wakefile[j], "r");
    if (!wp) {
      fprintf(stderr, "ERROR: Cannot open file %s for reading.\n", c->wakefile[j]);
      mbt_set_error("cannot open wake file %s", c->wakefile[j]);
      return MBT_EIO;
    }
    double *Gt = calloc(Nc, sizeof(double));
    double t = 0, w = 0;
    do {
      if (fscanf(wp, "%lf%lf", &t, &w) != 2) break;
    } while (t < 0);
    double last_t = 0, last_w = 0;
    int last_i = -1, eof = 0;
    for (int i = 0; i < Nc; i++) {
      if (eof) Gt[i] = Gt[i - 1];
      while (t <= i * dTau) {
        last_t = 1 * t;
        last_w = 1 * w;
        if (fscanf(wp, "%lf %lf\n", &t, &w) != 2) {
          eof = 1;
          break;
        }
      }
      while (t > (i + 1) * dTau) i++;
      for (int k = last_i + 1; k < i + 1; k++)
        if (k < Nc) Gt[k] = (w * (k * dTau - last_t) + last_w * (t - k * dTau)) / (t - last_t);
      if (i > last_i + 1) i--;
      last_i = 1 * i;
    }
    fclose(wp);
    int pl = c->wakefile_plane[j];
    if (pl == MBT_LON) {
      Gt[0] = Gt[0] / 2.;
      for (int i = 0; i < Nc; i++) Gl[i] += Gt[i];
    } else {
      double *G = pl == MBT_HOR ? Gh : Gv;
      for (int i = 0; i < Nc; i++) G[i] += Gt[i];
    }
    planes[pl]++;
    free(Gt);
  }
  if (c->EnableRW_short || c->EnableRW_short_LON) return greens_rw(c, Gl, Gv, Gh, planes);
  return MBT_OK;
}

/* ---------------------------------------- all bunches, one thread per bunch */

#include <pthread.h>
#include <unistd.h>

typedef struct gen_job {
  const mbt_config *c;
  double *out;
  int next;
  int rc;
  pthread_mutex_t mu;
} gen_job;

static void *gen_worker(void *arg)
{
  gen_job *j = (gen_job *)arg;
  const size_t per = 6 * (size_t)j->c->Np;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    int b = j->next++;
    int stop = j->rc != MBT_OK;
    pthread_mutex_unlock(&j->mu);
    if (stop || b >= j->c->Nbunch) break;
    int rc = mbt_generate_bunch(j->c, j->c->bucket[b], j->out + per * b);
    if (rc) {
      pthread_mutex_lock(&j->mu);
      j->rc = rc;
      pthread_mutex_unlock(&j->mu);
    }
  }
  return NULL;
}

/* Initial coordinates of every filled bunch into out[Nbunch][6][Np]; bunches are
 * independent streams (one reference rank each), so they are generated by up to
 * 16 host threads. */
int mbt_generate_all(const mbt_config *c, double *out)
{
  long ncpu = sysconf(_SC_NPROCESSORS_ONLN);
  int nt = ncpu > 16 ? 16 : (ncpu < 1 ? 1 : (int)ncpu);
  if (nt > c->Nbunch) nt = c->Nbunch;
  gen_job j = {c, out, 0, MBT_OK, PTHREAD_MUTEX_INITIALIZER};
  pthread_t th[16];
  int started = 0;
  for (int i = 0; i < nt; i++)
    if (pthread_create(&th[i], NULL, gen_worker, &j) == 0) started++;
  if (started == 0) gen_worker(&j);
  for (int i = 0; i < started; i++) pthread_join(th[i], NULL);
  return j.rc;
}

/* ------------------------------------------------- weak-strong (STRONG) set-up */

/* FBII grid of fbii_grid_setup_parameters, src/bunch.c:814-850: Ngrid 5 points
 * per axis (offset 2), local spacing sigma/3 [m], absolute spacing twice that. */
void mbt_ws_grid(const mbt_config *c, double *agLOCx, double *agLOCz)
{
  *agLOCz = c->sgm_zz / ((double)3) / MBT_FKILO;
  *agLOCx = c->sgm_xx / ((double)3) / MBT_FKILO;
}

/* Slice-centroid bunches of the STRONG model, as the reference's manager rank
 * makes them: strong_generate_eCMs (src/bunch_strong.c:60-216) draws Np
 * macro-electrons per plane from one c_fnorm stream (seeds iseed[plane]),
 * sorts them into 2 Nslc + 1 longitudinal cells of 0.3 sigma from -7 sigma
 * (cell index clamped to [0, 2 Nslc + 1]), averages x, x', z, z' per cell and
 * counts them on each cell's 5 x 5 local grid; then
 * generate_initial_general_distribution2 (src/bunch_strong.c:10-50) offsets
 * every filled bucket's centroids by four more draws of the same stream
 * (seed 571030).  The cells' tau and delta stay zero: the reference never
 * assigns them.
 *   slices[kb][6][S]  xtau xeps xx xp zz zp of bucket kb (filled buckets only)
 *   npop[S], nofe[S][5][5], dist[6][Np] (may be NULL)                          */
int mbt_ws_generate(const mbt_config *c, double *slices, int *npop, int *nofe, double *dist)
{
  const int Np = c->Np, S = 2 * c->Nslc + 1, H = (int)c->Nharm;
  if (Np <= 0 || c->Nslc < 0 || S > 50) {
    mbt_set_error("weak-strong slices: need Np > 0 and 2 Nslc + 1 <= 50 (Nslc %d)", c->Nslc);
    return MBT_EINVAL;
  }
  double *m = malloc(sizeof(double) * 6 * (size_t)Np);
  if (!m) return MBT_ENOMEM;
  double *tau = m, *del = m + Np, *x = m + 2 * (size_t)Np, *xp = m + 3 * (size_t)Np;
  double *z = m + 4 * (size_t)Np, *zp = m + 5 * (size_t)Np;
  gen_state s;
  memset(&s, 0, sizeof s);
  int rc = MBT_OK;
  if (Np == 1) {
    tau[0] = c->xtauCM_offset * MBT_FNANO;
    del[0] = c->xepsCM_offset;
    x[0] = c->xxCM_offset * MBT_FMILLI;
    xp[0] = c->xpCM_offset * MBT_FMILLI;
    z[0] = c->zzCM_offset * MBT_FMILLI;
    zp[0] = c->zpCM_offset * MBT_FMILLI;
  } else {
    if (c->nGen[0])
      for (int j = 0; j < Np; j++) {
        tau[j] = (c->xtauCM_offset + c->sgm_xtau * gen_fnorm(&s, c->iseed[0])) * MBT_FNANO;
        del[j] = c->xepsCM_offset + c->sgm_xeps * gen_fnorm(&s, c->iseed[0]);
      }
    else rc = read_plane_file(c->gendst_datafile[0], Np, MBT_FNANO, 1.0, tau, del);
    if (rc == MBT_OK) {
      if (c->nGen[1])
        for (int j = 0; j < Np; j++) {
          double a = c->sgm_xx * gen_fnorm(&s, c->iseed[1]), b = c->sgm_xp * gen_fnorm(&s, c->iseed[1]);
          x[j] = a * MBT_FMILLI;
          xp[j] = b * MBT_FMILLI;
        }
      else rc = read_plane_file(c->gendst_datafile[1], Np, MBT_FMILLI, MBT_FMILLI, x, xp);
    }
    if (rc == MBT_OK) {
      if (c->nGen[2])
        for (int j = 0; j < Np; j++) {
          double a = c->sgm_zz * gen_fnorm(&s, c->iseed[2]), b = c->sgm_zp * gen_fnorm(&s, c->iseed[2]);
          z[j] = a * MBT_FMILLI;
          zp[j] = b * MBT_FMILLI;
        }
      else rc = read_plane_file(c->gendst_datafile[2], Np, MBT_FMILLI, MBT_FMILLI, z, zp);
    }
  }
  if (rc != MBT_OK) {
    free(m);
    return rc;
  }
  if (dist) memcpy(dist, m, sizeof(double) * 6 * (size_t)Np);
  double cx[51] = {0}, cxp[51] = {0}, cz[51] = {0}, czp[51] = {0};
  int pop[52] = {0};
  unsigned char *cell = malloc((size_t)Np);
  for (int j = 0; j < Np; j++) {
    double T = -tau[j] / c->sgmatau;
    int ic = (int)((7.0 + T) / 0.3);
    if (ic < 1) ic = 0;
    if (ic > S) ic = S;
    cell[j] = (unsigned char)ic;
    pop[ic]++;
    cx[ic] = cx[ic] + x[j];
    cxp[ic] = cxp[ic] + xp[j];
    cz[ic] = cz[ic] + z[j];
    czp[ic] = czp[ic] + zp[j];
  }
  for (int ic = 0; ic < S; ic++)
    if (pop[ic]) {
      cx[ic] = cx[ic] / pop[ic];
      cxp[ic] = cxp[ic] / pop[ic];
      cz[ic] = cz[ic] / pop[ic];
      czp[ic] = czp[ic] / pop[ic];
    }
  double aLx, aLz;
  mbt_ws_grid(c, &aLx, &aLz);
  memset(nofe, 0, sizeof(int) * 25 * (size_t)S);
  for (int j = 0; j < Np; j++) {
    const int ic = cell[j];
    if (ic >= S) continue;
    int nz = (int)(2.0 * (z[j] - cz[ic]) / aLz), mz = nz >= 0 ? (nz + 1) / 2 : (nz - 1) / 2;
    int nx = (int)(2.0 * (x[j] - cx[ic]) / aLx), mx = nx >= 0 ? (nx + 1) / 2 : (nx - 1) / 2;
    mz = mz > 2 ? 2 : (mz < -2 ? -2 : mz);
    mx = mx > 2 ? 2 : (mx < -2 ? -2 : mx);
    nofe[(ic * 5 + mx + 2) * 5 + mz + 2]++;
  }
  for (int ic = 0; ic < S; ic++) npop[ic] = pop[ic];
  memset(slices, 0, sizeof(double) * 6 * (size_t)S * H);
  for (int kb = 0; kb < H; kb++) {
    if (!c->nfFill[kb]) continue;
    double ox = c->xxCM_offset * MBT_FMILLI * gen_fnorm(&s, 571030);
    double op = c->xpCM_offset * MBT_FMILLI * gen_fnorm(&s, 571030);
    double oz = c->zzCM_offset * MBT_FMILLI * gen_fnorm(&s, 571030);
    double oq = c->zpCM_offset * MBT_FMILLI * gen_fnorm(&s, 571030);
    double *b = slices + (size_t)kb * 6 * S;
    for (int ic = 0; ic < S; ic++) {
      b[2 * S + ic] = cx[ic] + ox;
      b[3 * S + ic] = cxp[ic] + op;
      b[4 * S + ic] = cz[ic] + oz;
      b[5 * S + ic] = czp[ic] + oq;
    }
  }
  free(cell);
  free(m);
  return MBT_OK;
}
