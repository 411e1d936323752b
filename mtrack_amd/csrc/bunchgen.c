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
  if (c->EnableRW_short || c->EnableRW_short_LON) {
    mbt_set_error("short-range resistive-wall Green functions (EnableRW_short) are not supported");
    return MBT_EUNSUPPORTED;
  }
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
