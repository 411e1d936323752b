/*
 * Weak-strong multi-bunch turn with fast beam-ion interaction (FBII) --
 * TEST INFRASTRUCTURE ONLY (the oracle; the product never links it).
 *
 * A plain-C FP64 restatement of the reference's BunchModel = STRONG path in
 * one process, in the MPI phase order of manager_weakstrong /
 * worker_weakstrong (src/manager_weakstrong.c:7-263,
 * src/worker_weakstrong.c:14-247):
 *
 *   set-up   strong_generate_eCMs (src/bunch_strong.c:60-216), the FBII grid
 *            (src/bunch.c:814-850), generate_initial_general_distribution2
 *            (src/bunch_strong.c:10-50), fbii_model_init
 *            (src/fbi_weakstrong.c:18-47)
 *   turn     phase 1, every filled bunch: transformSbunch_shortrange
 *            (src/rw_weakstrong.c:10-67), fbi_weakstrong_bunch
 *            (src/fbi_weakstrong.c:49-298), statistics on monitor turns
 *            (src/bunch.c:516-550); manager: CM history (src/bunch.c:94-115),
 *            sortCMdata / resortCMdata (src/bunch.c:117-224) with the 5000-entry
 *            message of src/manager_weakstrong.c:170-173;
 *            phase 2: transformSbunch_optic (src/rw_weakstrong.c:194-253),
 *            transformSbunch_longrange (:69-192), bunch_strong_updateCM
 *            (src/bunch_strong.c:221-266);
 *            phase 3, bunch by bunch in train order: fbi_weakstrong_train
 *            (src/fbi_weakstrong.c:301-451) with
 *            get_linearly_interpolated_Efields (:454-580).
 *
 * Bunch currents: the reference never fills ring.Ibunch in this mode (it stays
 * zero, so ions carry no charge and the RW kick vanishes); ORC_WS_IBUNCH_FILL
 * uses the filling's bunch currents instead (the weak-weak manager's rule,
 * src/manager_weakweak.c:58-93), which is what makes the physics non-trivial.
 * Faddeeva: orc_wofz (TOMS 680, oracle_wofz.c).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_sens.h"

#define WS_NC 50        /* slice arrays of bunch_strong_t, src/bunch.h:113-121 */
#define WS_OFF 2        /* grid offset (Ngrid 5), src/bunch.c:820-827 */
#define WS_NG 5
#define WS_HIST 30      /* CM history depth, src/bunch.h:155 */
#define WS_HMAX 500     /* buckets of the CM history */
#define WS_MSG 5000     /* doubles of the history message, src/manager_weakstrong.c:170 */
#define WS_DATA 15000
#define ORC_ION_BLOCK 4096 /* ions per block of the parallel train */

/* physical constants of src/def.h:21-27 */
#define K_C 2.997925e+08
#define K_EPS0 8.854187817620e-12
#define K_E 1.602e-19
#define K_ME 9.1e-31
#define K_RE 2.819e-15
#define K_MPOME 1.836e3

void *orc_rank_rng_new(void);
double orc_rank_c_fnorm(void *s, int iseed);

typedef struct {
  double xx[WS_NC], xp[WS_NC], zz[WS_NC], zp[WS_NC], xtau[WS_NC], xeps[WS_NC];
  double pos[3], slope[3];
} ws_bunch;

struct orc_ws {
  const mbt_config *c;
  int S, Nb, h;
  int npop[WS_NC + 1];
  int nofe[WS_NC][WS_NG][WS_NG];
  double agLOCx, agLOCz, agABSx, agABSz, deltT, factBE;
  double Ldmsgm, factF;
  double *Ibunch; /* [h] mA, per bucket */
  ws_bunch *b;    /* [h], filled buckets used */
  int *fill;
  double (*xxh)[WS_HIST], (*xph)[WS_HIST], (*zzh)[WS_HIST], (*zph)[WS_HIST]; /* manager's history [h] */
  /* ion cloud at the interaction point */
  long nF, capF;
  int kbF;
  double *NriF, *xxF, *xdF, *yyF, *ydF;
  double *train_buf; /* [WS_NC][ORC_ION_BLOCK][2] Nri * E products of one block of ions */
  /* one bunch's fields and local ions */
  double Ex[WS_NC * WS_NG * WS_NG], Ey[WS_NC * WS_NG * WS_NG];
  double Nri[WS_NC * WS_NG * WS_NG], xxi[WS_NC * WS_NG * WS_NG], xdi[WS_NC * WS_NG * WS_NG];
  double yyi[WS_NC * WS_NG * WS_NG], ydi[WS_NC * WS_NG * WS_NG];
  int Nions;
  /* per filled bunch, the fields and local ions of this turn (phase 1 -> phase 3) */
  double *Exb, *Eyb, *Nrib, *xxib, *xdib, *yyib, *ydib;
  long rev;
  int wofz_fail;
};

/* ---------------------------------------------------------------- set-up */

/* strong_generate_eCMs, src/bunch_strong.c:60-216 (nGen planes; one rank's stream) */
static int gen_ecms(orc_ws *w, void *rng, ws_bunch *b0, double *dist /* 6*Np or NULL */)
{
  const mbt_config *c = w->c;
  const int Np = c->Np, S = w->S;
  double *t0 = malloc(sizeof(double) * 6 * (size_t)Np);
  double *e0 = t0 + Np, *x0 = t0 + 2 * (size_t)Np, *xp0 = t0 + 3 * (size_t)Np;
  double *z0 = t0 + 4 * (size_t)Np, *zp0 = t0 + 5 * (size_t)Np;
  if (Np == 1) {
    t0[0] = c->xtauCM_offset * 1.0e-09;
    e0[0] = c->xepsCM_offset;
    x0[0] = c->xxCM_offset * 1.0e-03;
    xp0[0] = c->xpCM_offset * 1.0e-03;
    z0[0] = c->zzCM_offset * 1.0e-03;
    zp0[0] = c->zpCM_offset * 1.0e-03;
  } else {
    if (!c->nGen[0] || !c->nGen[1] || !c->nGen[2]) {
      free(t0);
      return -1;
    }
    for (int j = 0; j < Np; j++) {
      t0[j] = (c->xtauCM_offset + c->sgm_xtau * orc_rank_c_fnorm(rng, c->iseed[0])) * 1.0e-09;
      e0[j] = c->xepsCM_offset + c->sgm_xeps * orc_rank_c_fnorm(rng, c->iseed[0]);
    }
    for (int j = 0; j < Np; j++) {
      double a = c->sgm_xx * orc_rank_c_fnorm(rng, c->iseed[1]);
      double p = c->sgm_xp * orc_rank_c_fnorm(rng, c->iseed[1]);
      x0[j] = a * 1.0e-03;
      xp0[j] = p * 1.0e-03;
    }
    for (int j = 0; j < Np; j++) {
      double a = c->sgm_zz * orc_rank_c_fnorm(rng, c->iseed[2]);
      double p = c->sgm_zp * orc_rank_c_fnorm(rng, c->iseed[2]);
      z0[j] = a * 1.0e-03;
      zp0[j] = p * 1.0e-03;
    }
  }
  if (dist) memcpy(dist, t0, sizeof(double) * 6 * (size_t)Np);
  memset(b0, 0, sizeof *b0);
  memset(w->npop, 0, sizeof w->npop);
  int *cell = malloc(sizeof(int) * (size_t)Np);
  for (int j = 0; j < Np; j++) {
    double Tau = -t0[j] / c->sgmatau;
    int ic = (int)((7.0 + Tau) / 0.3);
    if (ic < 1) ic = 0;
    if (ic > 2 * c->Nslc + 1) ic = 2 * c->Nslc + 1;
    cell[j] = ic;
    w->npop[ic]++;
    if (ic < WS_NC) {
      b0->xx[ic] = b0->xx[ic] + x0[j];
      b0->xp[ic] = b0->xp[ic] + xp0[j];
      b0->zz[ic] = b0->zz[ic] + z0[j];
      b0->zp[ic] = b0->zp[ic] + zp0[j];
    }
  }
  for (int ic = 0; ic < S; ic++)
    if (w->npop[ic]) {
      b0->xx[ic] = b0->xx[ic] / w->npop[ic];
      b0->xp[ic] = b0->xp[ic] / w->npop[ic];
      b0->zz[ic] = b0->zz[ic] / w->npop[ic];
      b0->zp[ic] = b0->zp[ic] / w->npop[ic];
    }
  memset(w->nofe, 0, sizeof w->nofe);
  for (int ic = 0; ic < S; ic++)
    for (int j = 0; j < Np; j++) {
      if (cell[j] != ic) continue;
      int nsz = (int)(2.0 * (z0[j] - b0->zz[ic]) / w->agLOCz), msz = nsz >= 0 ? (nsz + 1) / 2 : (nsz - 1) / 2;
      if (msz > WS_OFF) msz = WS_OFF;
      if (msz < -WS_OFF) msz = -WS_OFF;
      int nsx = (int)(2.0 * (x0[j] - b0->xx[ic]) / w->agLOCx), msx = nsx >= 0 ? (nsx + 1) / 2 : (nsx - 1) / 2;
      if (msx > WS_OFF) msx = WS_OFF;
      if (msx < -WS_OFF) msx = -WS_OFF;
      w->nofe[ic][msx + WS_OFF][msz + WS_OFF]++;
    }
  free(cell);
  free(t0);
  return 0;
}

void orc_ws_destroy(orc_ws *w)
{
  if (!w) return;
  free(w->Ibunch);
  free(w->b);
  free(w->fill);
  free(w->xxh);
  free(w->xph);
  free(w->zzh);
  free(w->zph);
  free(w->NriF);
  free(w->train_buf);
  free(w->xxF);
  free(w->xdF);
  free(w->yyF);
  free(w->ydF);
  free(w->Exb);
  free(w->Eyb);
  free(w->Nrib);
  free(w->xxib);
  free(w->xdib);
  free(w->yyib);
  free(w->ydib);
  free(w);
}

orc_ws *orc_ws_create(const mbt_config *c, int ibunch_mode, double *dist)
{
  if (c->Nslc < 0 || 2 * c->Nslc + 1 > WS_NC || (int)c->Nharm > WS_HMAX || c->Nmlt + 2 > WS_HIST) return NULL;
  orc_ws *w = calloc(1, sizeof *w);
  w->c = c;
  w->S = 2 * c->Nslc + 1;
  w->Nb = c->Nbunch;
  w->h = (int)c->Nharm;
  /* fbii_grid_setup_parameters, src/bunch.c:814-850 */
  w->agLOCz = c->sgm_zz / ((double)3) / 1.0e+03;
  w->agLOCx = c->sgm_xx / ((double)3) / 1.0e+03;
  w->agABSz = 2.0 * w->agLOCz;
  w->agABSx = 2.0 * w->agLOCx;
  w->deltT = 0.3 * c->sgmatau;
  w->factBE = sqrt(2 * (c->sgm_xx * c->sgm_xx - c->sgm_zz * c->sgm_zz)) / 1.0e+03;
  /* fbii_model_init, src/fbi_weakstrong.c:18-47 */
  {
    double Pave = 2.7e-9, Pu20 = 5 * Pave, Pm = 0.15 * Pu20 / 1.3332, dm = 3.22e22 * Pm, sigmam = 1.78;
    double Lu20 = 354.0;
    w->Ldmsgm = Lu20 * dm * sigmam * 1e-22;
    w->factF = K_E / (28.0 * K_MPOME * K_ME);
  }
  w->Ibunch = calloc(w->h, sizeof(double));
  if (ibunch_mode == ORC_WS_IBUNCH_FILL) memcpy(w->Ibunch, c->Ibunch, sizeof(double) * w->h);
  w->fill = malloc(sizeof(int) * w->h);
  memcpy(w->fill, c->nfFill, sizeof(int) * w->h);
  w->b = calloc(w->h, sizeof(ws_bunch));
  w->xxh = calloc(w->h, sizeof *w->xxh);
  w->xph = calloc(w->h, sizeof *w->xph);
  w->zzh = calloc(w->h, sizeof *w->zzh);
  w->zph = calloc(w->h, sizeof *w->zph);
  const size_t per = (size_t)w->S * WS_NG * WS_NG;
  w->capF = (long)w->Nb * (long)per + 1;
  w->NriF = malloc(sizeof(double) * w->capF);
  w->train_buf = malloc(sizeof(double) * WS_NC * ORC_ION_BLOCK * 2);
  w->xxF = malloc(sizeof(double) * w->capF);
  w->xdF = malloc(sizeof(double) * w->capF);
  w->yyF = malloc(sizeof(double) * w->capF);
  w->ydF = malloc(sizeof(double) * w->capF);
  w->Exb = malloc(sizeof(double) * per * w->h);
  w->Eyb = malloc(sizeof(double) * per * w->h);
  w->Nrib = malloc(sizeof(double) * per * w->h);
  w->xxib = malloc(sizeof(double) * per * w->h);
  w->xdib = malloc(sizeof(double) * per * w->h);
  w->yyib = malloc(sizeof(double) * per * w->h);
  w->ydib = malloc(sizeof(double) * per * w->h);
  /* the manager's rank: eCMs (written out there) then the per-bucket CM offsets
   * of generate_initial_general_distribution2, src/bunch_strong.c:10-50 */
  void *rng = orc_rank_rng_new();
  ws_bunch b0;
  if (gen_ecms(w, rng, &b0, dist)) {
    free(rng);
    orc_ws_destroy(w);
    return NULL;
  }
  for (int kb = 0; kb < w->h; kb++) {
    if (!w->fill[kb]) continue;
    double xo = c->xxCM_offset * 1.0e-03 * orc_rank_c_fnorm(rng, 571030);
    double po = c->xpCM_offset * 1.0e-03 * orc_rank_c_fnorm(rng, 571030);
    double zo = c->zzCM_offset * 1.0e-03 * orc_rank_c_fnorm(rng, 571030);
    double qo = c->zpCM_offset * 1.0e-03 * orc_rank_c_fnorm(rng, 571030);
    ws_bunch *b = &w->b[kb];
    for (int ic = 0; ic < w->S; ic++) {
      b->xtau[ic] = b0.xtau[ic];
      b->xeps[ic] = b0.xeps[ic];
      b->xx[ic] = b0.xx[ic] + xo;
      b->xp[ic] = b0.xp[ic] + po;
      b->zz[ic] = b0.zz[ic] + zo;
      b->zp[ic] = b0.zp[ic] + qo;
    }
  }
  free(rng);
  return w;
}

/* ------------------------------------------------------------- the turn */

/* bunch_strong_updateCM, src/bunch_strong.c:221-266 */
static void update_cm(const orc_ws *w, ws_bunch *b, int plane)
{
  const double *u = plane == 0 ? b->xtau : plane == 1 ? b->xx : b->zz;
  const double *p = plane == 0 ? b->xeps : plane == 1 ? b->xp : b->zp;
  double uu = 0.0, up = 0.0;
  int n = 0;
  for (int ic = 0; ic < w->S; ic++) {
    uu += u[ic] * ((double)w->npop[ic]);
    up += p[ic] * ((double)w->npop[ic]);
    n += w->npop[ic];
  }
  b->pos[plane] = uu / ((double)n);
  b->slope[plane] = up / ((double)n);
}

/* transformSbunch_shortrange, src/rw_weakstrong.c:10-67 (old xeps in the tau update) */
static void shortrange(const orc_ws *w, ws_bunch *b)
{
  const mbt_config *c = w->c;
  for (int ic = 0; ic < w->S; ic++) {
    double gain = c->Vrf0 / (c->E0 * 1.0e+03) * sens_ulp(sin(c->wrf * b->xtau[ic] + c->phai0), b->xtau[ic]);
    double e0 = b->xeps[ic];
    b->xeps[ic] = e0 + gain - c->Urad;
    b->xtau[ic] = b->xtau[ic] - e0 * c->T0 * c->ac;
  }
  update_cm(w, b, 0);
  update_cm(w, b, 1);
  update_cm(w, b, 2);
}

/* get_linearly_interpolated_Efields, src/fbi_weakstrong.c:454-580 */
static void interp(const orc_ws *w, const double *Ex, const double *Ey, int ic, double xx, double yy, double E[2])
{
  const int off = WS_OFF;
  const double ax = w->agABSx, az = w->agABSz;
  #ifdef ORC_SENS_RECIP
  const double iax = 1.0 / ax, iaz = 1.0 / az;
  int k1x = (int)(xx * iax), k1z = (int)(yy * iaz), k2x, k2z;
#else
  int k1x = (int)(xx / ax), k1z = (int)(yy / az), k2x, k2z;
#endif
  const double *ex = Ex + ic * WS_NG * WS_NG, *ey = Ey + ic * WS_NG * WS_NG;
#define IDX(kx, kz) (((kx) + off) * WS_NG + ((kz) + off))
  double sx, sz; /* +1 / -1: direction of the interpolation */
  if (yy >= 0) {
    k2z = k1z < off ? k1z + 1 : k1z;
    if (k1z > off) k1z = off;
    if (k2z > off) k2z = off;
    sz = 1.0;
  } else {
    k2z = k1z > -off ? k1z - 1 : k1z;
    if (k1z < -off) k1z = -off;
    if (k2z < -off) k2z = -off;
    sz = -1.0;
  }
  if (xx >= 0) {
    k2x = k1x < off ? k1x + 1 : k1x;
    if (k1x > off) k1x = off;
    if (k2x > off) k2x = off;
    sx = 1.0;
  } else {
    k2x = k1x > -off ? k1x - 1 : k1x;
    if (k1x < -off) k1x = -off;
    if (k2x < -off) k2x = -off;
    sx = -1.0;
  }
  const int i1 = IDX(k1x, k1z), i2 = IDX(k2x, k1z), i3 = IDX(k1x, k2z), i4 = IDX(k2x, k2z);
#undef IDX
  #ifdef ORC_SENS_RECIP
  const double fx = (xx - k1x * ax) * iax, fz = (yy - k1z * az) * iaz;
#else
  const double fx = (xx - k1x * ax) / ax, fz = (yy - k1z * az) / az;
#endif
  double y5, y6, x5, x6;
  if (sx > 0) {
    y5 = ey[i1] + fx * (ey[i2] - ey[i1]);
    y6 = ey[i3] + fx * (ey[i4] - ey[i3]);
    x5 = ex[i1] + fx * (ex[i2] - ex[i1]);
    x6 = ex[i3] + fx * (ex[i4] - ex[i3]);
  } else {
    y5 = ey[i1] - fx * (ey[i2] - ey[i1]);
    y6 = ey[i3] - fx * (ey[i4] - ey[i3]);
    x5 = ex[i1] - fx * (ex[i2] - ex[i1]);
    x6 = ex[i3] - fx * (ex[i4] - ex[i3]);
  }
  if (sz > 0) {
    E[1] = y5 + fz * (y6 - y5);
    E[0] = x5 + fz * (x6 - x5);
  } else {
    E[1] = y5 - fz * (y6 - y5);
    E[0] = x5 - fz * (x6 - x5);
  }
}

/* fbi_weakstrong_bunch, src/fbi_weakstrong.c:49-298: fields on the absolute
 * grid (Bassetti-Erskine) and the ions this bunch creates, kicked by later
 * slices of the same bunch */
static void fbi_bunch(orc_ws *w, const ws_bunch *b, int kb)
{
  const mbt_config *c = w->c;
  const double Nelec = (w->Ibunch[kb] * c->T0 / (c->Np * K_E * 1.0e+03));
  const double Convion = Nelec * w->Ldmsgm;
  const double sxx = c->sgm_xx, szz = c->sgm_zz, fBE = w->factBE;
  int ic = 0;
  for (int cell = 0; cell < w->S; cell++) {
    double dQ = Nelec * w->npop[cell] * K_E;
    double factdQ = dQ / (2 * K_EPS0 * sqrt(M_PI) * fBE * K_C * w->deltT);
    for (int msx = -WS_OFF; msx < WS_OFF + 1; msx++) {
      double xabs = msx * w->agABSx, Dx = xabs - b->xx[cell];
      double xi1 = Dx / fBE, xi2 = xi1 * szz / sxx;
      for (int msz = -WS_OFF; msz < WS_OFF + 1; msz++) {
        double yabs = msz * w->agABSz, Dy = yabs - b->zz[cell];
        double yi1 = Dy / fBE, yi2 = yi1 * sxx / szz;
        double u1[2], u2[2];
        if (orc_wofz(xi1, yi1, &u1[0], &u1[1]) || orc_wofz(xi2, yi2, &u2[0], &u2[1])) w->wofz_fail++;
        double g = sens_ulp_exp(-(Dx * Dx * 1.0e+06 / (2 * sxx * sxx) + Dy * Dy * 1.0e+06 / (2 * szz * szz)));
        w->Ex[ic] = -factdQ * (u1[1] - g * u2[1]);
        w->Ey[ic] = -factdQ * (u1[0] - g * u2[0]);
        ic++;
      }
    }
  }
  int ip = 0;
  for (int cell = 0; cell < w->S; cell++) {
    int ip0 = ip;
    for (int msx = -WS_OFF; msx < WS_OFF + 1; msx++)
      for (int msz = -WS_OFF; msz < WS_OFF + 1; msz++) {
        w->Nri[ip] = Convion * ((double)w->nofe[cell][msx + WS_OFF][msz + WS_OFF]);
        w->xxi[ip] = b->xx[cell] + msx * w->agLOCx;
        w->xdi[ip] = 0.0;
        w->yyi[ip] = b->zz[cell] + msz * w->agLOCz;
        w->ydi[ip] = 0.0;
        ip++;
      }
    for (int q = 0; q < ip0; q++) {
      double E[2];
      interp(w, w->Ex, w->Ey, cell, w->xxi[q], w->yyi[q], E);
      double ax = w->factF * E[0];
      w->xdi[q] = w->xdi[q] + w->deltT * ax;
      w->xxi[q] = w->xxi[q] + w->deltT * w->xdi[q];
      double ay = w->factF * E[1];
      w->ydi[q] = w->ydi[q] + w->deltT * ay;
      w->yyi[q] = w->yyi[q] + w->deltT * w->ydi[q];
    }
  }
  w->Nions = ip;
}

/* bunch_strong_distribution_statistics, src/bunch.c:516-550: out[18] =
 * pos[3] pos_sigma[3] slope[3] slope_sigma[3] ampinv[3] ampinv_cm[3] */
static void strong_stats(const orc_ws *w, ws_bunch *b, double *out)
{
  const mbt_config *c = w->c;
  memset(out, 0, sizeof(double) * 18);
  for (int pl = 0; pl < 3; pl++) {
    if (!c->TrackPlane[pl]) continue;
    const double *u = pl == 0 ? b->xtau : pl == 1 ? b->xx : b->zz;
    const double *p = pl == 0 ? b->xeps : pl == 1 ? b->xp : b->zp;
    update_cm(w, b, pl);
    double avr = b->pos[pl], avs = b->slope[pl];
    double sq = 0.0, sqs = 0.0;
    int ws = 0, wss = 0;
    for (int ic = 0; ic < w->S; ic++) {
      sq += (u[ic] - avr) * (u[ic] - avr) * w->npop[ic];
      ws += w->npop[ic];
    }
    for (int ic = 0; ic < w->S; ic++) {
      sqs += (p[ic] - avs) * (p[ic] - avs) * w->npop[ic];
      wss += w->npop[ic];
    }
    double amp = 0.0;
    int n = 0;
    for (int ic = 0; ic < w->S; ic++) {
      amp += w->npop[ic] * (c->gamma1[pl] * pow(u[ic], 2) + 2.0 * c->alpha1[pl] * u[ic] * p[ic] +
                            c->beta1[pl] * pow(p[ic], 2));
      n += w->npop[ic];
    }
    out[pl] = avr;
    out[3 + pl] = sqrt(sq / ((double)ws));
    out[6 + pl] = avs;
    out[9 + pl] = sqrt(sqs / ((double)wss));
    out[12 + pl] = amp / ((double)n);
    out[15 + pl] = c->gamma1[pl] * pow(avr, 2) + 2.0 * c->alpha1[pl] * avr * avs + c->beta1[pl] * pow(avs, 2);
  }
}

/* transformSbunch_optic, src/rw_weakstrong.c:194-253 */
static void optic(const orc_ws *w, ws_bunch *b)
{
  const mbt_config *c = w->c;
  double PsiV0 = 2.0 * M_PI * c->QV0 / 1.0, PsiH0 = 2.0 * M_PI * c->QH0 / 1.0;
  double cv = cos(PsiV0), sv = sin(PsiV0);
  double v11 = cv + c->alpha1[2] * sv, v12 = c->beta1[2] * sv, v21 = -c->gamma1[2] * sv, v22 = cv - c->alpha1[2] * sv;
  double ch = cos(PsiH0), sh = sin(PsiH0);
  double h11 = ch + c->alpha1[1] * sh, h12 = c->beta1[1] * sh, h21 = -c->gamma1[1] * sh, h22 = ch - c->alpha1[1] * sh;
  for (int ic = 0; ic < w->S; ic++) {
    double y = v11 * b->zz[ic] + v12 * b->zp[ic];
    double yp = v21 * b->zz[ic] + v22 * b->zp[ic];
    b->zz[ic] = y;
    b->zp[ic] = yp;
    double x = h11 * b->xx[ic] + h12 * b->xp[ic];
    double xp = h21 * b->xx[ic] + h22 * b->xp[ic];
    b->xx[ic] = x;
    b->xp[ic] = xp;
  }
}

/* transformSbunch_longrange, src/rw_weakstrong.c:69-192, on the worker's copy
 * of the history (hist[jb][j], zero where jb + j h >= 5000) */
static void longrange(const orc_ws *w, ws_bunch *b, int kb, int plane, long in)
{
  const mbt_config *c = w->c;
  double Psi0 = 2.0 * M_PI * (plane == 1 ? c->QH0 : c->QV0) / 1.0;
  double beff3 = pow(plane == 1 ? c->beffH[0] : c->beffV[0], 3);
  if (c->triggerRW && in < 50) beff3 = pow(2.0e-3, 3);
  const double Pref0 = c->T0 / c->E0 / c->fNp / 1.0e+12;
  const double rhorw = 2.8e-08, kappa = 1.0 / rhorw;
  const double fNharm = ((double)c->Nharm), Dub0 = c->Lc / fNharm;
  const double prfrw = 2.0 * c->R / (beff3) / sqrt(M_PI * kappa * K_EPS0) / 1.0;
  double Psi = 0.0 + Psi0;
  double cp = cos(Psi), sp = sin(Psi);
  double a11 = cp + c->alpha1[plane] * sp, a12 = +c->beta1[plane] * sp;
  double *up = plane == 1 ? b->xp : b->zp;
  const double(*uh)[WS_HIST] = plane == 1 ? (const double(*)[WS_HIST])w->xxh : (const double(*)[WS_HIST])w->zzh;
  const double(*ph)[WS_HIST] = plane == 1 ? (const double(*)[WS_HIST])w->xph : (const double(*)[WS_HIST])w->zph;
  for (int jb = 0; jb < w->h; jb++) {
    if (!w->fill[jb]) continue;
    double ujb[WS_HIST];
    for (int j = 0; j < c->Nmlt + 2; j++) {
      int in_msg = jb + j * w->h < WS_MSG;
      double u = in_msg ? uh[jb][j] : 0.0, p = in_msg ? ph[jb][j] : 0.0;
      ujb[j] = a11 * u + a12 * p;
    }
    double Dup = 0.0, D0;
    if (jb == kb) D0 = c->Lc;
    else if (jb > kb) D0 = ((double)(jb - kb)) * Dub0;
    else D0 = ((double)(w->h + jb - kb)) * Dub0;
    for (int j = 0; j < c->Nmlt + 1; j++) {
      double Du = D0 + ((double)j) * c->Lc;
      double Dt = Du / K_C;
      double W = prfrw / sqrt(Dt);
      int jm = jb > kb ? j : j + 1;
      double dip = c->fNp * ujb[jm];
      Dup = Dup + dip * W;
    }
    Dup = Pref0 * w->Ibunch[jb] * Dup;
    if (in < c->Nmlt) Dup = 0.0;
    for (int ic = 0; ic < w->S; ic++) up[ic] += Dup;
  }
}

/* fbi_weakstrong_train, src/fbi_weakstrong.c:301-451, for filled bunch ib in bucket kb */
static void fbi_train(orc_ws *w, ws_bunch *b, int ib, int kb, const double *Ex, const double *Ey, const double *Nri,
                      const double *xxi, const double *xdi, const double *yyi, const double *ydi)
{
  const mbt_config *c = w->c;
  if (ib == 0) {
    w->nF = 0;
    for (int q = 0; q < w->Nions; q++) {
      w->NriF[w->nF] = Nri[q];
      w->yyF[w->nF] = yyi[q];
      w->ydF[w->nF] = ydi[q];
      w->xxF[w->nF] = xxi[q];
      w->xdF[w->nF] = xdi[q];
      w->nF++;
    }
    w->kbF = kb;
    return;
  }
  const double factjp0 = 1e+2 * 4.0 * M_PI * K_EPS0 * K_RE / c->Gamma;
  const double Nelec = (w->Ibunch[kb] * c->T0 / (c->Np * K_E * 1.0e+03));
  const double Dtime = fabs((double)(kb - w->kbF)) * (c->T0 / c->h);
  for (long q = 0; q < w->nF; q++) {
    w->xxF[q] = w->xxF[q] + Dtime * w->xdF[q];
    w->yyF[q] = w->yyF[q] + Dtime * w->ydF[q];
  }
  /* The reference loops slice-major (every ion of the cloud at slice 0, then at
   * slice 1, ...).  An ion's path through the slices depends on nothing but its
   * own state and the fields, so the ions are taken in blocks, every ion of a
   * block carried through all slices on the threads at hand, with each
   * Nri * E product kept; the per-slice sums then add those products in ion
   * order, slice by slice -- the reference's own operations in its order, so
   * the result is bitwise the same on any thread count. */
  double sx[WS_NC], sy[WS_NC];
  for (int cell = 0; cell < w->S; cell++) sx[cell] = sy[cell] = 0.0;
  double *pe = w->train_buf;
  for (long q0 = 0; q0 < w->nF; q0 += ORC_ION_BLOCK) {
    const long nq = w->nF - q0 < ORC_ION_BLOCK ? w->nF - q0 : ORC_ION_BLOCK;
#pragma omp parallel for schedule(static)
    for (long j = 0; j < nq; j++) {
      const long q = q0 + j;
      double xx = w->xxF[q], xd = w->xdF[q], yy = w->yyF[q], yd = w->ydF[q];
      const double nri = w->NriF[q];
      for (int cell = 0; cell < w->S; cell++) {
        double E[2];
        interp(w, Ex, Ey, cell, xx, yy, E);
        double ax = w->factF * E[0];
        xd = xd + w->deltT * ax;
        xx = xx + w->deltT * xd;
        double ay = w->factF * E[1];
        yd = yd + w->deltT * ay;
        yy = yy + w->deltT * yd;
        pe[((size_t)cell * ORC_ION_BLOCK + j) * 2] = nri * E[0];
        pe[((size_t)cell * ORC_ION_BLOCK + j) * 2 + 1] = nri * E[1];
      }
      w->xxF[q] = xx;
      w->xdF[q] = xd;
      w->yyF[q] = yy;
      w->ydF[q] = yd;
    }
#pragma omp parallel for schedule(static)
    for (int cell = 0; cell < w->S; cell++) {
      const double *p = pe + (size_t)cell * ORC_ION_BLOCK * 2;
#ifdef ORC_SENS_TREE
      sx[cell] = sx[cell] + sens_pairwise(p, nq, 2);
      sy[cell] = sy[cell] + sens_pairwise(p + 1, nq, 2);
#else
      double a = sx[cell], bb = sy[cell];
      for (long j = 0; j < nq; j++) {
        a = a + p[2 * j];
        bb = bb + p[2 * j + 1];
      }
      sx[cell] = a;
      sy[cell] = bb;
#endif
    }
  }
  for (int cell = 0; cell < w->S; cell++) {
    double dQ = Nelec * w->npop[cell] * K_E;
    double factjp = dQ > 0.0 ? factjp0 / dQ : 0.0;
    double fx = factjp * sx[cell];
    double fy = factjp * sy[cell];
    b->xp[cell] = b->xp[cell] + K_C * w->deltT * fx;
    b->zp[cell] = b->zp[cell] + K_C * w->deltT * fy;
  }
  for (int q = 0; q < w->Nions; q++) {
    w->NriF[w->nF] = Nri[q];
    w->yyF[w->nF] = yyi[q];
    w->ydF[w->nF] = ydi[q];
    w->xxF[w->nF] = xxi[q];
    w->xdF[w->nF] = xdi[q];
    w->nF++;
  }
  w->kbF = kb;
}

/* bunch_CM_history_append, src/bunch.c:94-115 */
static void hist_append(orc_ws *w, int kb, const ws_bunch *b)
{
  const int NmT = w->c->NmultiT;
  for (int i = 0; i < NmT; i++) {
    int j = NmT - i - 2;
    if (j >= 0) {
      w->xxh[kb][j + 1] = w->xxh[kb][j];
      w->xph[kb][j + 1] = w->xph[kb][j];
      w->zzh[kb][j + 1] = w->zzh[kb][j];
      w->zph[kb][j + 1] = w->zph[kb][j];
    } else {
      w->xxh[kb][0] = b->pos[1];
      w->xph[kb][0] = b->slope[1];
      w->zzh[kb][0] = b->pos[2];
      w->zph[kb][0] = b->slope[2];
    }
  }
}

/* Tracks nturns turns.  Outputs per turn (any may be NULL):
 *   cm[t][ib][4]   the CM (x, x', z, z') each bunch sends after phase 1
 *   mon[t][ib][18] statistics of monitor turns (rev % NrevMon == 0), else zero
 *   ion[t][ib][3]  after bunch ib's train step: sum Nri x, sum Nri y, sum Nri
 *                  over the ion cloud (the sums of writeoutCMImotions)
 *   zcm[t][kb]     the slice-averaged z of every bucket at the end of the turn
 *                  (writeoutVCMmotionsPM2), 0 for empty buckets */
int orc_ws_track(orc_ws *w, long nturns, double *cm, double *mon, double *ion, double *zcm)
{
  const mbt_config *c = w->c;
  const size_t per = (size_t)w->S * WS_NG * WS_NG;
  for (long t = 0; t < nturns; t++, w->rev++) {
    const long rev = w->rev;
    int ib = 0;
    for (int kb = 0; kb < w->h; kb++) {
      if (!w->fill[kb]) continue;
      ws_bunch *b = &w->b[kb];
      shortrange(w, b);
      if (c->EnableFBII) {
        fbi_bunch(w, b, kb);
        memcpy(w->Exb + per * kb, w->Ex, sizeof(double) * per);
        memcpy(w->Eyb + per * kb, w->Ey, sizeof(double) * per);
        memcpy(w->Nrib + per * kb, w->Nri, sizeof(double) * per);
        memcpy(w->xxib + per * kb, w->xxi, sizeof(double) * per);
        memcpy(w->xdib + per * kb, w->xdi, sizeof(double) * per);
        memcpy(w->yyib + per * kb, w->yyi, sizeof(double) * per);
        memcpy(w->ydib + per * kb, w->ydi, sizeof(double) * per);
      }
      if (cm) {
        double *o = cm + ((size_t)t * w->Nb + ib) * 4;
        o[0] = b->pos[1];
        o[1] = b->slope[1];
        o[2] = b->pos[2];
        o[3] = b->slope[2];
      }
      if (rev % c->NrevMon == 0 && mon) strong_stats(w, b, mon + ((size_t)t * w->Nb + ib) * 18);
      else if (mon) memset(mon + ((size_t)t * w->Nb + ib) * 18, 0, sizeof(double) * 18);
      hist_append(w, kb, b);
      ib++;
    }
    for (int kb = 0; kb < w->h; kb++) {
      if (!w->fill[kb]) continue;
      ws_bunch *b = &w->b[kb];
      optic(w, b);
      if (c->EnableRW_long) {
        if (c->TrackPlane[1]) longrange(w, b, kb, 1, rev);
        if (c->TrackPlane[2]) longrange(w, b, kb, 2, rev);
      }
      update_cm(w, b, 1);
      update_cm(w, b, 2);
    }
    w->nF = 0;
    w->kbF = 0;
    ib = 0;
    for (int kb = 0; kb < w->h; kb++) {
      if (!w->fill[kb]) continue;
      if (c->EnableFBII) {
        w->Nions = (int)per;
        fbi_train(w, &w->b[kb], ib, kb, w->Exb + per * kb, w->Eyb + per * kb, w->Nrib + per * kb,
                  w->xxib + per * kb, w->xdib + per * kb, w->yyib + per * kb, w->ydib + per * kb);
      }
      if (ion) {
        double sx = 0.0, sy = 0.0, sn = 0.0;
#ifdef ORC_SENS_TREE
        sens_cloud_sums(w->NriF, w->xxF, w->yyF, w->nF, &sx, &sy, &sn);
#else
        for (long q = 0; q < w->nF; q++) {
          sx = sx + w->NriF[q] * w->xxF[q];
          sy = sy + w->NriF[q] * w->yyF[q];
          sn = sn + w->NriF[q];
        }
#endif
        double *o = ion + ((size_t)t * w->Nb + ib) * 3;
        o[0] = sx;
        o[1] = sy;
        o[2] = sn;
      }
      ib++;
    }
    if (zcm)
      for (int kb = 0; kb < w->h; kb++) {
        double z = 0.0;
        if (w->fill[kb]) {
          for (int ic = 0; ic < w->S; ic++) z = z + w->b[kb].zz[ic];
          z = z / ((double)w->S);
        }
        zcm[(size_t)t * w->h + kb] = z;
      }
  }
  return w->wofz_fail ? 1 : 0;
}

/* Slice arrays of filled bunch ordinal ib: out[6][S] = xtau xeps xx xp zz zp */
int orc_ws_get_slices(const orc_ws *w, int ib, double *out)
{
  int k = 0;
  for (int kb = 0; kb < w->h; kb++) {
    if (!w->fill[kb]) continue;
    if (k++ != ib) continue;
    const ws_bunch *b = &w->b[kb];
    const double *src[6] = {b->xtau, b->xeps, b->xx, b->xp, b->zz, b->zp};
    for (int a = 0; a < 6; a++) memcpy(out + (size_t)a * w->S, src[a], sizeof(double) * w->S);
    return 0;
  }
  return -1;
}

/* npop[S] and Nofe[S][5][5] of the generated bunch */
int orc_ws_population(const orc_ws *w, int *npop, int *nofe)
{
  for (int ic = 0; ic < w->S; ic++) npop[ic] = w->npop[ic];
  if (nofe)
    for (int ic = 0; ic < w->S; ic++)
      for (int a = 0; a < WS_NG; a++)
        for (int q = 0; q < WS_NG; q++) nofe[(ic * WS_NG + a) * WS_NG + q] = w->nofe[ic][a][q];
  return w->S;
}
