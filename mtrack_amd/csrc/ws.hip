// The weak-strong multi-bunch turn with fast beam-ion interaction (FBII) on one
// MI355X: the reference's BunchModel = STRONG path, manager_weakstrong +
// worker_weakstrong (src/manager_weakstrong.c:7-263, src/worker_weakstrong.c:14-247).
//
// Each filled bunch is 2 Nslc + 1 slice centroids (tau, delta, x, x', z, z')
// resident in HBM.  One turn is four launch groups on one stream:
//
//   k_ws_phase1  one block per bunch: synchrotron step of the centroids
//                (transformSbunch_shortrange, src/rw_weakstrong.c:10-67), bunch
//                centroid, CM history append (src/bunch.c:94-115), monitor-turn
//                statistics (src/bunch.c:516-550); with FBII, one thread per
//                point, the Bassetti-Erskine field of every slice on the 5 x 5
//                absolute grid (two Faddeeva evaluations per point), then
//                k_ws_create: the ions each bunch creates, one thread each,
//                kicked by the bunch's later slices (fbi_weakstrong_bunch,
//                src/fbi_weakstrong.c:49-298);
//   k_ws_phase2  one block per bunch: betatron rotation of the centroids
//                (transformSbunch_optic, :194-253) and the resistive-wall kick of
//                every other bucket's centroid history (transformSbunch_longrange,
//                :69-192), one lane per source bucket;
//   train        the reference's manager relay (fbi_weakstrong_train
//                :301-451): the accumulated ion cloud drifts to each bunch in
//                train order, passes its slices (bilinear field interpolation,
//                src/fbi_weakstrong.c:454-580), kicks the slice centroids with
//                the summed ion field, and the bunch's own ions join the cloud.
//                An ion's path depends only on the field tables phase 1 made
//                and on the drift times, never on the kicks the train gives the
//                electrons, so the whole relay is one launch with one lane per
//                ion over every later bunch (k_ws_ions: per-slice wave sums,
//                block partials per bunch), a per-bunch, per-quantity sum of
//                the partials (k_ws_ksum) and the kicks (in k_ws_end);
//                k_ws_cloud_reset + k_ws_train keep the bunch-by-bunch relay
//                (one launch per bunch) for MBT_WS_SERIAL_TRAIN and for trains
//                whose partials would not fit;
//   k_ws_end     the ion-parallel train's kicks and cloud sums, and the
//                slice-averaged vertical centroid of every bucket.
//
// Everything but the ion-field sums, sin and the Faddeeva function keeps the
// reference's operation order; constants derived with libm are computed on the
// host with glibc.  Results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mbt_host.h"
#include "wofz_dev.h"

#pragma clang fp contract(off)

#define WS_S 50       // slice capacity of the reference's bunch_strong_t
#define WS_PTS 1250   // 50 slices x 25 grid points
#define WS_TQ (2 * WS_S + 3 + 12) // per bunch in WSD::tot: 2 S + 3 train sums, 4 x 3 own-ion sums
#define WS_HIST 30    // CM history depth (src/bunch.h:155)
#define WS_MSG 5000   // doubles of the history message (src/manager_weakstrong.c:170-173)
#define WS_T 256      // turns of outputs per device buffer block
#define WS_WG 256
// device memory the ion-parallel train may use for its per-bunch partials
#define WS_TRAIN_PARTIAL_MAX (8LL << 30)

#define K_C 2.997925e+08
#define K_EPS0 8.854187817620e-12
#define K_E 1.602e-19

#define CK(x)                                                                                    \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      mbt_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);     \
      return MBT_EDEVICE;                                                                        \
    }                                                                                            \
  } while (0)

extern "C" int mbt_ws_generate(const mbt_config *c, double *slices, int *npop, int *nofe, double *dist);
extern "C" void mbt_ws_grid(const mbt_config *c, double *agLOCx, double *agLOCz);

struct WSP {
  int S, G, h, Nmlt, NmultiT, fbii, rw, track_h, track_v, track[3], Nions;
  int mixed; // Faddeeva function of the field tables in FP32 (MBT_WS_MIXED)
  double eps, wrf, phai0, Urad, T0, ac;                 // synchrotron step
  double gamma1[3], alpha1[3], beta1[3];                // statistics
  double v11, v12, v21, v22, h11, h12, h21, h22;        // one-turn rotation (glibc cos/sin)
  double rwa11[2], rwa12[2], prfrw[2], prfrw_trig[2];   // RW: [HOR, VER]
  int triggerRW;
  double Pref0, fNp, Dub0, Lc;
  double factBE, sxx, szz, deltT, aAx, aAz, aLx, aLz, den_dQ; // FBII fields
  double iaAx, iaAz;                                          // 1 / aAx, 1 / aAz
  double factF, factjp0, CdeltT;
  double dtF; // deltT * factF: the ion train's velocity kick per unit field
};

struct WSD {
  double *sl;       // [G][6][S] xtau xeps xx xp zz zp
  const int *npop;  // [S]
  const int *nofe;  // [S*25]
  const int *bucket, *fill; // [G], [h]
  const double *Ibunch, *nelec, *convion; // [h] mA, [G], [G]
  double *hist;     // [4][h][WS_HIST] x x' z z' (the manager's history)
  double2 *Exy;     // [G][S*25] (Ex, Ey) of every grid point
  double *li;       // [5][G][S*25] local ions: Nri x x' y y'
  double *cl;       // [5][cap] ion cloud
  long long cap;
  double *part;     // [blocks][S][2]
  double *psum;     // [blocks][3]
  unsigned *done;
  const double *dtime;    // [G] drift time from the previous bunch
  double *tpart;          // ion-parallel train: per bunch ib, [2 S + 3][blocks with an ion ahead of ib]
  const long long *toff;  // [G] offset of bunch ib's partials in tpart
  int nblk_ions;          // blocks of k_ws_ions
  double *tot;            // ion-parallel train: [G][WS_TQ] per-bunch sums of the partials and own-ion sums
  double *cm, *mon, *ion, *zcm; // output blocks [WS_T][G][4|18|3], [WS_T][h]
};

// ------------------------------------------------------------ device helpers

// A double from another lane of the same row of 16 by a DPP move (no LDS
// permute): CTRL 0x140 = row_mirror (lane 15 - i), 0x141 = row_half_mirror
// (lane 7 - i within each 8), 0x4E = quad_perm [2,3,0,1], 0xB1 = quad_perm
// [1,0,3,2].  Every partner is an involution, so both lanes of a pair form the
// same sum and a butterfly over these four steps leaves the same total in all
// 16 lanes, as the xor butterfly does.  Every partner is a lane of the row, so
// bound_ctrl changes no value; it spares the move of the unused old operand.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Sum over the row of 16 lanes, in every lane of the row.
__device__ __forceinline__ double rowsum16(double v)
{
  v += dpp_d<0x140>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0xB1>(v);
  return v;
}

__device__ __forceinline__ double wsum64(double v)
{
  v += __shfl_xor(v, 32, 64);
  v += __shfl_xor(v, 16, 64);
  return rowsum16(v);
}

// One halving exchange over lane distance D = 32 or 16 without selects: lanes
// with bit D clear return a + (a of the lane D up), lanes with it set (b of the
// lane D down) + b.  v_permlane32_swap exchanges lanes 32-63 of its first
// operand with lanes 0-31 of its second, v_permlane16_swap the odd rows of 16
// of the first with the even rows of the second; one per 32-bit half.
template <int D>
__device__ __forceinline__ double swap_sum(double a, double b)
{
  const unsigned long long ia = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ib = (unsigned long long)__double_as_longlong(b);
  unsigned lo0, lo1, hi0, hi1;
  if (D == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ia, (unsigned)ib, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ia >> 32), (unsigned)(ib >> 32), false, false);
    lo0 = lo[0], lo1 = lo[1], hi0 = hi[0], hi1 = hi[1];
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ia, (unsigned)ib, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ia >> 32), (unsigned)(ib >> 32), false, false);
    lo0 = lo[0], lo1 = lo[1], hi0 = hi[0], hi1 = hi[1];
  }
  const double x = __longlong_as_double((long long)(((unsigned long long)hi0 << 32) | lo0));
  const double y = __longlong_as_double((long long)(((unsigned long long)hi1 << 32) | lo1));
  return x + y;
}

// Eight sums over the 64 lanes of a wave at once: three halving exchanges (xor
// 32, 16, 8) split the eight values between lane groups, three more (xor 4, 2,
// 1) finish each; lanes 8k .. 8k+7 return the total of v[k].  Ten exchanges
// instead of 48 for eight separate reductions; a fixed tree, so reproducible.
__device__ __forceinline__ double wsum64x8(const double v[8])
{
  const bool b3 = threadIdx.x & 8;
  double w[4], x[2];
#pragma unroll
  for (int k = 0; k < 4; k++) w[k] = swap_sum<32>(v[k], v[4 + k]);
#pragma unroll
  for (int k = 0; k < 2; k++) x[k] = swap_sum<16>(w[k], w[2 + k]);
  // bit 3 by the row mirror (its partner 15 - i has bit 3 flipped), then the
  // remaining three bits within each half row
  double y = (b3 ? x[1] : x[0]) + dpp_d<0x140>(b3 ? x[0] : x[1]);
  y += dpp_d<0x141>(y);
  y += dpp_d<0x4E>(y);
  y += dpp_d<0xB1>(y);
  return y;
}

// bunch_strong_updateCM, src/bunch_strong.c:221-266 (sequential, one lane)
__device__ void ws_cm(const double *u, const double *p, const int *npop, int S, double &pos, double &slope)
{
  double uu = 0.0, up = 0.0;
  int n = 0;
  for (int ic = 0; ic < S; ic++) {
    uu += u[ic] * ((double)npop[ic]);
    up += p[ic] * ((double)npop[ic]);
    n += npop[ic];
  }
  pos = uu / ((double)n);
  slope = up / ((double)n);
}

// get_linearly_interpolated_Efields, src/fbi_weakstrong.c:454-580, on one
// slice's 25 field values, stored as (Ex, Ey) pairs so that one 16-byte LDS
// read fetches both fields of a grid point.  The grid quotients multiply by the
// reciprocal spacings (iax = 1 / aAx, iaz = 1 / aAz) in place of the
// reference's four FP64 divisions per interaction; such a quotient can differ
// from the divided one in its last bit, which moves a cell choice only where
// the bilinear field is continuous.
__device__ __forceinline__ void ws_interp(const double2 *t, double aAx, double aAz, double iax, double iaz, double xx,
                                          double yy, double &Ex, double &Ey)
{
  // the bilinear steps as fused multiply-adds (one rounding less each)
#pragma clang fp contract(fast)
  // Grid point (kx, kz), kx, kz in [-2, 2], is entry 5 kx + kz + 12.  The cell
  // is k1 = trunc(x / d) clamped to the grid, its partner one step further from
  // the centre in the direction of the coordinate's sign (the same point at the
  // clamped edge), as in the reference's two sign cases.  The fraction is
  // |x - k1 d| / d: the reference's fraction, negated where it interpolates
  // towards the negative side (the sign flips are exact).
  const int k1x = min(max((int)(xx * iax), -2), 2);
  const int k1z = min(max((int)(yy * iaz), -2), 2);
  const int k2x = min(max(k1x + (xx >= 0 ? 1 : -1), -2), 2);
  const int k2z = min(max(k1z + (yy >= 0 ? 1 : -1), -2), 2);
  const double2 *tc = t + 12;
  const double2 e1 = tc[5 * k1x + k1z], e2 = tc[5 * k2x + k1z], e3 = tc[5 * k1x + k2z], e4 = tc[5 * k2x + k2z];
  const double sx = fabs((xx - k1x * aAx) * iax), sz = fabs((yy - k1z * aAz) * iaz);
  const double y5 = e1.y + sx * (e2.y - e1.y);
  const double y6 = e3.y + sx * (e4.y - e3.y);
  const double x5 = e1.x + sx * (e2.x - e1.x);
  const double x6 = e3.x + sx * (e4.x - e3.x);
  Ey = y5 + sz * (y6 - y5);
  Ex = x5 + sz * (x6 - x5);
}

// ------------------------------------------------------------------- phase 1

// Per bunch (blocks 0 .. G-1): the synchrotron step of every slice, the
// centroids it reports, the manager's history and the monitor statistics.
// With FBII every thread of the grid (G S 25 / WS_WG blocks, at least G) also
// evaluates the Bassetti-Erskine field at one absolute grid point of one slice
// of one bunch; the field needs the slices' transverse positions only, which
// the synchrotron step leaves alone.
__global__ __launch_bounds__(WS_WG) void k_ws_phase1(WSP P, WSD D, long long rev, int slot, int monitor)
{
  const int tid = threadIdx.x, S = P.S;
  if (blockIdx.x < P.G) {
    const int b = blockIdx.x;
    const int kb = D.bucket[b];
    double *xtau = D.sl + (size_t)b * 6 * S, *xeps = xtau + S, *xx = xeps + S, *xp = xx + S, *zz = xp + S,
           *zp = zz + S;
    for (int ic = tid; ic < S; ic += WS_WG) {
      const double e0 = xeps[ic];
      const double gain = P.eps * sin(P.wrf * xtau[ic] + P.phai0);
      xeps[ic] = e0 + gain - P.Urad;
      xtau[ic] = xtau[ic] - e0 * P.T0 * P.ac;
    }
    __syncthreads();
    if (tid == 0) {
      double pos[3], slope[3];
      ws_cm(xtau, xeps, D.npop, S, pos[0], slope[0]);
      ws_cm(xx, xp, D.npop, S, pos[1], slope[1]);
      ws_cm(zz, zp, D.npop, S, pos[2], slope[2]);
      double *o = D.cm + ((size_t)slot * P.G + b) * 4;
      o[0] = pos[1];
      o[1] = slope[1];
      o[2] = pos[2];
      o[3] = slope[2];
      // the manager's history of this bucket: shift by one, newest at 0
      const double nw[4] = {pos[1], slope[1], pos[2], slope[2]};
      for (int a = 0; a < 4; a++) {
        double *hrow = D.hist + ((size_t)a * P.h + kb) * WS_HIST;
        for (int i = 0; i < P.NmultiT; i++) {
          const int j = P.NmultiT - i - 2;
          if (j >= 0) hrow[j + 1] = hrow[j];
          else hrow[0] = nw[a];
        }
      }
      if (monitor) {
        double *m = D.mon + ((size_t)slot * P.G + b) * 18;
        const double *U[3] = {xtau, xx, zz}, *Q[3] = {xeps, xp, zp};
        for (int pl = 0; pl < 3; pl++) {
          for (int k = 0; k < 6; k++) m[3 * k + pl] = 0.0;
          if (!P.track[pl]) continue;
          const double *u = U[pl], *p = Q[pl];
          double avr, avs;
          ws_cm(u, p, D.npop, S, avr, avs);
          double sq = 0.0, sqs = 0.0, amp = 0.0;
          int w1 = 0, w2 = 0, n = 0;
          for (int ic = 0; ic < S; ic++) {
            sq += (u[ic] - avr) * (u[ic] - avr) * D.npop[ic];
            w1 += D.npop[ic];
          }
          for (int ic = 0; ic < S; ic++) {
            sqs += (p[ic] - avs) * (p[ic] - avs) * D.npop[ic];
            w2 += D.npop[ic];
          }
          for (int ic = 0; ic < S; ic++) {
            amp += D.npop[ic] * (P.gamma1[pl] * (u[ic] * u[ic]) + 2.0 * P.alpha1[pl] * u[ic] * p[ic] +
                                 P.beta1[pl] * (p[ic] * p[ic]));
            n += D.npop[ic];
          }
          m[pl] = avr;
          m[3 + pl] = sqrt(sq / ((double)w1));
          m[6 + pl] = avs;
          m[9 + pl] = sqrt(sqs / ((double)w2));
          m[12 + pl] = amp / ((double)n);
          m[15 + pl] = P.gamma1[pl] * (avr * avr) + 2.0 * P.alpha1[pl] * avr * avs + P.beta1[pl] * (avs * avs);
        }
      }
    }
  }
  if (!P.fbii) return;
  // ---- Bassetti-Erskine field of one slice at one absolute grid point
  const int npts = S * 25;
  const long long gq = (long long)blockIdx.x * WS_WG + tid;
  if (gq >= (long long)P.G * npts) return;
  const int b = (int)(gq / npts), q = (int)(gq - (long long)b * npts);
  const double *xx = D.sl + (size_t)b * 6 * S + 2 * S, *zz = xx + 2 * S;
  const int cell = q / 25, r = q % 25, msx = r / 5 - 2, msz = r % 5 - 2;
  const double dQ = D.nelec[b] * D.npop[cell] * K_E;
  const double fdQ = dQ / P.den_dQ;
  const double Dx = msx * P.aAx - xx[cell];
  const double xi1 = Dx / P.factBE, xi2 = xi1 * P.szz / P.sxx;
  const double Dy = msz * P.aAz - zz[cell];
  const double yi1 = Dy / P.factBE, yi2 = yi1 * P.sxx / P.szz;
  double u1, v1, u2, v2;
  if (P.mixed) {
    float a1, b1, a2, b2;
    mbt_wofz_dev<float>((float)xi1, (float)yi1, a1, b1);
    mbt_wofz_dev<float>((float)xi2, (float)yi2, a2, b2);
    u1 = a1;
    v1 = b1;
    u2 = a2;
    v2 = b2;
  } else {
    mbt_wofz_dev<double>(xi1, yi1, u1, v1);
    mbt_wofz_dev<double>(xi2, yi2, u2, v2);
  }
  const double g = exp(-(Dx * Dx * 1.0e+06 / (2 * P.sxx * P.sxx) + Dy * Dy * 1.0e+06 / (2 * P.szz * P.szz)));
  const double Ex = -fdQ * (v1 - g * v2), Ey = -fdQ * (u1 - g * u2);
  D.Exy[gq] = make_double2(Ex, Ey);
}

// The ions every bunch creates, one per thread: one per local grid point of
// each slice, kicked by the bunch's later slices (their field tables read from
// the one k_ws_phase1 wrote).
__global__ __launch_bounds__(WS_WG) void k_ws_create(WSP P, WSD D)
{
  const int S = P.S, npts = S * 25;
  const long long gq = (long long)blockIdx.x * WS_WG + threadIdx.x;
  if (gq >= (long long)P.G * npts) return;
  const int b = (int)(gq / npts), q = (int)(gq - (long long)b * npts);
  const double *xx = D.sl + (size_t)b * 6 * S + 2 * S, *zz = xx + 2 * S;
  const double2 *exy = D.Exy + (size_t)b * npts;
  const int cell = q / 25, r = q % 25, msx = r / 5 - 2, msz = r % 5 - 2;
  const double Nri = D.convion[b] * ((double)D.nofe[q]);
  double x = xx[cell] + msx * P.aLx, xd = 0.0, y = zz[cell] + msz * P.aLz, yd = 0.0;
  for (int c = cell + 1; c < S; c++) {
    double Ex, Ey;
    ws_interp(exy + c * 25, P.aAx, P.aAz, P.iaAx, P.iaAz, x, y, Ex, Ey);
    const double ax = P.factF * Ex;
    xd = xd + P.deltT * ax;
    x = x + P.deltT * xd;
    const double ay = P.factF * Ey;
    yd = yd + P.deltT * ay;
    y = y + P.deltT * yd;
  }
  const size_t GP = (size_t)P.G * npts;
  D.li[gq] = Nri;
  D.li[GP + gq] = x;
  D.li[2 * GP + gq] = xd;
  D.li[3 * GP + gq] = y;
  D.li[4 * GP + gq] = yd;
}

// ------------------------------------------------------------------- phase 2

__global__ __launch_bounds__(WS_WG) void k_ws_phase2(WSP P, WSD D, long long rev)
{
  extern __shared__ double dup[]; // [2][h]
  const int b = blockIdx.x, tid = threadIdx.x, S = P.S, h = P.h;
  const int kb = D.bucket[b];
  double *xx = D.sl + (size_t)b * 6 * S + 2 * S, *xp = xx + S, *zz = xp + S, *zp = zz + S;
  for (int ic = tid; ic < S; ic += WS_WG) {
    const double y = P.v11 * zz[ic] + P.v12 * zp[ic];
    const double yp = P.v21 * zz[ic] + P.v22 * zp[ic];
    zz[ic] = y;
    zp[ic] = yp;
    const double x = P.h11 * xx[ic] + P.h12 * xp[ic];
    const double xq = P.h21 * xx[ic] + P.h22 * xp[ic];
    xx[ic] = x;
    xp[ic] = xq;
  }
  if (!P.rw) return;
  // the kick of every filled source bucket jb, one lane each; the worker's copy
  // of the history holds zeros where jb + j h falls outside the 5000-double message
  for (int pl = 0; pl < 2; pl++) {
    if (!(pl == 0 ? P.track_h : P.track_v)) continue;
    const double *uh = D.hist + (size_t)(pl == 0 ? 0 : 2) * h * WS_HIST;
    const double *ph = uh + (size_t)h * WS_HIST;
    const double prf = (P.triggerRW && rev < 50) ? P.prfrw_trig[pl] : P.prfrw[pl];
    for (int jb = tid; jb < h; jb += WS_WG) {
      double Dup = 0.0;
      if (D.fill[jb]) {
        double D0;
        if (jb == kb) D0 = P.Lc;
        else if (jb > kb) D0 = ((double)(jb - kb)) * P.Dub0;
        else D0 = ((double)(h + jb - kb)) * P.Dub0;
        for (int j = 0; j < P.Nmlt + 1; j++) {
          const double Du = D0 + ((double)j) * P.Lc;
          const double Dt = Du / K_C;
          const double W = prf / sqrt(Dt);
          const int jm = jb > kb ? j : j + 1;
          const bool in = jb + jm * h < WS_MSG;
          const double u = in ? uh[(size_t)jb * WS_HIST + jm] : 0.0, p = in ? ph[(size_t)jb * WS_HIST + jm] : 0.0;
          const double ujb = P.rwa11[pl] * u + P.rwa12[pl] * p;
          const double dip = P.fNp * ujb;
          Dup = Dup + dip * W;
        }
        Dup = P.Pref0 * D.Ibunch[jb] * Dup;
        if (rev < P.Nmlt) Dup = 0.0;
      }
      dup[pl * h + jb] = Dup;
    }
  }
  __syncthreads();
  for (int pl = 0; pl < 2; pl++) {
    if (!(pl == 0 ? P.track_h : P.track_v)) continue;
    double *up = pl == 0 ? xp : zp;
    for (int ic = tid; ic < S; ic += WS_WG) {
      double a = up[ic];
      for (int jb = 0; jb < h; jb++)
        if (D.fill[jb]) a += dup[pl * h + jb];
      up[ic] = a;
    }
  }
}

// ------------------------------------------------------------------ the train

// head bunch: the cloud becomes its ions (src/fbi_weakstrong.c:335-353)
__global__ __launch_bounds__(WS_WG) void k_ws_cloud_reset(WSP P, WSD D, int slot)
{
  __shared__ double red[WS_WG / 64][3];
  const int tid = threadIdx.x, npts = P.S * 25;
  const size_t GP = (size_t)P.G * npts;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int q = tid; q < npts; q += WS_WG) {
    for (int a = 0; a < 5; a++) D.cl[a * D.cap + q] = D.li[a * GP + q];
    s0 += D.li[q] * D.li[GP + q];
    s1 += D.li[q] * D.li[3 * GP + q];
    s2 += D.li[q];
  }
  s0 = wsum64(s0);
  s1 = wsum64(s1);
  s2 = wsum64(s2);
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = s0;
    red[tid >> 6][1] = s1;
    red[tid >> 6][2] = s2;
  }
  __syncthreads();
  if (tid < 3) {
    double s = 0.0;
    for (int w = 0; w < WS_WG / 64; w++) s += red[w][tid];
    D.ion[((size_t)slot * P.G) * 3 + tid] = s;
  }
}

// bunch ib (bucket kb) meets the cloud of nF ions that drifted Dtime since the
// previous bunch; the last block applies the slice kicks and appends the
// bunch's ions
__global__ __launch_bounds__(WS_WG) void k_ws_train(WSP P, WSD D, int ib, long long nF, double Dtime, int slot)
{
  __shared__ double2 exy_s[WS_PTS];
  __shared__ double red[WS_WG / 64][2];
  __shared__ int last;
  const int tid = threadIdx.x, S = P.S, npts = S * 25;
  const size_t GP = (size_t)P.G * npts;
  for (int q = tid; q < npts; q += WS_WG) exy_s[q] = D.Exy[(size_t)ib * npts + q];
  __syncthreads();
  const long long q = (long long)blockIdx.x * WS_WG + tid;
  const bool act = q < nF;
  double Nri = 0.0, x = 0.0, xd = 0.0, y = 0.0, yd = 0.0;
  if (act) {
    Nri = D.cl[q];
    x = D.cl[D.cap + q];
    xd = D.cl[2 * D.cap + q];
    y = D.cl[3 * D.cap + q];
    yd = D.cl[4 * D.cap + q];
    x = x + Dtime * xd;
    y = y + Dtime * yd;
  }
  double *part = D.part + (size_t)blockIdx.x * S * 2;
  for (int c = 0; c < S; c++) {
    double sx = 0.0, sy = 0.0;
    if (act) {
      double Ex, Ey;
      ws_interp(exy_s + c * 25, P.aAx, P.aAz, P.iaAx, P.iaAz, x, y, Ex, Ey);
      const double ax = P.factF * Ex;
      xd = xd + P.deltT * ax;
      x = x + P.deltT * xd;
      const double ay = P.factF * Ey;
      yd = yd + P.deltT * ay;
      y = y + P.deltT * yd;
      sx = Nri * Ex;
      sy = Nri * Ey;
    }
    sx = wsum64(sx);
    sy = wsum64(sy);
    if ((tid & 63) == 0) {
      red[tid >> 6][0] = sx;
      red[tid >> 6][1] = sy;
    }
    __syncthreads();
    if (tid < 2) {
      double s = 0.0;
      for (int w = 0; w < WS_WG / 64; w++) s += red[w][tid];
      part[2 * c + tid] = s;
    }
    __syncthreads();
  }
  double s0 = act ? Nri * x : 0.0, s1 = act ? Nri * y : 0.0, s2 = act ? Nri : 0.0;
  if (act) {
    D.cl[D.cap + q] = x;
    D.cl[2 * D.cap + q] = xd;
    D.cl[3 * D.cap + q] = y;
    D.cl[4 * D.cap + q] = yd;
  }
  s0 = wsum64(s0);
  s1 = wsum64(s1);
  s2 = wsum64(s2);
  __shared__ double red3[WS_WG / 64][3];
  if ((tid & 63) == 0) {
    red3[tid >> 6][0] = s0;
    red3[tid >> 6][1] = s1;
    red3[tid >> 6][2] = s2;
  }
  __syncthreads();
  if (tid < 3) {
    double s = 0.0;
    for (int w = 0; w < WS_WG / 64; w++) s += red3[w][tid];
    D.psum[(size_t)blockIdx.x * 3 + tid] = s;
  }
  // ---- the last block to finish: fixed-order sums, slice kicks, append
  __threadfence();
  __syncthreads();
  if (tid == 0) last = atomicAdd(D.done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  __threadfence();
  const int kb = D.bucket[ib];
  const double nel = D.nelec[ib];
  double *xp = D.sl + (size_t)ib * 6 * S + 3 * S, *zp = xp + 2 * S;
  for (int c = tid; c < S; c += WS_WG) {
    double sx = 0.0, sy = 0.0;
    for (unsigned k = 0; k < gridDim.x; k++) {
      sx += D.part[((size_t)k * S + c) * 2];
      sy += D.part[((size_t)k * S + c) * 2 + 1];
    }
    const double dQ = nel * D.npop[c] * K_E;
    const double fjp = dQ > 0.0 ? P.factjp0 / dQ : 0.0;
    sx = fjp * sx;
    sy = fjp * sy;
    xp[c] = xp[c] + P.CdeltT * sx;
    zp[c] = zp[c] + P.CdeltT * sy;
  }
  // append the bunch's ions and finish the cloud sums
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  for (int p = tid; p < npts; p += WS_WG) {
    const size_t o = (size_t)ib * npts + p;
    for (int a = 0; a < 5; a++) D.cl[a * D.cap + nF + p] = D.li[a * GP + o];
    t0 += D.li[o] * D.li[GP + o];
    t1 += D.li[o] * D.li[3 * GP + o];
    t2 += D.li[o];
  }
  t0 = wsum64(t0);
  t1 = wsum64(t1);
  t2 = wsum64(t2);
  if ((tid & 63) == 0) {
    red3[tid >> 6][0] = t0;
    red3[tid >> 6][1] = t1;
    red3[tid >> 6][2] = t2;
  }
  __syncthreads();
  if (tid < 3) {
    double s = 0.0;
    for (unsigned k = 0; k < gridDim.x; k++) s += D.psum[(size_t)k * 3 + tid];
    for (int w = 0; w < WS_WG / 64; w++) s += red3[w][tid];
    D.ion[((size_t)slot * P.G + ib) * 3 + tid] = s;
  }
  if (tid == 0) {
    *D.done = 0;
    (void)kb;
  }
}

// The ion train's field table of a slice in LDS, over the 5 x 5 grid with its
// edge rows and columns repeated once more: entry 7 kx + kz (kx in 0 .. 5,
// kz in 0 .. 6) of the first WS_QT double2 holds the field (Ex, Ey) at the
// point (clamp(kx - 3), clamp(kz - 3)) of the grid, the same entry of the next
// WS_QT its step to the next point in x, E(kx + 1, kz) - E(kx, kz) (zero past
// the edge).  The steps are the subtractions of the bilinear formula, made once
// per table instead of once per ion.  Two arrays of 16-byte entries rather than
// 32-byte (field, step) pairs: the lanes of one read then spread over all
// sixteen 16-byte bank groups of the LDS, not eight (measured: a sixth of the
// bank-conflict cycles, SQ_LDS_BANK_CONFLICT).
#define WS_QT 42 // entries per slice
#define WS_CH 16 // slices per staged chunk of a bunch's tables (a multiple of 4)

// ws_interp on a slice's table, with the cell anchored at the floor of the
// grid quotient u: the cell [f, f + 1], f = floor(u), at weight u - f.  The
// reference anchors at the grid point nearer zero (trunc(u)) and weighs toward
// the next one out; for u >= 0 that is the same cell and steps, for u < 0 the
// same bilinear interpolant taken from its other end.  The floor and the
// weight come from t = u + 3 in one rounding (a fused multiply-add), as its
// integer part and v_fract_f64: the weight then carries the rounding of t (at
// most 4.4e-16 for |u| < 3) instead of that of u.  Quotients past the grid take
// the cell [-3, -2] or [2, 3], whose outer point repeats the edge (a zero step),
// so the field there is the edge value, as the reference's clamped partner
// gives.  No sign tests; the four reads are fixed offsets from one address.
__device__ __forceinline__ void ws_interp_p(const double2 *t, double iax, double iaz, double xx, double yy,
                                            double &Ex, double &Ey)
{
  // the bilinear steps as fused multiply-adds (one rounding less each)
#pragma clang fp contract(fast)
  const double tx = fma(xx, iax, 3.0), tz = fma(yy, iaz, 3.0);
  const double sx = __builtin_amdgcn_fract(tx), sz = __builtin_amdgcn_fract(tz);
  const int kx = min(max((int)tx, 0), 5), kz = min(max((int)tz, 0), 5);
  const double2 *e = t + 7 * kx + kz;
  const double2 e1 = e[0], e3 = e[1], d1 = e[WS_QT], d3 = e[WS_QT + 1];
  const double y5 = e1.y + sx * d1.y;
  const double y6 = e3.y + sx * d3.y;
  const double x5 = e1.x + sx * d1.x;
  const double x6 = e3.x + sx * d3.x;
  Ey = y5 + sz * (y6 - y5);
  Ex = x5 + sz * (x6 - x5);
}

// One slice c of a bunch for the lane's ion: the field at the ion, the ion's
// kick, and the ion's share Nri E of the slice's kick.  The velocity kick
// deltT (factF E) of the reference is taken as (deltT factF) E, one fused
// multiply-add with the product precomputed (a rounding-level change).
__device__ __forceinline__ void ion_slice(const WSP &P, const double2 *qt, int c, double Nri, double &x, double &xd,
                                          double &y, double &yd, double &sx, double &sy)
{
  // the ion's kick as fused multiply-adds
#pragma clang fp contract(fast)
  double Ex, Ey;
  ws_interp_p(qt + c * 2 * WS_QT, P.iaAx, P.iaAz, x, y, Ex, Ey);
  xd = xd + P.dtF * Ex;
  x = x + P.deltT * xd;
  yd = yd + P.dtF * Ey;
  y = y + P.deltT * yd;
  sx = Nri * Ex;
  sy = Nri * Ey;
}

// One group of four slices c0 .. c0 + 3 of a bunch for the lane's ion: the
// kicks, then the group's eight field sums (Nri Ex, Nri Ey per slice) reduced
// over the wave into wred.  qt holds the tables of the slices from cq on.
// ALL: every lane's ion is in the cloud (lanes past the last ion carry Nri = 0
// from a finite state, so they add exact zeros); FULL: all four slices exist.
template <bool ALL, bool FULL>
__device__ __forceinline__ void ions_group(const WSP &P, const double2 *qt, int cq, int c0, int S, bool act,
                                           double Nri, double &x, double &xd, double &y, double &yd,
                                           double (*wred)[WS_WG / 64], int lane, int wv)
{
  double v[8];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int c = c0 + k;
    double sx = 0.0, sy = 0.0;
    if ((ALL || act) && (FULL || c < S)) ion_slice(P, qt, c - cq, Nri, x, xd, y, yd, sx, sy);
    v[2 * k] = sx;
    v[2 * k + 1] = sy;
  }
  const double t = wsum64x8(v);
  const int idx = lane >> 3, c = c0 + (idx >> 1);
  if ((lane & 7) == 0 && (FULL || c < S)) wred[2 * c + (idx & 1)][wv] = t;
}

// The last r = 1 or 2 slices c0 .. c0 + r - 1 of a bunch and, after them, the
// cloud sums Nri x, Nri y, Nri: one eight-value reduction for both.
template <bool ALL>
__device__ __forceinline__ void ions_tail(const WSP &P, const double2 *qt, int cq, int c0, int r, int S, bool act,
                                          double Nri, double &x, double &xd, double &y, double &yd,
                                          double (*wred)[WS_WG / 64], int lane, int wv)
{
  double v[8];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    double sx = 0.0, sy = 0.0;
    if ((ALL || act) && k < r) ion_slice(P, qt, c0 + k - cq, Nri, x, xd, y, yd, sx, sy);
    v[2 * k] = sx;
    v[2 * k + 1] = sy;
  }
  v[4] = ALL || act ? Nri * x : 0.0;
  v[5] = ALL || act ? Nri * y : 0.0;
  v[6] = ALL || act ? Nri : 0.0;
  v[7] = 0.0;
  const double t = wsum64x8(v);
  const int m = lane >> 3;
  if ((lane & 7) == 0) {
    if (m < 2 * r)
      wred[2 * c0 + m][wv] = t;
    else if (m >= 4 && m < 7)
      wred[2 * S + m - 4][wv] = t;
  }
}

// Slices cb .. ce - 1 of a bunch for the lane's ion (qt: their tables from cb
// on; cb a multiple of 4) and, when they end the bunch (ce = S), the cloud
// sums Nri x, Nri y, Nri after it, over the wave into wred.
template <bool ALL>
__device__ __forceinline__ void ions_chunk(const WSP &P, const double2 *qt, int cb, int ce, bool act, double Nri,
                                           double &x, double &xd, double &y, double &yd,
                                           double (*wred)[WS_WG / 64], int lane, int wv)
{
  const int S = P.S;
  int c0 = cb;
  for (; c0 + 4 <= ce; c0 += 4) ions_group<ALL, true>(P, qt, cb, c0, S, act, Nri, x, xd, y, yd, wred, lane, wv);
  if (ce < S) return;
  const int r = S - c0;
  if (r == 1 || r == 2) {
    ions_tail<ALL>(P, qt, cb, c0, r, S, act, Nri, x, xd, y, yd, wred, lane, wv);
    return;
  }
  if (r) ions_group<ALL, false>(P, qt, cb, c0, S, act, Nri, x, xd, y, yd, wred, lane, wv);
  double s0 = ALL || act ? Nri * x : 0.0, s1 = ALL || act ? Nri * y : 0.0, s2 = ALL || act ? Nri : 0.0;
  s0 = wsum64(s0);
  s1 = wsum64(s1);
  s2 = wsum64(s2);
  if (lane == 0) {
    wred[2 * S][wv] = s0;
    wred[2 * S + 1][wv] = s1;
    wred[2 * S + 2][wv] = s2;
  }
}

// Ion-parallel train: lane q carries ion q = j * Nions + p (ion p made by bunch
// j, its state after the creator's later slices) through bunches j+1 .. G-1,
// each reached after its drift time.  Per bunch ib the block writes the 2 S
// slice sums of Nri * E and the cloud sums Nri x, Nri y, Nri after the bunch,
// each the fixed-order sum of its four waves' sums.  The bunches up to the
// block's last creator (some lanes' ions not yet made) run with per-lane
// predicates, the rest without.  The tables of a bunch are staged in chunks
// of WS_CH slices (dynamic LDS: min(S, WS_CH) * 2 WS_QT double2, 21.5 KB at
// most), so that six blocks fit a CU: 80 registers a lane allow six waves per
// SIMD, and the 1,422 blocks of the 356-bunch train are all resident at once.
__global__ __launch_bounds__(WS_WG) void k_ws_ions(WSP P, WSD D)
{
  extern __shared__ double2 qt_s[];
  __shared__ double wred[2 * WS_S + 3][WS_WG / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, S = P.S, npts = S * 25, NQ = 2 * S + 3;
  const size_t GP = (size_t)P.G * npts;
  const long long nI = (long long)(P.G - 1) * npts;
  // Dispatch order: a block's work is the number of bunches after its ions'
  // creator, longest first.  Blocks are dealt to the CUs round-robin, so
  // blocks c, c + 256, c + 512, ... share a CU: in each run of 512 blocks
  // (or the last, shorter run, if it holds more than 256) block c takes the
  // logical block 255 - c and block c + 256 the logical block 256 + c, so
  // that every CU holds pairs of the same total work instead of the longest
  // trains on the first CUs (measured 4 % on the 356-bunch train).  Every
  // logical block computes and writes what it would at its own index.
  int L = blockIdx.x;
  {
    const int c = L & 255, k = L >> 8, base = (k >> 1) * 512;
    if ((int)gridDim.x - base > 256) L = (k & 1) ? base + 256 + c : base + 255 - c;
  }
  const long long q0 = (long long)L * WS_WG, q = q0 + tid;
  const int jmin = (int)(q0 / npts);
  const int jmax = (int)(min(q0 + WS_WG - 1, nI - 1) / npts);
  const bool valid = q < nI;
  const int j = valid ? (int)(q / npts) : P.G;
  double Nri = 0.0, x = 0.0, xd = 0.0, y = 0.0, yd = 0.0;
  if (valid) {
    Nri = D.li[q];
    x = D.li[GP + q];
    xd = D.li[2 * GP + q];
    y = D.li[3 * GP + q];
    yd = D.li[4 * GP + q];
  }
  // threads 0 .. 6 WS_QT - 1 stage entry r of the tables of slices sl0,
  // sl0 + 6, ... of a chunk: a fixed grid point and its x neighbour per thread
  const int sl0 = tid / WS_QT, r = tid % WS_QT, gz = min(max(r % 7 - 3, -2), 2);
  const int src = 12 + 5 * min(max(r / 7 - 3, -2), 2) + gz, src2 = 12 + 5 * min(max(r / 7 - 2, -2), 2) + gz;
  for (int ib = jmin + 1; ib < P.G; ib++) {
    const bool act = ib > jmax || j < ib;
    if (act) {
      const double dt = D.dtime[ib];
      x = x + dt * xd;
      y = y + dt * yd;
    }
    const double2 *exy = D.Exy + (size_t)ib * npts;
    for (int cb = 0; cb < S; cb += WS_CH) {
      const int ce = min(S, cb + WS_CH);
      __syncthreads(); // the previous chunk's tables (and bunch's sums) have been read
      if (tid < 6 * WS_QT)
        for (int c = cb + sl0; c < ce; c += 6) {
          const double2 a = exy[c * 25 + src], b = exy[c * 25 + src2];
          qt_s[(c - cb) * 2 * WS_QT + r] = a;
          qt_s[(c - cb) * 2 * WS_QT + WS_QT + r] = make_double2(b.x - a.x, b.y - a.y);
        }
      __syncthreads();
      if (ib <= jmax)
        ions_chunk<false>(P, qt_s, cb, ce, act, Nri, x, xd, y, yd, wred, lane, wv);
      else
        ions_chunk<true>(P, qt_s, cb, ce, true, Nri, x, xd, y, yd, wred, lane, wv);
    }
    __syncthreads();
    // quantity-major, so that k_ws_ksum reads each quantity's partials contiguously
    const long long nbi = min(((long long)ib * npts + WS_WG - 1) / WS_WG, (long long)D.nblk_ions);
    for (int i = tid; i < NQ; i += WS_WG) {
      double v = 0.0;
      for (int w = 0; w < WS_WG / 64; w++) v += wred[i][w];
      D.tpart[D.toff[ib] + i * nbi + L] = v;
    }
  }
}

// Per bunch ib (blockIdx.x) and group of KSUM_Q quantities (blockIdx.y): the
// blocks' partials summed in a fixed order, one quantity per wave (lane l adds
// blocks l, l + 64, ... in turn, then the wave sum), into D.tot[ib][i].  The
// last group of each bunch sums instead the bunch's own ions (Nri x, Nri y,
// Nri over its S * 25 ions, four waves' sums in order) into D.tot[ib][NQ + 3 w
// + m], which k_ws_end adds to the cloud sums after the bunch.
#define KSUM_Q 8
__global__ __launch_bounds__(64 * KSUM_Q) void k_ws_ksum(WSP P, WSD D)
{
  const int ib = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, S = P.S, npts = S * 25;
  const int NQ = 2 * S + 3;
  double *tot = D.tot + (size_t)ib * WS_TQ;
  if (blockIdx.y == gridDim.y - 1) {
    if (tid >= WS_WG) return;
    const size_t GP = (size_t)P.G * npts;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int p = tid; p < npts; p += WS_WG) {
      const size_t o = (size_t)ib * npts + p;
      t0 += D.li[o] * D.li[GP + o];
      t1 += D.li[o] * D.li[3 * GP + o];
      t2 += D.li[o];
    }
    t0 = wsum64(t0);
    t1 = wsum64(t1);
    t2 = wsum64(t2);
    if (lane == 0) {
      tot[NQ + 3 * wv] = t0;
      tot[NQ + 3 * wv + 1] = t1;
      tot[NQ + 3 * wv + 2] = t2;
    }
    return;
  }
  const int i = blockIdx.y * KSUM_Q + wv;
  if (ib == 0 || i >= NQ) return;
  long long nb = ((long long)ib * npts + WS_WG - 1) / WS_WG; // blocks of k_ws_ions ahead of ib
  if (nb > D.nblk_ions) nb = D.nblk_ions;
  const double *tp = D.tpart + D.toff[ib] + i * nb;
  double v = 0.0;
#pragma unroll 4
  for (long long k = lane; k < nb; k += 64) v += tp[k];
  v = wsum64(v);
  if (lane == 0) tot[i] = v;
}

// Blocks 0 .. nkick-1 (the ion-parallel train, nkick = G): bunch b's slice
// kicks from the ion field sums and the cloud sums after it with its own ions
// appended (the sums of k_ws_ksum).  The blocks after them: the slice-averaged
// vertical centroid of every bucket (writeoutVCMmotionsPM2,
// src/fbi_weakstrong.c:688-737), which the kicks (on the slopes) leave alone.
__global__ void k_ws_end(WSP P, WSD D, const int *bunch_of_bucket, int slot, int nkick)
{
  const int tid = threadIdx.x, S = P.S;
  if ((int)blockIdx.x < nkick) {
    const int ib = blockIdx.x, NQ = 2 * S + 3;
    const double *tot = D.tot + (size_t)ib * WS_TQ;
    if (ib > 0) {
      const double nel = D.nelec[ib];
      double *xp = D.sl + (size_t)ib * 6 * S + 3 * S, *zp = xp + 2 * S;
      for (int c = tid; c < S; c += blockDim.x) {
        const double dQ = nel * D.npop[c] * K_E;
        const double fjp = dQ > 0.0 ? P.factjp0 / dQ : 0.0;
        const double sx = fjp * tot[2 * c], sy = fjp * tot[2 * c + 1];
        xp[c] = xp[c] + P.CdeltT * sx;
        zp[c] = zp[c] + P.CdeltT * sy;
      }
    }
    if (tid < 3) {
      double v = ib > 0 ? tot[2 * S + tid] : 0.0;
      for (int w = 0; w < WS_WG / 64; w++) v += tot[NQ + 3 * w + tid];
      D.ion[((size_t)slot * P.G + ib) * 3 + tid] = v;
    }
    return;
  }
  const int kb = (blockIdx.x - nkick) * blockDim.x + tid;
  if (kb >= P.h) return;
  double z = 0.0;
  const int b = bunch_of_bucket[kb];
  if (b >= 0) {
    const double *zz = D.sl + (size_t)b * 6 * P.S + 4 * P.S;
    for (int ic = 0; ic < P.S; ic++) z = z + zz[ic];
    z = z / ((double)P.S);
  }
  D.zcm[(size_t)slot * P.h + kb] = z;
}

// ================================================================ host side

struct mbt_ws {
  mbt_config cfg;
  int device;
  hipStream_t stream;
  WSP P;
  WSD D;
  int *d_bucket, *d_fill, *d_npop, *d_nofe, *d_bob;
  double *d_Ibunch, *d_nelec, *d_convion;
  int *bucket_h, *fill_h;
  int npop_h[WS_S];
  double *dist; // generated macro-electrons [6][Np]
  double *nelec_h;
  long long rev;
  int maxblk;
  int serial;          // bunch-by-bunch relay instead of the ion-parallel train
  double *d_dtime;
  long long *d_toff;
  int timing;         // HIP events around the ion train of every turn
  hipEvent_t *ev;     // [2 WS_T] train start / end per turn of a block
  double train_ms;    // summed train time of the last mbt_ws_track call
  long train_launches;
  long async_n;       // turns enqueued by mbt_ws_track_async, not yet collected (0: none)
};

// Tracking calls that would overwrite the output block of turns enqueued by
// mbt_ws_track_async fail while those turns are pending.
static int ws_no_pending(const mbt_ws *w)
{
  if (w->async_n) {
    mbt_set_error("%ld turns enqueued by mbt_ws_track_async are pending: collect them with mbt_ws_track_wait",
                  w->async_n);
    return MBT_EINVAL;
  }
  return MBT_OK;
}

extern "C" void mbt_ws_destroy(mbt_ws *w)
{
  if (!w) return;
  (void)hipSetDevice(w->device);
  if (w->stream) (void)hipStreamSynchronize(w->stream); // turns of mbt_ws_track_async may still run
  void *p[] = {w->D.sl,  w->D.hist, w->D.Exy, w->D.li,     w->D.cl,     w->D.part,    w->D.psum,
               w->D.done, w->D.cm,  w->D.mon,  w->D.ion, w->D.zcm,    w->d_bucket, w->d_fill,    w->d_npop,
               w->d_nofe, w->d_bob, w->d_Ibunch, w->d_nelec, w->d_convion, w->D.tpart, w->d_dtime, w->d_toff,
               w->D.tot};
  for (size_t i = 0; i < sizeof p / sizeof p[0]; i++)
    if (p[i]) (void)hipFree(p[i]);
  if (w->ev)
    for (int i = 0; i < 2 * WS_T; i++) (void)hipEventDestroy(w->ev[i]);
  free(w->ev);
  if (w->stream) (void)hipStreamDestroy(w->stream);
  free(w->bucket_h);
  free(w->fill_h);
  free(w->dist);
  free(w->nelec_h);
  free(w);
}

template <typename T> static int walloc(T **p, size_t n)
{
  if (n == 0) n = 1;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
    mbt_set_error("hipMalloc of %zu bytes failed", n * sizeof(T));
    return MBT_ENOMEM;
  }
  return hipMemset(*p, 0, n * sizeof(T)) == hipSuccess ? MBT_OK : MBT_EDEVICE;
}

#define WA(p, n)                                                                                 \
  do {                                                                                           \
    int rc_ = walloc(&(p), (size_t)(n));                                                         \
    if (rc_) {                                                                                   \
      mbt_ws_destroy(w);                                                                         \
      return rc_;                                                                                \
    }                                                                                            \
  } while (0)
#define WUP(dst, src, n) CK(hipMemcpyAsync((dst), (src), (size_t)(n) * sizeof(*(src)), hipMemcpyHostToDevice, w->stream))

extern "C" int mbt_ws_create(const mbt_config *c, int device, int flags, mbt_ws **out)
{
  *out = NULL;
  if (c->bunch_model != 1) {
    mbt_set_error("mbt_ws_create needs BunchModel = STRONG");
    return MBT_EINVAL;
  }
  const int S = 2 * c->Nslc + 1, H = (int)c->Nharm, G = c->Nbunch;
  if (c->Nslc < 0 || S > WS_S || H > 500 || G <= 0 || c->NmultiT > WS_HIST || c->Nmlt + 2 > WS_HIST ||
      c->NrevMon <= 0) {
    mbt_set_error("weak-strong limits: 2 Nslc + 1 <= 50, h <= 500, Nmlt + 2 <= 30, NrevMon > 0");
    return MBT_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
    mbt_set_error("no HIP device %d", device);
    return MBT_EDEVICE;
  }
  mbt_ws *w = (mbt_ws *)calloc(1, sizeof *w);
  w->cfg = *c;
  w->cfg.nfFill = NULL;
  w->cfg.Ibunch = NULL;
  w->cfg.bucket = NULL;
  w->cfg.bunch_Ib = NULL;
  w->cfg.Ib_frac = NULL;
  w->cfg.bunch_out = NULL;
  w->device = device;
  CK(hipSetDevice(device));
  CK(hipStreamCreate(&w->stream));
  // ---- slices, populations, grid counts (host, the reference's generator)
  double *slices = (double *)malloc(sizeof(double) * 6 * (size_t)S * H);
  int nofe[WS_PTS];
  w->dist = (double *)malloc(sizeof(double) * 6 * (size_t)(c->Np > 0 ? c->Np : 1));
  int rc = mbt_ws_generate(c, slices, w->npop_h, nofe, w->dist);
  if (rc) {
    free(slices);
    mbt_ws_destroy(w);
    return rc;
  }
  w->bucket_h = (int *)malloc(sizeof(int) * G);
  w->fill_h = (int *)malloc(sizeof(int) * H);
  memcpy(w->bucket_h, c->bucket, sizeof(int) * G);
  memcpy(w->fill_h, c->nfFill, sizeof(int) * H);
  double *Ib = (double *)calloc(H, sizeof(double));
  if (flags & MBT_WS_BUNCH_CURRENTS) memcpy(Ib, c->Ibunch, sizeof(double) * H);
  // ---- constants (glibc on the host, the reference's expressions)
  WSP &P = w->P;
  P.S = S;
  P.G = G;
  P.h = H;
  P.Nmlt = c->Nmlt;
  P.NmultiT = c->NmultiT;
  P.fbii = c->EnableFBII != 0;
  P.rw = c->EnableRW_long != 0;
  P.track_h = c->TrackPlane[1] != 0;
  P.track_v = c->TrackPlane[2] != 0;
  for (int p = 0; p < 3; p++) P.track[p] = c->TrackPlane[p] != 0;
  P.Nions = S * 25;
  P.mixed = (flags & MBT_WS_MIXED) ? 1 : 0;
  P.eps = c->Vrf0 / (c->E0 * MBT_FKILO);
  P.wrf = c->wrf;
  P.phai0 = c->phai0;
  P.Urad = c->Urad;
  P.T0 = c->T0;
  P.ac = c->ac;
  for (int p = 0; p < 3; p++) {
    P.gamma1[p] = c->gamma1[p];
    P.alpha1[p] = c->alpha1[p];
    P.beta1[p] = c->beta1[p];
  }
  {
    double PsiV0 = 2.0 * M_PI * c->QV0 / 1.0, PsiH0 = 2.0 * M_PI * c->QH0 / 1.0;
    double cv = cos(PsiV0), sv = sin(PsiV0), ch = cos(PsiH0), sh = sin(PsiH0);
    P.v11 = cv + c->alpha1[2] * sv;
    P.v12 = c->beta1[2] * sv;
    P.v21 = -c->gamma1[2] * sv;
    P.v22 = cv - c->alpha1[2] * sv;
    P.h11 = ch + c->alpha1[1] * sh;
    P.h12 = c->beta1[1] * sh;
    P.h21 = -c->gamma1[1] * sh;
    P.h22 = ch - c->alpha1[1] * sh;
    for (int pl = 0; pl < 2; pl++) {
      const int plane = pl + 1;
      double Psi0 = 2.0 * M_PI * (plane == 1 ? c->QH0 : c->QV0) / 1.0;
      double Psi = 0.0 + Psi0;
      double cp = cos(Psi), sp = sin(Psi);
      P.rwa11[pl] = cp + c->alpha1[plane] * sp;
      P.rwa12[pl] = +c->beta1[plane] * sp;
      const double kappa = 1.0 / 2.8e-08;
      double beff3 = pow(plane == 1 ? c->beffH[0] : c->beffV[0], 3);
      P.prfrw[pl] = 2.0 * c->R / (beff3) / sqrt(M_PI * kappa * K_EPS0) / 1.0;
      beff3 = pow(2.0e-3, 3);
      P.prfrw_trig[pl] = 2.0 * c->R / (beff3) / sqrt(M_PI * kappa * K_EPS0) / 1.0;
    }
  }
  P.triggerRW = c->triggerRW;
  P.Pref0 = c->T0 / c->E0 / c->fNp / MBT_FTERA;
  P.fNp = c->fNp;
  P.Dub0 = c->Lc / ((double)c->Nharm);
  P.Lc = c->Lc;
  P.factBE = sqrt(2 * (c->sgm_xx * c->sgm_xx - c->sgm_zz * c->sgm_zz)) / MBT_FKILO;
  P.sxx = c->sgm_xx;
  P.szz = c->sgm_zz;
  P.deltT = 0.3 * c->sgmatau;
  mbt_ws_grid(c, &P.aLx, &P.aLz);
  P.aAx = 2.0 * P.aLx;
  P.aAz = 2.0 * P.aLz;
  P.iaAx = 1.0 / P.aAx;
  P.iaAz = 1.0 / P.aAz;
  P.den_dQ = 2 * K_EPS0 * sqrt(M_PI) * P.factBE * K_C * P.deltT;
  {
    double Pave = 2.7e-9, Pu20 = 5 * Pave, Pm = 0.15 * Pu20 / 1.3332, dm = 3.22e22 * Pm, sigmam = 1.78;
    double Lu20 = 354.0, Ldmsgm = Lu20 * dm * sigmam * 1e-22;
    P.factF = K_E / (28.0 * 1.836e3 * 9.1e-31);
    P.factjp0 = 1e+2 * 4.0 * M_PI * K_EPS0 * 2.819e-15 / c->Gamma;
    P.CdeltT = K_C * P.deltT;
    P.dtF = P.deltT * P.factF;
    w->nelec_h = (double *)malloc(sizeof(double) * G);
    double *conv = (double *)malloc(sizeof(double) * G);
    for (int b = 0; b < G; b++) {
      w->nelec_h[b] = (Ib[c->bucket[b]] * c->T0 / (c->Np * K_E * MBT_FKILO));
      conv[b] = w->nelec_h[b] * Ldmsgm;
    }
    // ---- device memory
    const size_t npts = (size_t)S * 25;
    WA(w->D.sl, (size_t)G * 6 * S);
    WA(w->d_npop, S);
    WA(w->d_nofe, npts);
    WA(w->d_bucket, G);
    WA(w->d_fill, H);
    WA(w->d_bob, H);
    WA(w->d_Ibunch, H);
    WA(w->d_nelec, G);
    WA(w->d_convion, G);
    WA(w->D.hist, (size_t)4 * H * WS_HIST);
    WA(w->D.Exy, (size_t)G * npts);
    WA(w->D.li, (size_t)5 * G * npts);
    // train: drift times, then the ion-parallel form's partials (one block
    // row per bunch for every block holding an ion made ahead of it), or the
    // relay's cloud when those would not fit
    double *dt = (double *)calloc(G, sizeof(double));
    for (int ib = 1; ib < G; ib++) dt[ib] = fabs((double)(w->bucket_h[ib] - w->bucket_h[ib - 1])) * (c->T0 / c->h);
    WA(w->d_dtime, G);
    WUP(w->d_dtime, dt, G);
    CK(hipStreamSynchronize(w->stream));
    free(dt);
    const int NQ = 2 * S + 3;
    const long long nI = (long long)(G - 1) * (long long)npts;
    w->D.nblk_ions = (int)((nI + WS_WG - 1) / WS_WG);
    long long *toff = (long long *)malloc(sizeof(long long) * G), ntp = 0;
    for (int ib = 0; ib < G; ib++) {
      toff[ib] = ntp;
      long long nb = ((long long)ib * (long long)npts + WS_WG - 1) / WS_WG;
      if (nb > w->D.nblk_ions) nb = w->D.nblk_ions;
      ntp += nb * NQ;
    }
    w->serial = (flags & MBT_WS_SERIAL_TRAIN) || G < 2 || ntp * (long long)sizeof(double) > WS_TRAIN_PARTIAL_MAX;
    if (!w->serial && P.fbii) {
      WA(w->D.tpart, ntp);
      WA(w->D.tot, (size_t)G * WS_TQ);
      WA(w->d_toff, G);
      WUP(w->d_toff, toff, G);
      CK(hipStreamSynchronize(w->stream));
    } else if (P.fbii) {
      w->D.cap = (long long)G * npts;
      WA(w->D.cl, (size_t)5 * w->D.cap);
      w->maxblk = (int)((w->D.cap + WS_WG - 1) / WS_WG);
      WA(w->D.part, (size_t)w->maxblk * S * 2);
      WA(w->D.psum, (size_t)w->maxblk * 3);
    }
    free(toff);
    WA(w->D.done, 1);
    WA(w->D.cm, (size_t)WS_T * G * 4);
    WA(w->D.mon, (size_t)WS_T * G * 18);
    WA(w->D.ion, (size_t)WS_T * G * 3);
    WA(w->D.zcm, (size_t)WS_T * H);
    double *sl = (double *)malloc(sizeof(double) * 6 * (size_t)S * G);
    int *bob = (int *)malloc(sizeof(int) * H);
    for (int kb = 0; kb < H; kb++) bob[kb] = -1;
    for (int b = 0; b < G; b++) {
      memcpy(sl + (size_t)b * 6 * S, slices + (size_t)c->bucket[b] * 6 * S, sizeof(double) * 6 * S);
      bob[c->bucket[b]] = b;
    }
    WUP(w->D.sl, sl, (size_t)G * 6 * S);
    WUP(w->d_npop, w->npop_h, S);
    WUP(w->d_nofe, nofe, npts);
    WUP(w->d_bucket, w->bucket_h, G);
    WUP(w->d_fill, w->fill_h, H);
    WUP(w->d_bob, bob, H);
    WUP(w->d_Ibunch, Ib, H);
    WUP(w->d_nelec, w->nelec_h, G);
    WUP(w->d_convion, conv, G);
    CK(hipStreamSynchronize(w->stream));
    free(sl);
    free(bob);
    free(conv);
  }
  free(slices);
  free(Ib);
  w->D.npop = w->d_npop;
  w->D.nofe = w->d_nofe;
  w->D.bucket = w->d_bucket;
  w->D.fill = w->d_fill;
  w->D.Ibunch = w->d_Ibunch;
  w->D.nelec = w->d_nelec;
  w->D.convion = w->d_convion;
  w->D.dtime = w->d_dtime;
  w->D.toff = w->d_toff;
  *out = w;
  return MBT_OK;
}

static int ws_download(mbt_ws *w, long n, double *cm, double *mon, double *ion, double *zcm)
{
  CK(hipStreamSynchronize(w->stream));
  const int G = w->P.G, H = w->P.h;
  if (cm) CK(hipMemcpy(cm, w->D.cm, sizeof(double) * 4 * G * n, hipMemcpyDeviceToHost));
  if (mon) CK(hipMemcpy(mon, w->D.mon, sizeof(double) * 18 * G * n, hipMemcpyDeviceToHost));
  if (ion) CK(hipMemcpy(ion, w->D.ion, sizeof(double) * 3 * G * n, hipMemcpyDeviceToHost));
  if (zcm) CK(hipMemcpy(zcm, w->D.zcm, sizeof(double) * H * n, hipMemcpyDeviceToHost));
  return MBT_OK;
}

// Enqueues n <= WS_T turns on the object's stream, their outputs in slots
// 0 .. n-1 of the device's output block (with HIP events around each ion train
// when timing is on).
static int ws_enqueue(mbt_ws *w, long n)
{
  const int G = w->P.G, H = w->P.h;
  const size_t lds2 = sizeof(double) * 2 * H;
  // phase 1: a block per bunch, and with FBII one thread per grid point of
  // every slice of every bunch (for the fields, then the ions)
  const long long np = (long long)G * w->P.S * 25;
  const int nbp = (int)((np + WS_WG - 1) / WS_WG), nb1 = w->P.fbii && nbp > G ? nbp : G;
  const int nkick = w->P.fbii && !w->serial ? G : 0;
  CK(hipMemsetAsync(w->D.mon, 0, sizeof(double) * 18 * G * n, w->stream));
  CK(hipMemsetAsync(w->D.ion, 0, sizeof(double) * 3 * G * n, w->stream));
  for (long t = 0; t < n; t++) {
    const long long rev = w->rev;
    const int monitor = rev % w->cfg.NrevMon == 0;
    hipLaunchKernelGGL(k_ws_phase1, dim3(nb1), dim3(WS_WG), 0, w->stream, w->P, w->D, rev, (int)t, monitor);
    CK(hipGetLastError());
    if (w->P.fbii) {
      // the ions start from the slices' positions before the betatron rotation
      hipLaunchKernelGGL(k_ws_create, dim3(nbp), dim3(WS_WG), 0, w->stream, w->P, w->D);
      CK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_ws_phase2, dim3(G), dim3(WS_WG), lds2, w->stream, w->P, w->D, rev);
    CK(hipGetLastError());
    if (w->timing && w->P.fbii) CK(hipEventRecord(w->ev[2 * t], w->stream));
    if (w->P.fbii && !w->serial) {
      hipLaunchKernelGGL(k_ws_ions, dim3(w->D.nblk_ions), dim3(WS_WG), sizeof(double2) * 2 * WS_QT * min(w->P.S, WS_CH),
                         w->stream, w->P, w->D);
      CK(hipGetLastError());
      if (w->timing) CK(hipEventRecord(w->ev[2 * t + 1], w->stream));
      hipLaunchKernelGGL(k_ws_ksum, dim3(G, (2 * w->P.S + 3 + KSUM_Q - 1) / KSUM_Q + 1), dim3(64 * KSUM_Q), 0,
                         w->stream, w->P, w->D);
      CK(hipGetLastError());
    } else if (w->P.fbii) {
      hipLaunchKernelGGL(k_ws_cloud_reset, dim3(1), dim3(WS_WG), 0, w->stream, w->P, w->D, (int)t);
      CK(hipGetLastError());
      for (int ib = 1; ib < G; ib++) {
        const long long nF = (long long)ib * w->P.Nions;
        const double Dtime =
            fabs((double)(w->bucket_h[ib] - w->bucket_h[ib - 1])) * (w->cfg.T0 / w->cfg.h);
        const int nblk = (int)((nF + WS_WG - 1) / WS_WG);
        hipLaunchKernelGGL(k_ws_train, dim3(nblk), dim3(WS_WG), 0, w->stream, w->P, w->D, ib, nF, Dtime, (int)t);
        CK(hipGetLastError());
      }
      if (w->timing) CK(hipEventRecord(w->ev[2 * t + 1], w->stream));
    }
    hipLaunchKernelGGL(k_ws_end, dim3(nkick + (H + 255) / 256), dim3(256), 0, w->stream, w->P, w->D, w->d_bob, (int)t,
                       nkick);
    CK(hipGetLastError());
    w->rev++;
  }
  return MBT_OK;
}

extern "C" int mbt_ws_track(mbt_ws *w, long nturns, double *cm, double *mon, double *ion, double *zcm)
{
  int rc;
  if ((rc = ws_no_pending(w))) return rc;
  CK(hipSetDevice(w->device));
  (void)hipGetLastError();
  const int G = w->P.G, H = w->P.h;
  long done = 0;
  w->train_ms = 0.0;
  w->train_launches = 0;
  while (done < nturns) {
    long n = nturns - done;
    if (n > WS_T) n = WS_T;
    if ((rc = ws_enqueue(w, n))) return rc;
    if (w->timing && w->P.fbii) {
      CK(hipStreamSynchronize(w->stream));
      for (long t = 0; t < n; t++) {
        float ms = 0.0f;
        CK(hipEventElapsedTime(&ms, w->ev[2 * t], w->ev[2 * t + 1]));
        w->train_ms += ms;
        w->train_launches++;
      }
    }
    rc = ws_download(w, n, cm ? cm + (size_t)done * G * 4 : NULL, mon ? mon + (size_t)done * G * 18 : NULL,
                     ion ? ion + (size_t)done * G * 3 : NULL, zcm ? zcm + (size_t)done * H : NULL);
    if (rc) return rc;
    done += n;
  }
  return MBT_OK;
}

extern "C" long mbt_ws_async_max(const mbt_ws *w)
{
  (void)w;
  return WS_T;
}

extern "C" int mbt_ws_track_async(mbt_ws *w, long nturns)
{
  int rc;
  if ((rc = ws_no_pending(w))) return rc;
  if (nturns < 1 || nturns > WS_T) {
    mbt_set_error("mbt_ws_track_async: 1 .. %d turns per call", WS_T);
    return MBT_EINVAL;
  }
  if (w->timing) {
    mbt_set_error("mbt_ws_track_async: not with train timing on");
    return MBT_EINVAL;
  }
  CK(hipSetDevice(w->device));
  (void)hipGetLastError(); // launch checks below must not report an earlier call's error
  if ((rc = ws_enqueue(w, nturns))) return rc;
  w->async_n = nturns;
  return MBT_OK;
}

extern "C" int mbt_ws_track_wait(mbt_ws *w, double *cm, double *mon, double *ion, double *zcm)
{
  if (!w->async_n) {
    mbt_set_error("mbt_ws_track_wait: no turns pending");
    return MBT_EINVAL;
  }
  CK(hipSetDevice(w->device));
  const long n = w->async_n;
  w->async_n = 0;
  return ws_download(w, n, cm, mon, ion, zcm);
}

extern "C" int mbt_ws_get_slices(const mbt_ws *w, int ib, double *out)
{
  if (ib < 0 || ib >= w->P.G) return MBT_EINVAL;
  CK(hipSetDevice(w->device));
  CK(hipStreamSynchronize(w->stream));
  CK(hipMemcpy(out, w->D.sl + (size_t)ib * 6 * w->P.S, sizeof(double) * 6 * w->P.S, hipMemcpyDeviceToHost));
  return MBT_OK;
}

extern "C" int mbt_ws_population(const mbt_ws *w, int *npop, int *nslices)
{
  for (int i = 0; i < w->P.S; i++) npop[i] = w->npop_h[i];
  if (nslices) *nslices = w->P.S;
  return MBT_OK;
}

extern "C" int mbt_ws_distribution(const mbt_ws *w, double *dist)
{
  memcpy(dist, w->dist, sizeof(double) * 6 * (size_t)w->cfg.Np);
  return MBT_OK;
}

extern "C" long mbt_ws_turn(const mbt_ws *w) { return (long)w->rev; }

extern "C" int mbt_ws_set_timing(mbt_ws *w, int on)
{
  CK(hipSetDevice(w->device));
  if (on && !w->ev) {
    w->ev = (hipEvent_t *)calloc(2 * WS_T, sizeof(hipEvent_t));
    if (!w->ev) return MBT_ENOMEM;
    for (int i = 0; i < 2 * WS_T; i++) CK(hipEventCreate(&w->ev[i]));
  }
  w->timing = on ? 1 : 0;
  return MBT_OK;
}

extern "C" int mbt_ws_train_time(const mbt_ws *w, double *ms, long *launches)
{
  if (!w || !ms || !launches) {
    mbt_set_error("mbt_ws_train_time: null argument");
    return MBT_EINVAL;
  }
  *ms = w->train_ms;
  *launches = w->train_launches;
  return MBT_OK;
}
