// The multi-bunch tracking turn on one MI355X (gfx950).
//
// Every bunch is resident in HBM as six FP64 coordinate arrays (tau, delta, x,
// x', z, z'), bunches back to back.  One turn is four kernels on one stream:
//
//   k_lr      long-range resonator phasor recurrence over all buckets and the
//             RF-feedback solve (construct_wake_phasor src/transform_weak.c:961-1048,
//             src/worker_weakweak.c:268-278): sequential by nature, one lane per
//             resonator, bit-identical arithmetic to the reference;
//   k_turn    one pass over every macroparticle: pending RW kick, selffield
//             kick (src/transform_weak.c:697-731), one-turn map
//             (src/transform_weak.c:21-192), shifted moment sums and the next
//             turn's slice histogram with LDS-privatised exact integer sums;
//   k_reduce  per bunch: statistics (src/statistics.c:91-195), RW dipole history,
//             LR histograms (src/transform_weak.c:839-865) and the selffield
//             convolutions for the next turn (src/transform_weak.c:633-731);
//   k_rw      per bunch: resistive-wall long-range sum over (Nmlt+1) h past
//             buckets (src/transform_weak.c:780-833), applied as a pending kick
//             at the start of the next pass; k_end advances the turn counter.
//
// Determinism: histogram dipole sums are accumulated as exact 128-bit fixed-point
// integers (LDS then global integer atomics), moment sums by fixed-order
// reductions, so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "mbt_host.h"

#pragma clang fp contract(off)

#define WG 256
#define CHUNK_ITEMS 8
#define NMOM 15
#define STATS_TURNS 256
#define MAXLR (3 * MBT_MAX_LR)

#define CK(x)                                                                                    \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      mbt_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);     \
      return MBT_EDEVICE;                                                                        \
    }                                                                                            \
  } while (0)

struct DevParams {
  int Nb, h, Nc, Nmlt;
  int track_h, track_v;
  int qe, ideal_hc, amp_dq;
  int kickL, kickV, kickH; // selffield kick of the plane is applied
  int convL;               // longitudinal Green function present (PlaneL > 0)
  int scan;
  long long NrevScan;
  double scan_step;
  long long Np;
  int sf_active;
  int rw;
  int n_ahc;
  int monitor_every;
  unsigned seed;
  int m_aHC;
  double ahc_amp[MBT_MAX_AHC], ahc_wn[MBT_MAX_AHC], ahc_phi[MBT_MAX_AHC], ahc_nHC[MBT_MAX_AHC];
  double E0, E0k; // E0 [GeV], E0 * FKILO
  double RsisL, aindL;
  double wrf, Urad, damp, excite, qVp, qVs, qHp, qHs, T0, ac;
  double PsiV0, PsiH0, Gziz, Gzix;
  double aV, bV, gV, aH, bH, gH, dispH, disppH;
  double CVH, CVV, CHH, CHV;
  double HC_k, m_wrf, m_phi;
  double sigma_tau, Nsigma, dT;
  double gamma1[3], alpha1[3], beta1[3];
  double fNp;
  double rwconst[2][2]; // [plane HOR/VER][trigger off/on]
};

struct TurnState {
  long long rev;
  int m;
  int has_fb;
  double phai0, Vrf0;
  double phai0_prev, Vrf0_prev; // before this turn's RF-feedback solve
  double Iring, scan_val;
  long long fb_count;
};

struct LRConst {
  int plane, l, fb;
  double pr0, pr1;   // one bin
  double ge, gc, gs; // bucket gap
  double ee, ec, es; // empty bucket
  double pb0, pb1, pa0, pa1;
  double fac;
};

// ---------------------------------------------------------------- device math

// Philox4x32-10 and the QE normals: identical algorithm to the oracle's
// orc_qe_normals (float Box-Muller from explicit-FMA polynomials), so the
// normals are bit-identical on CPU and GPU.
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ float qe_logf(float u)
{
  uint32_t i = __float_as_uint(u);
  int e = (int)((i >> 23) & 0xff) - 127;
  float m = __uint_as_float((i & 0x007fffffu) | 0x3f800000u);
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

__device__ __forceinline__ void qe_sincos2pi(float t, float *s, float *c)
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
  float so, co;
  if (iq == 0) { so = sn; co = cs; }
  else if (iq == 1) { so = cs; co = -sn; }
  else if (iq == 2) { so = -sn; co = -cs; }
  else { so = -cs; co = sn; }
  *s = so;
  *c = co;
}

__device__ __forceinline__ float u01(uint32_t x) { return ((float)((x >> 8) + 1u)) * 5.9604645e-8f; }

__device__ __forceinline__ void bm_pair(uint32_t a, uint32_t b, double &n0, double &n1)
{
  float r = sqrtf(-2.0f * qe_logf(u01(a)));
  float s, c;
  qe_sincos2pi(u01(b), &s, &c);
  n0 = (double)(r * c);
  n1 = (double)(r * s);
}

__device__ __forceinline__ void qe_normals(uint32_t seed, uint32_t bucket, unsigned long long rev, uint32_t j, double n[5])
{
  uint32_t c[4] = {j, (uint32_t)rev, (uint32_t)(rev >> 32), 0u};
  philox(c, seed, bucket);
  bm_pair(c[0], c[1], n[0], n[1]);
  bm_pair(c[2], c[3], n[2], n[3]);
  uint32_t d[4] = {j, (uint32_t)rev, (uint32_t)(rev >> 32), 1u};
  philox(d, seed, bucket);
  double t;
  bm_pair(d[0], d[1], n[4], t);
}

// Slice index of src/transform_weak.c:611-613 with x86-64 conversion semantics
// (NaN and out-of-range values give INT_MIN, as cvttsd2si does).
__device__ __noinline__ int bin_of(double tau, double sgm, double Nsigma, double dT)
{
  double Tau = -tau / sgm;
  double v = (Tau + Nsigma) / dT + 1;
  if (!(v > -2147483649.0 && v < 2147483648.0)) return INT_MIN;
  return (int)v;
}

// Exact 128-bit fixed point, quantum 2^-100: q = round(x 2^100) for |x| < 2^20.
__device__ __forceinline__ void fx_split(double x, unsigned long long &lo, unsigned long long &hi)
{
  unsigned long long bits = (unsigned long long)__double_as_longlong(x);
  int e = (int)((bits >> 52) & 0x7ff);
  unsigned long long m = (bits & 0xfffffffffffffULL) | (1ULL << 52);
  lo = 0;
  hi = 0;
  if (e == 0) return;
  int s = e - 975;
  if (s >= 64) {
    hi = m << (s - 64);
  } else if (s > 0) {
    lo = m << s;
    hi = m >> (64 - s);
  } else if (s == 0) {
    lo = m;
  } else if (s > -64) {
    int r = -s;
    lo = (m >> r) + ((m >> (r - 1)) & 1ULL);
  } else if (s == -64) {
    lo = (m >> 63) & 1ULL;
  }
  if (bits >> 63) { // negate
    lo = ~lo + 1ULL;
    hi = ~hi + (lo == 0ULL ? 1ULL : 0ULL);
  }
}

__device__ __forceinline__ void fx_atomic_add(unsigned long long *plo, unsigned long long *phi, unsigned long long lo,
                                              unsigned long long hi)
{
  unsigned long long old = atomicAdd(plo, lo);
  unsigned long long carry = (old + lo) < old ? 1ULL : 0ULL;
  atomicAdd(phi, hi + carry);
}

__device__ __forceinline__ double fx_to_double(unsigned long long lo, unsigned long long hi)
{
  long long h = (long long)hi;
  double v = (double)h * 18446744073709551616.0 + (double)lo;
  return v * 7.888609052210118e-31; // 2^-100
}

// ------------------------------------------------------------------ k_turn

enum { MODE_TURN = 0, MODE_STATS = 1, MODE_FINAL = 2 };

struct HistPtr {
  unsigned int *cnt;        // [Nb*Nc]
  unsigned long long *sx;   // [Nb*Nc*2] lo,hi
  unsigned long long *sz;   // [Nb*Nc*2]
  unsigned int *tcnt;       // [Nb*2] low, high
  unsigned long long *tsx;  // [Nb*2*2]
  unsigned long long *tsz;  // [Nb*2*2]
  int *bad;                 // [Nb]
};

struct Coords {
  double *tau, *del, *x, *xp, *z, *zp;
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sin and cos of x for |x| < 2^19 pi/2: Cody-Waite reduction by pi/2 in three
// parts (fdlibm's pio2_1, pio2_2, pio2_2t; the first two carry 33 bits, so k*part
// is exact for |k| < 2^20), then Taylor polynomials to degree 17/18 on
// [-pi/4, pi/4] (truncation < 1e-19), cos with the 1 - z/2 correction.  Larger
// arguments fall back to the library sincos.
__device__ __noinline__ double2 sincos_lib(double x)
{
  double2 r;
  sincos(x, &r.x, &r.y);
  return r;
}

__device__ __forceinline__ void sincos_cw(double x, double *s, double *c)
{
  if (!(fabs(x) < 823549.6)) {
    const double2 r = sincos_lib(x);
    *s = r.x;
    *c = r.y;
    return;
  }
  const double k = rint(x * 6.36619772367581382433e-01);
  double r = fma(-k, 1.57079632673412561417e+00, x);
  r = fma(-k, 6.07710050630396597660e-11, r);
  r = fma(-k, 2.02226624879595063154e-21, r);
  const double z = r * r;
  double ps = fma(z, 2.8114572543455206e-15, -7.647163731819816e-13);
  ps = fma(z, ps, 1.6059043836821613e-10);
  ps = fma(z, ps, -2.505210838544172e-08);
  ps = fma(z, ps, 2.7557319223985893e-06);
  ps = fma(z, ps, -0.0001984126984126984);
  ps = fma(z, ps, 0.008333333333333333);
  ps = fma(z, ps, -0.16666666666666666);
  const double sn = fma(r * z, ps, r);
  double pc = fma(z, -1.5619206968586225e-16, 4.779477332387385e-14);
  pc = fma(z, pc, -1.1470745597729725e-11);
  pc = fma(z, pc, 2.08767569878681e-09);
  pc = fma(z, pc, -2.755731922398589e-07);
  pc = fma(z, pc, 2.48015873015873e-05);
  pc = fma(z, pc, -0.001388888888888889);
  pc = fma(z, pc, 0.041666666666666664);
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double cs = w + (((1.0 - w) - hz) + (z * z) * pc);
  // quadrant without branches: swap on odd q, sin negative for q = 2,3, cos for q = 1,2
  const int q = ((int)k) & 3;
  const double a = (q & 1) ? cs : sn;
  const double bq = (q & 1) ? sn : cs;
  *s = (q & 2) ? -a : a;
  *c = ((q + 1) & 2) ? -bq : bq;
}

__device__ __forceinline__ double sin_cw(double x)
{
  double s, c;
  sincos_cw(x, &s, &c);
  return s;
}

// Slice index with reciprocal multiplies, falling back to the exact reference
// expression (two correctly rounded divides) whenever the truncation could
// depend on the last bits: the approximate value is within 1e-12 of the exact
// one for |v| < 1e6.
struct BinK {
  double sgm, inv_sgm, Nsigma, dT, inv_dT;
};
__device__ __forceinline__ int bin_fast(double tau, const BinK &k)
{
  const double vf = fma(fma(-tau, k.inv_sgm, k.Nsigma), k.inv_dT, 1.0);
  if (fabs(vf) < 1.0e6) {
    const double lo = trunc(vf - 1.0e-9), hi = trunc(vf + 1.0e-9);
    if (lo == hi) return (int)lo;
  }
  return bin_of(tau, k.sgm, k.Nsigma, k.dT);
}

// Block index -> work item, so that consecutive work items (the chunks of one
// bunch) run on the same XCD (blocks are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ int xcd_remap(int bid, int n)
{
  const int x = bid & 7, i = bid >> 3, q = n >> 3, r = n & 7;
  return x * q + (x < r ? x : r) + i;
}

template <int MODE, bool QE, bool DIP, bool GEN, bool MON>
__global__ __launch_bounds__(WG) void k_turn(DevParams P, Coords C, const long long *wg_first, const long long *wg_end,
                                             const int *wg_bunch, const int *bucket, const double *shift,
                                             const double *sfL, const double *sfV, const double *sfH,
                                             const double *lrL, const double *lrV, const double *lrH,
                                             const double *pending, double *partial, HistPtr H,
                                             const TurnState *ts)
{
  extern __shared__ unsigned long long smem[];
  const int Nc = P.Nc;
  // LDS: sx lo/hi [2Nc], sz lo/hi [2Nc], trash sums [8], kick tables [3Nc], counts [Nc] + trash [2]
  unsigned long long *lsx = smem;
  unsigned long long *lsz = lsx + 2 * Nc;
  unsigned long long *ltr = lsz + 2 * Nc;
  double *kL = (double *)(ltr + 8);
  double *kV = kL + Nc;
  double *kH = kV + Nc;
  unsigned int *lcnt = (unsigned int *)(kH + Nc);
  unsigned int *ltc = lcnt + Nc;
  __shared__ double red[WG / 64][NMOM];
  __shared__ int lbad;

  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const int b = wg_bunch[wg];
  const int kb = bucket[b];
  const long long j0 = wg_first[wg], j1 = wg_end[wg];
  const int tid = threadIdx.x;
  const long long rev = ts->rev;
  const bool do_hist = Nc > 0 && MODE != MODE_FINAL;

  if (DIP)
    for (int i = tid; i < 2 * Nc; i += WG) {
      lsx[i] = 0;
      lsz[i] = 0;
    }
  for (int i = tid; i < Nc; i += WG) {
    lcnt[i] = 0;
    if (MODE == MODE_TURN) {
      if (P.kickL) kL[i] = sfL[(size_t)b * Nc + i] + lrL[(size_t)kb * Nc + i];
      if (P.kickV) kV[i] = sfV[(size_t)b * Nc + i] + lrV[(size_t)kb * Nc + i];
      if (P.kickH) kH[i] = sfH[(size_t)b * Nc + i] + lrH[(size_t)kb * Nc + i];
    }
  }
  if (tid < 8) ltr[tid] = 0;
  if (tid < 2) ltc[tid] = 0;
  if (tid == 0) lbad = 0;
  __syncthreads();

  const BinK bk = {P.sigma_tau, 1.0 / P.sigma_tau, P.Nsigma, P.dT, 1.0 / P.dT};
  double sh[6];
#pragma unroll
  for (int c = 0; c < 6; c++) sh[c] = shift[b * 6 + c];
  const bool rw_h = MODE != MODE_STATS && P.rw && P.track_h;
  const bool rw_v = MODE != MODE_STATS && P.rw && P.track_v;
  const double rwx = rw_h ? pending[2 * b] : 0.0;
  const double rwz = rw_v ? pending[2 * b + 1] : 0.0;
  const double eps_const = ts->Vrf0 / P.E0k;
  const double phai0 = ts->phai0;
  const long long jb0 = (long long)b * P.Np;

  double s1[6] = {0, 0, 0, 0, 0, 0}, s2[6] = {0, 0, 0, 0, 0, 0}, am[3] = {0, 0, 0};

  for (long long j = j0 + tid; j < j1; j += WG) {
    double tau = C.tau[j], del = C.del[j], x = C.x[j], xp = C.xp[j], z = C.z[j], zp = C.zp[j];
    if (rw_h) xp += rwx;
    if (rw_v) zp += rwz;
    if (MODE == MODE_TURN) {
      if (P.sf_active) {
        int bn = bin_fast(tau, bk);
        bn = bn < 0 ? 0 : (bn >= Nc ? Nc - 1 : bn);
        if (P.kickL) del += kL[bn];
        if (P.kickV) zp += kV[bn];
        if (P.kickH) xp += kH[bn];
      }
      // ---- one-turn map, src/transform_weak.c:21-192 (operation order kept)
      const double tau0 = tau, del0 = del, x0 = x, xp0 = xp, z0 = z, zp0 = zp;
      double g;
      if (!GEN) {
        g = eps_const * sin_cw(P.wrf * tau0 + phai0);
      } else if (!P.ideal_hc) {
        g = eps_const * sin_cw(P.wrf * tau0 + phai0);
        for (int a = 0; a < P.n_ahc; a++)
          g += P.ahc_amp[a] * sin_cw(P.ahc_wn[a] * tau0 + P.ahc_phi[a] + P.ahc_nHC[a] * kb * 2 * M_PI);
      } else {
        g = eps_const * (sin_cw(P.wrf * tau0 + phai0) + P.HC_k * sin_cw(P.m_wrf * tau0 + P.m_phi));
      }
      del = del0 + g - P.Urad;
      double nq[5], rff = 0.0;
      if (QE) {
        qe_normals(P.seed, (uint32_t)kb, (unsigned long long)rev, (uint32_t)(j - jb0), nq);
        del = del * P.damp + P.excite * nq[0];
        rff = (1.0 + del0) / (1.0 + del0 + g);
      }
      tau = tau0 - del * P.T0 * P.ac;
      if (P.track_v) {
        double dQ = 0.0;
        if (GEN && P.amp_dq) {
          double EV = P.gV * z0 * z0 + 2 * P.aV * z0 * zp0 + P.bV * zp0 * zp0;
          double EH = P.gV * x0 * x0 + 2 * P.aV * x0 * xp0 + P.bV * xp0 * xp0;
          dQ = P.CVH * EH + P.CVV * EV;
        }
        double sn, cs;
        sincos_cw(P.PsiV0 * (1.0 + P.Gziz * del0 + dQ), &sn, &cs);
        double a11 = cs + P.aV * sn, a12 = P.bV * sn, a21 = -P.gV * sn, a22 = cs - P.aV * sn;
        z = a11 * z0 + a12 * zp0;
        zp = a21 * z0 + a22 * zp0;
        if (QE) {
          z = z + P.qVp * nq[1];
          zp = rff * zp + P.qVs * nq[2];
        }
      }
      if (P.track_h) {
        double dQ = 0.0;
        if (GEN && P.amp_dq) {
          double EV = P.gV * z0 * z0 + 2 * P.aV * z0 * zp0 + P.bV * zp0 * zp0;
          double EH = P.gV * x0 * x0 + 2 * P.aV * x0 * xp0 + P.bV * xp0 * xp0;
          dQ = P.CHH * EH + P.CHV * EV;
        }
        double sn, cs;
        sincos_cw(P.PsiH0 * (1.0 + P.Gzix * del0 + dQ), &sn, &cs);
        double a11 = cs + P.aH * sn;
        double a12 = P.bH * sn;
        double a13 = (1.0 - a11) * P.dispH - a12 * P.disppH;
        double a21 = -P.gH * sn;
        double a22 = cs - P.aH * sn;
        double a23 = -a21 * P.dispH + (1.0 - a22) * P.disppH;
        x = a11 * x0 + a12 * xp0 + a13 * del;
        xp = a21 * x0 + a22 * xp0 + a23 * del;
        if (QE) {
          x = x + P.qHp * nq[3];
          xp = rff * xp + P.qHs * nq[4];
        }
      }
    }
    if (MODE != MODE_STATS) {
      C.tau[j] = tau;
      C.del[j] = del;
      C.x[j] = x;
      C.xp[j] = xp;
      C.z[j] = z;
      C.zp[j] = zp;
    }
    // ---- moments about the per-bunch shift (reassociated: no bitwise target)
    const double v[6] = {tau, del, x, xp, z, zp};
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const double d = v[c] - sh[c];
      s1[c] += d;
      s2[c] = fma(d, d, s2[c]);
    }
    if (MON) {
      am[0] += P.gamma1[0] * (tau * tau) + 2.0 * P.alpha1[0] * tau * del + P.beta1[0] * (del * del);
      am[1] += P.gamma1[1] * (x * x) + 2.0 * P.alpha1[1] * x * xp + P.beta1[1] * (xp * xp);
      am[2] += P.gamma1[2] * (z * z) + 2.0 * P.alpha1[2] * z * zp + P.beta1[2] * (zp * zp);
    }
    // ---- next turn's slice histogram
    if (do_hist) {
      const int bn = bin_fast(tau, bk);
      const bool inr = bn >= 0 && bn < Nc;
      if (inr) atomicAdd(&lcnt[bn], 1u);
      else atomicAdd(&ltc[bn < 0 ? 0 : 1], 1u);
      if (DIP) {
        unsigned long long xl, xh, zl, zh;
        if (!(fabs(x) < 1048576.0) || !(fabs(z) < 1048576.0)) {
          lbad = 1;
          xl = xh = zl = zh = 0;
        } else {
          fx_split(x, xl, xh);
          fx_split(z, zl, zh);
        }
        if (inr) {
          fx_atomic_add(&lsx[2 * bn], &lsx[2 * bn + 1], xl, xh);
          fx_atomic_add(&lsz[2 * bn], &lsz[2 * bn + 1], zl, zh);
        } else {
          const int t = bn < 0 ? 0 : 1;
          fx_atomic_add(&ltr[4 * t], &ltr[4 * t + 1], xl, xh);
          fx_atomic_add(&ltr[4 * t + 2], &ltr[4 * t + 3], zl, zh);
        }
      }
    }
  }

  // ---- block reduction of the moment sums in a fixed order
  double vals[NMOM];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    vals[c] = s1[c];
    vals[6 + c] = s2[c];
  }
  vals[12] = am[0];
  vals[13] = am[1];
  vals[14] = am[2];
  const int lane = tid & 63, wv = tid >> 6;
#pragma unroll
  for (int c = 0; c < NMOM; c++) {
    if (!MON && c >= 12) continue;
    double sum = wave_sum(vals[c]);
    if (lane == 0) red[wv][c] = sum;
  }
  __syncthreads();
  if (tid < NMOM) {
    double sum = 0.0;
    if (MON || tid < 12)
      for (int w = 0; w < WG / 64; w++) sum += red[w][tid];
    partial[(size_t)wg * NMOM + tid] = sum;
  }
  // ---- flush the block histogram with exact integer atomics
  if (do_hist) {
    for (int i = tid; i < Nc; i += WG) {
      const unsigned int n = lcnt[i];
      if (n == 0) continue;
      const size_t gi = (size_t)b * Nc + i;
      atomicAdd(&H.cnt[gi], n);
      if (DIP) {
        fx_atomic_add(&H.sx[2 * gi], &H.sx[2 * gi + 1], lsx[2 * i], lsx[2 * i + 1]);
        fx_atomic_add(&H.sz[2 * gi], &H.sz[2 * gi + 1], lsz[2 * i], lsz[2 * i + 1]);
      }
    }
    if (tid < 2 && ltc[tid] > 0) {
      const int t = tid;
      atomicAdd(&H.tcnt[2 * b + t], ltc[t]);
      if (DIP) {
        fx_atomic_add(&H.tsx[4 * b + 2 * t], &H.tsx[4 * b + 2 * t + 1], ltr[4 * t], ltr[4 * t + 1]);
        fx_atomic_add(&H.tsz[4 * b + 2 * t], &H.tsz[4 * b + 2 * t + 1], ltr[4 * t + 2], ltr[4 * t + 3]);
      }
    }
    if (DIP && tid == 0 && lbad) H.bad[b] = 1;
  }
}

// ---------------------------------------------------------------- k_reduce

struct RedOut {
  double *stats;     // [STATS_TURNS][Nb][12]
  double *amp;       // [STATS_TURNS][Nb][6]
  double *shift;     // [Nb][6]
  double *rwhist;    // [(Nmlt+2)h][2]
  double *fnp_ring;  // [3][h*Nc]
  double *sfL, *sfV, *sfH;
  double *fnp_out, *dipV_out, *dipH_out; // last clamped histograms [Nb*Nc]
  int *trash_tot;    // [Nb*2] cumulative (consumed)
  int *trash_pend;   // [Nb*2]
  const double *Gl, *Gv, *Gh;
  const double *bunch_Ib;
  const int *np;
};

// phase: 0 = means only (shift <- mean), 1 = initial statistics + RW history
// fill, 2 = turn, 3 = final statistics
__global__ __launch_bounds__(WG) void k_reduce(DevParams P, const double *partial, const int *bunch_wg0, const int *bucket,
                                               HistPtr H, RedOut R, const TurnState *ts, int phase, int lr_L,
                                               int lr_H, int lr_V)
{
  extern __shared__ double sm[];
  const int Nc = P.Nc;
  double *fnp = sm;
  double *dV = fnp + Nc;
  double *dH = dV + Nc;
  __shared__ double mom[NMOM];
  __shared__ int icmin_s, icmax_s;
  const int b = blockIdx.x, tid = threadIdx.x, kb = bucket[b];
  const long long rev = ts->rev;
  const int w0 = bunch_wg0[b], w1 = bunch_wg0[b + 1];
  const double Np = (double)R.np[b];
  if (tid < NMOM) {
    double s = 0.0;
    for (int w = w0; w < w1; w++) s += partial[(size_t)w * NMOM + tid];
    mom[tid] = s;
  }
  __syncthreads();
  const int slot = (int)(((rev % STATS_TURNS) + STATS_TURNS) % STATS_TURNS);
  if (tid == 0) {
    double st[12];
    const int track[3] = {1, P.track_h, P.track_v};
    for (int pl = 0; pl < 3; pl++) {
      int cp = 2 * pl, cs = 2 * pl + 1;
      double mp = mom[cp] / Np, ms = mom[cs] / Np;
      double pos = R.shift[b * 6 + cp] + mp, slope = R.shift[b * 6 + cs] + ms;
      double vp = mom[6 + cp] / Np - mp * mp, vs = mom[6 + cs] / Np - ms * ms;
      st[pl] = track[pl] ? pos : 0.0;
      st[3 + pl] = track[pl] ? sqrt(vp > 0.0 ? vp : 0.0) : 0.0;
      st[6 + pl] = track[pl] ? slope : 0.0;
      st[9 + pl] = track[pl] ? sqrt(vs > 0.0 ? vs : 0.0) : 0.0;
    }
    if (phase == 0) {
      for (int pl = 0; pl < 3; pl++) {
        R.shift[b * 6 + 2 * pl] = st[pl];
        R.shift[b * 6 + 2 * pl + 1] = st[6 + pl];
      }
    } else {
      double *o = R.stats + ((size_t)slot * P.Nb + b) * 12;
      for (int i = 0; i < 12; i++) o[i] = st[i];
      for (int pl = 0; pl < 3; pl++) {
        R.shift[b * 6 + 2 * pl] = st[pl];
        R.shift[b * 6 + 2 * pl + 1] = st[6 + pl];
      }
      if (phase == 2 && (rev + 1) % P.monitor_every == 0) {
        double *a = R.amp + ((size_t)slot * P.Nb + b) * 6;
        for (int pl = 0; pl < 3; pl++) {
          a[pl] = track[pl] ? mom[12 + pl] / Np : 0.0;
          a[3 + pl] = track[pl] ? P.gamma1[pl] * (st[pl] * st[pl]) + 2.0 * P.alpha1[pl] * st[pl] * st[6 + pl] +
                                      P.beta1[pl] * (st[6 + pl] * st[6 + pl])
                                : 0.0;
        }
      }
      if (P.rw && phase != 3) {
        long long n = (long long)(P.Nmlt + 2) * P.h;
        long long k = (phase == 1 ? 0 : -rev * P.h) + kb;
        long long idx = ((k % n) + n) % n;
        R.rwhist[2 * idx] = st[1];
        R.rwhist[2 * idx + 1] = st[2];
        if (phase == 1)
          for (int jj = 1; jj < P.Nmlt; jj++) {
            long long i2 = ((long long)jj * P.h + kb) % n;
            R.rwhist[2 * i2] = st[1];
            R.rwhist[2 * i2 + 1] = st[2];
          }
      }
    }
  }
  if (phase == 0 || Nc <= 0) {
    // consume (and clear) the histogram only on the stats passes
    if (phase == 0 && Nc > 0) {
      for (int i = tid; i < Nc; i += WG) {
        size_t g = (size_t)b * Nc + i;
        H.cnt[g] = 0;
        H.sx[2 * g] = H.sx[2 * g + 1] = 0;
        H.sz[2 * g] = H.sz[2 * g + 1] = 0;
      }
      if (tid < 2) {
        H.tcnt[2 * b + tid] = 0;
        H.tsx[4 * b + 2 * tid] = H.tsx[4 * b + 2 * tid + 1] = 0;
        H.tsz[4 * b + 2 * tid] = H.tsz[4 * b + 2 * tid + 1] = 0;
      }
      if (tid == 0) H.bad[b] = 0;
    }
    return;
  }
  // ---- histograms: clamped (selffield) and dropped (LR resonators)
  const bool bad = H.bad[b] != 0;
  for (int i = tid; i < Nc; i += WG) {
    size_t g = (size_t)b * Nc + i;
    double cnt = (double)H.cnt[g];
    double sx = bad ? NAN : fx_to_double(H.sx[2 * g], H.sx[2 * g + 1]);
    double sz = bad ? NAN : fx_to_double(H.sz[2 * g], H.sz[2 * g + 1]);
    if (lr_L) R.fnp_ring[(size_t)kb * Nc + i] = cnt;
    if (lr_H) R.fnp_ring[(size_t)P.h * Nc + (size_t)kb * Nc + i] = sx;
    if (lr_V) R.fnp_ring[2 * (size_t)P.h * Nc + (size_t)kb * Nc + i] = sz;
    // clamped: out-of-range particles join the edge bins
    unsigned long long xl = H.sx[2 * g], xh = H.sx[2 * g + 1], zl = H.sz[2 * g], zh = H.sz[2 * g + 1];
    unsigned int c = H.cnt[g];
    if (i == 0 || i == Nc - 1) {
      int t = (i == 0) ? 0 : 1;
      // bin 0 takes the low trash, bin Nc-1 the high trash (both if Nc == 1)
      for (int tt = 0; tt < 2; tt++) {
        if (!((tt == 0 && i == 0) || (tt == 1 && i == Nc - 1))) continue;
        c += H.tcnt[2 * b + tt];
        unsigned long long al = H.tsx[4 * b + 2 * tt], ah = H.tsx[4 * b + 2 * tt + 1];
        unsigned long long nl = xl + al;
        xh = xh + ah + (nl < xl ? 1ULL : 0ULL);
        xl = nl;
        al = H.tsz[4 * b + 2 * tt];
        ah = H.tsz[4 * b + 2 * tt + 1];
        nl = zl + al;
        zh = zh + ah + (nl < zl ? 1ULL : 0ULL);
        zl = nl;
      }
      (void)t;
    }
    fnp[i] = (double)c;
    dV[i] = bad ? NAN : fx_to_double(zl, zh);
    dH[i] = bad ? NAN : fx_to_double(xl, xh);
    R.fnp_out[g] = fnp[i];
    R.dipV_out[g] = dV[i];
    R.dipH_out[g] = dH[i];
  }
  if (tid == 0) {
    icmin_s = Nc - 1;
    icmax_s = 0;
    // trash counters: the previous pending binning is consumed by this turn's kick
    if (phase == 2) {
      R.trash_tot[2 * b] += R.trash_pend[2 * b];
      R.trash_tot[2 * b + 1] += R.trash_pend[2 * b + 1];
    }
    R.trash_pend[2 * b] = (int)H.tcnt[2 * b];
    R.trash_pend[2 * b + 1] = (int)H.tcnt[2 * b + 1];
  }
  __syncthreads();
  // clear accumulators for the next pass
  for (int i = tid; i < Nc; i += WG) {
    size_t g = (size_t)b * Nc + i;
    H.cnt[g] = 0;
    H.sx[2 * g] = H.sx[2 * g + 1] = 0;
    H.sz[2 * g] = H.sz[2 * g + 1] = 0;
  }
  if (tid < 2) {
    H.tcnt[2 * b + tid] = 0;
    H.tsx[4 * b + 2 * tid] = H.tsx[4 * b + 2 * tid + 1] = 0;
    H.tsz[4 * b + 2 * tid] = H.tsz[4 * b + 2 * tid + 1] = 0;
  }
  if (tid == 0) {
    H.bad[b] = 0;
    // icellmin / icellmax searches of src/transform_weak.c:633-651
    int mn = Nc - 1, mx = 0;
    for (int i = 0; i < mn; i++)
      if (fnp[i] > 0) {
        mn = i;
        break;
      }
    for (int i = Nc - 1; i > mx; i--)
      if (fnp[i] > 0) {
        mx = i;
        break;
      }
    icmin_s = mn;
    icmax_s = mx;
  }
  __syncthreads();
  const int icmin = icmin_s, icmax = icmax_s;
  // the next turn's kick uses the current after this turn's scan step (src/tracking.c:43-59)
  double Ib = R.bunch_Ib[b];
  if (phase == 2 && P.scan == 1 && (rev + 1) % P.NrevScan == 0) Ib += P.scan_step / (((double)P.Nb) * 1.0e3);
  const double factG = -P.T0 * Ib / (R.np[b] * P.E0 * 1.0e9);
  const double sgm = P.sigma_tau, sgm2 = sgm * sgm, dT2 = P.dT * P.dT;
  for (int i = tid; i < Nc; i += WG) {
    double GL1 = 0.0, GV = 0.0, GH = 0.0;
    if (P.convL && i >= icmin) {
      double Gn = 0.0;
      for (int k = i; k >= icmin; k--) Gn += fnp[k] * R.Gl[i - k];
      GL1 = 1 * Gn;
      // resistive and inductive smoothing, src/transform_weak.c:670-683; reads
      // outside [0, Ncell) count as zero
#define F(q) (((q) >= 0 && (q) < Nc) ? fnp[q] : 0.0)
      if (P.RsisL > 0 && i >= icmin + 1 && i <= icmax) {
        double nt = (0.5 * F(i - 1) + F(i) + 0.5 * F(i + 1)) / 2.0;
        GL1 += P.RsisL * nt / (sgm * P.dT);
      }
      if (P.aindL > 0 && i >= icmin + 1 && i <= icmax) {
        double n0 = (0.25 * F(i - 3) + 0.5 * F(i - 2) + F(i - 1) + 0.5 * F(i) + 0.25 * F(i + 1)) / 2.5;
        double n1 = (0.25 * F(i - 1) + 0.5 * F(i) + F(i + 1) + 0.5 * F(i + 2) + 0.25 * F(i + 3)) / 2.5;
        GL1 += (n1 - n0) / (2.0 * sgm2 * dT2) * P.aindL;
      }
#undef F
    }
    if (P.kickV) {
      double Gn = 0.0;
      for (int k = i; k >= 0; k--) Gn += dV[k] * R.Gv[i - k];
      GV = 1 * Gn;
    }
    if (P.kickH && i >= 1) {
      double Gn = 0.0;
      for (int k = i; k >= 0; k--) Gn += dH[k] * R.Gh[i - k];
      GH = 1 * Gn;
    }
    R.sfL[(size_t)b * Nc + i] = factG * GL1;
    R.sfV[(size_t)b * Nc + i] = -factG * GV;
    R.sfH[(size_t)b * Nc + i] = -factG * GH;
  }
}

// ------------------------------------------------------------------- k_rw

__global__ __launch_bounds__(WG) void k_rw(DevParams P, const double *rwhist, const double *sqrt_tt, const double *Ibunch,
                                           const int *nfFill, const int *bucket, double *pending, const TurnState *ts)
{
  __shared__ double red[WG / 64][2];
  const int b = blockIdx.x, tid = threadIdx.x, kb = bucket[b];
  const long long rev = ts->rev;
  const int h = P.h;
  const long long n = (long long)(P.Nmlt + 2) * h;
  const long long M = (long long)(P.Nmlt + 1) * h;
  // index of tap m in the cyclic history: (-rev h + kb + m) mod n, bucket (kb + m) mod h,
  // advanced incrementally (stride WG) instead of by 64-bit divisions per tap
  const long long m0 = 1 + tid;
  long long k0 = (-rev * h + kb + m0) % n;
  long long idx = k0 < 0 ? k0 + n : k0;
  int in_mod = (int)((kb + m0) % h);
  const long long step_n = WG % n;
  const int step_h = WG % h;
  double ax = 0.0, az = 0.0;
  for (long long m = m0; m <= M; m += WG) {
    if (nfFill[in_mod]) {
      const double I = Ibunch[in_mod];
      const double st = sqrt_tt[m];
      ax += P.fNp * rwhist[2 * idx] * I / st;
      az += P.fNp * rwhist[2 * idx + 1] * I / st;
    }
    idx += step_n;
    if (idx >= n) idx -= n;
    in_mod += step_h;
    if (in_mod >= h) in_mod -= h;
  }
  ax = wave_sum(ax);
  az = wave_sum(az);
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = ax;
    red[tid >> 6][1] = az;
  }
  __syncthreads();
  if (tid == 0) {
    double sx = 0, sz = 0;
    for (int w = 0; w < WG / 64; w++) {
      sx += red[w][0];
      sz += red[w][1];
    }
    int trig = ts->m < 50 ? 1 : 0;
    pending[2 * b] = P.track_h ? sx * P.rwconst[0][trig] : 0.0;
    pending[2 * b + 1] = P.track_v ? sz * P.rwconst[1][trig] : 0.0;
  }
}

// Turn end: RW trigger counter, current scan (src/tracking.c:43-59), turn counter.
__global__ void k_end(DevParams P, TurnState *ts, double *bunch_Ib, double *Ibunch, const int *nfFill, int scan,
                      long long NrevScan, double scan_step)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  long long rev = ts->rev;
  if (P.rw) ts->m++;
  if ((rev + 1) % NrevScan == 0) {
    if (scan == 1) {
      ts->Iring += scan_step;
      ts->scan_val = ts->Iring;
      for (int k = 0; k < P.Nb; k++) bunch_Ib[k] += scan_step / (((double)P.Nb) * 1.0e3);
      for (int i = 0; i < P.h; i++)
        if (nfFill[i]) Ibunch[i] += scan_step / ((double)P.Nb);
    }
    ts->m = 0;
  }
  ts->rev = rev + 1;
}

// ------------------------------------------------------------------- k_lr

// One lane per long-range resonator: the recurrence of construct_wake_phasor
// (src/transform_weak.c:961-1048) or, with init_turns > 0, wake_phasor_init
// (:869-952).  Constants come from the host (glibc exp/cos/sin), the update
// order is the reference's, so the phasors are bit-identical.
__global__ void k_lr(int nres, const LRConst *lc, const double *fnp_ring, const double *Ibunch, const int *nfFill,
                     int h, int Nc, double *phasor_end, double *lrc, double *fb_hist, int fb_len, TurnState *ts,
                     int init_turns)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nres) return;
  const LRConst L = lc[r];
  const double *fnp = fnp_ring + (size_t)L.plane * h * Nc;
  double *out = lrc + (size_t)r * h * Nc;
  double Vo0, Vo1, Vn0, Vn1;
  if (init_turns) {
    Vo0 = 0.0;
    Vo1 = 0.0;
  } else {
    Vo0 = phasor_end[2 * r];
    Vo1 = phasor_end[2 * r + 1];
  }
  Vn0 = phasor_end[2 * r];
  Vn1 = phasor_end[2 * r + 1];
  const bool fb = L.fb && ts->has_fb;
  long long cnt = ts->fb_count;
  const int nturn = init_turns ? init_turns : 1;
  for (int k = 0; k < nturn; k++)
    for (int m = 0; m < h; m++) {
      int i = h - m - 1;
      if (fb) {
        double vr = Vn0 * L.pb0 - Vn1 * L.pb1, vi = Vn0 * L.pb1 + Vn1 * L.pb0;
        unsigned key = ((unsigned)cnt % (unsigned)fb_len) * 2;
        fb_hist[key] = sqrt(vr * vr + vi * vi);
        fb_hist[key + 1] = atan2(vr, vi);
        cnt++;
      }
      if (nfFill[i] == 1) {
        double Ii = Ibunch[i];
        const double *f = fnp + (size_t)i * Nc;
        double *o = out + (size_t)i * Nc;
        for (int j = 0; j < Nc; j++) {
          Vn0 = (Vo0 * L.pr0 - Vo1 * L.pr1) - Ii * f[j] * L.fac;
          Vn1 = (Vo0 * L.pr1 + Vo1 * L.pr0);
          if (!init_turns) o[j] = Vn0 + 0.5 * Ii * f[j] * L.fac;
          Vo0 = Vn0;
          Vo1 = Vn1;
        }
        Vn0 = (Vo0 * L.ge * L.gc - Vo1 * L.ge * L.gs);
        Vn1 = (Vo0 * L.ge * L.gs + Vo1 * L.ge * L.gc);
        Vo0 = Vn0;
        Vo1 = Vn1;
      } else {
        Vn0 = (Vo0 * L.ee * L.ec - Vo1 * L.ee * L.es);
        Vn1 = (Vo0 * L.ee * L.es + Vo1 * L.ee * L.ec);
        Vo0 = Vn0;
        Vo1 = Vn1;
      }
      if (fb) {
        double vr = Vn0 * L.pa0 - Vn1 * L.pa1, vi = Vn0 * L.pa1 + Vn1 * L.pa0;
        unsigned key = ((unsigned)cnt % (unsigned)fb_len) * 2;
        fb_hist[key] = sqrt(vr * vr + vi * vi);
        fb_hist[key + 1] = atan2(vr, vi);
        cnt++;
      }
    }
  phasor_end[2 * r] = Vn0;
  phasor_end[2 * r + 1] = Vn1;
  if (fb) ts->fb_count = cnt;
}

// lr_wake[plane][i][j] = sum over the plane's resonators in order of their terms
__global__ void k_lr_sum(int nres, const LRConst *lc, const double *lrc, const int *nfFill, int h, int Nc, double *lr_wake)
{
  long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long per = (long long)h * Nc;
  if (e >= 3 * per) return;
  int plane = (int)(e / per);
  long long ij = e % per;
  int i = (int)(ij / Nc);
  if (!nfFill[i]) {
    lr_wake[e] = 0.0;
    return;
  }
  double s = 0.0;
  for (int r = 0; r < nres; r++)
    if (lc[r].plane == plane) s += lrc[(size_t)r * per + ij];
  lr_wake[e] = s;
}

// RF feedback: rffb_calc_mean_voltage_phase (src/feedback_rf.c:21-34) and the
// solve of src/worker_weakweak.c:268-278.
__global__ void k_fb_solve(TurnState *ts, const double *fb_hist, int fb_len, double E0, double vrf_design,
                           double phi0_design, int switch_off)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  ts->phai0_prev = ts->phai0;
  ts->Vrf0_prev = ts->Vrf0;
  if (!ts->has_fb) return;
  double vs = 0, ps = 0;
  for (int m = 0; m < fb_len; m++) {
    vs += fb_hist[2 * m];
    ps += fb_hist[2 * m + 1];
  }
  double vmbar = vs / fb_len, phibar = ps / fb_len;
  vmbar = vmbar * E0 * 1.0e3;
  double num = vrf_design * sin(phi0_design) - vmbar * sin(phibar);
  double ph = atan2(num, vrf_design * cos(phi0_design) - vmbar * cos(phibar));
  ts->phai0 = ph;
  ts->Vrf0 = num / sin(ph);
  if (switch_off) ts->has_fb = 0;
}

__global__ void k_record_rf(const TurnState *ts, double *rfbuf)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int slot = (int)(ts->rev % STATS_TURNS);
    rfbuf[2 * slot] = ts->phai0;
    rfbuf[2 * slot + 1] = ts->Vrf0;
  }
}

// =============================================================== host side

struct mbt_ring {
  mbt_config cfg;
  int *nfFill_h, *bucket_h, *bunch_out_h;
  double *Ibunch_h, *bunch_Ib_h, *Ib_frac_h;
  int device;
  hipStream_t stream;
  DevParams P;
  int Nb, Nc, h;
  long long Ntot;
  int nwg;
  Coords C;
  long long *wg_first, *wg_end;
  int *wg_bunch, *bunch_wg0, *bucket, *np, *nfFill;
  double *partial, *shift;
  HistPtr H;
  RedOut R;
  double *sfL, *sfV, *sfH, *lr_wake, *fnp_ring, *lrc;
  double *Gl, *Gv, *Gh;
  double *pending, *rwhist, *sqrt_tt, *Ibunch, *bunch_Ib;
  double *stats, *amp, *rfbuf;
  double *fnp_out, *dipV_out, *dipH_out;
  int *trash_tot, *trash_pend;
  TurnState *ts;
  LRConst *lc;
  int nres;
  double *phasor_end, *fb_hist;
  int fb_len, fb_off;
  int lr_plane[3];
  size_t lds_turn, lds_red;
  long long rev_host;
  hipEvent_t ev[8];
  hipEvent_t *evpool; // 5 per turn of a stats chunk when timing
  double kt_ms[4];
  long kt_launches[4];
  int timing;
};

static size_t turn_lds(int Nc)
{
  return (size_t)(4 * Nc + 8) * 8 + (size_t)3 * Nc * 8 + (size_t)(Nc + 2) * 4 + 16;
}

static void fill_params(mbt_ring *r)
{
  const mbt_config *c = &r->cfg;
  DevParams &P = r->P;
  memset(&P, 0, sizeof P);
  P.Nb = c->Nbunch;
  P.h = (int)c->Nharm;
  P.Nc = c->Ncell;
  P.Nmlt = c->Nmlt;
  P.track_h = c->TrackPlane[1] != 0;
  P.track_v = c->TrackPlane[2] != 0;
  P.qe = c->EnableQuantum != 0;
  P.ideal_hc = c->EnableIdealHC != 0;
  P.amp_dq = c->AmpDependdQV != 0;
  P.rw = c->EnableRW_long != 0;
  P.n_ahc = c->n_ahc;
  P.monitor_every = (int)c->NrevMon;
  P.seed = (unsigned)c->iseed[0];
  for (int a = 0; a < c->n_ahc; a++) {
    P.ahc_amp[a] = c->ahc[a].Vpeak / (c->E0 * MBT_FGIGA);
    P.ahc_wn[a] = c->wrf * c->ahc[a].nHC;
    P.ahc_phi[a] = c->ahc[a].phi_aHC;
    P.ahc_nHC[a] = c->ahc[a].nHC;
  }
  P.E0 = c->E0;
  P.E0k = c->E0 * MBT_FKILO;
  P.RsisL = c->RsisL;
  P.aindL = c->aindL;
  P.Np = c->Np;
  P.scan = c->scan;
  P.NrevScan = c->NrevScan;
  P.scan_step = c->scan_step;
  P.wrf = c->wrf;
  P.Urad = c->Urad;
  P.damp = 1.0 - c->De;
  P.excite = 2.0 * c->slope_sgm[0] * sqrt(c->T0 / c->taue);
  P.qVp = c->pos_sgm[2] * sqrt(2.0 * c->T0 / c->tauz);
  P.qVs = c->slope_sgm[2] * sqrt(2.0 * c->T0 / c->tauz);
  P.qHp = c->pos_sgm[1] * sqrt(2.0 * c->T0 / c->taux);
  P.qHs = c->slope_sgm[1] * sqrt(2.0 * c->T0 / c->taux);
  P.T0 = c->T0;
  P.ac = c->ac;
  P.PsiV0 = 2.0 * M_PI * c->QV0;
  P.PsiH0 = 2.0 * M_PI * c->QH0;
  P.Gziz = c->Gziz;
  P.Gzix = c->Gzix;
  P.aV = c->alpha1[2];
  P.bV = c->beta1[2];
  P.gV = c->gamma1[2];
  P.aH = c->alpha1[1];
  P.bH = c->beta1[1];
  P.gH = c->gamma1[1];
  P.dispH = c->dispH1;
  P.disppH = c->disppH1;
  P.CVH = c->CVH;
  P.CVV = c->CVV;
  P.CHH = c->CHH;
  P.CHV = c->CHV;
  P.HC_k = c->HC_k;
  P.m_wrf = c->m_aHC * c->wrf;
  P.m_phi = c->m_aHC * c->phi_n;
  P.sigma_tau = c->sigma_tau;
  P.Nsigma = c->Nsigma;
  P.dT = c->dT;
  for (int p = 0; p < 3; p++) {
    P.gamma1[p] = c->gamma1[p];
    P.alpha1[p] = c->alpha1[p];
    P.beta1[p] = c->beta1[p];
  }
  P.fNp = c->fNp;
  for (int p = 0; p < 2; p++)
    for (int t = 0; t < 2; t++) {
      double beff3 = p == 0 ? c->beffH[0] : c->beffV[0];
      if (t && c->triggerRW) beff3 /= 2;
      P.rwconst[p][t] = c->T0 / c->E0 / c->fNp / MBT_FTERA * c->Lc / (M_PI * pow(beff3, 3)) *
                        sqrt(MBT_Z0 * MBT_C_LIGHT * c->rhorw / M_PI);
    }
}

template <typename T> static int dalloc(T **p, size_t n)
{
  if (n == 0) n = 1;
  if (hipMalloc((void **)p, n * sizeof(T)) != hipSuccess) {
    mbt_set_error("hipMalloc of %zu bytes failed", n * sizeof(T));
    return MBT_ENOMEM;
  }
  if (hipMemset(*p, 0, n * sizeof(T)) != hipSuccess) return MBT_EDEVICE;
  return MBT_OK;
}

#define DA(p, n)                                                                                 \
  do {                                                                                           \
    int rc_ = dalloc(&(p), (size_t)(n));                                                         \
    if (rc_) {                                                                                   \
      mbt_ring_destroy(r);                                                                       \
      return rc_;                                                                                \
    }                                                                                            \
  } while (0)

static void copy_cfg(mbt_ring *r, const mbt_config *c)
{
  r->cfg = *c;
  int H = (int)c->Nharm, Nb = c->Nbunch;
  r->nfFill_h = (int *)malloc(H * sizeof(int));
  memcpy(r->nfFill_h, c->nfFill, H * sizeof(int));
  r->Ibunch_h = (double *)malloc(H * sizeof(double));
  memcpy(r->Ibunch_h, c->Ibunch, H * sizeof(double));
  r->bucket_h = (int *)malloc(Nb * sizeof(int));
  memcpy(r->bucket_h, c->bucket, Nb * sizeof(int));
  r->bunch_Ib_h = (double *)malloc(Nb * sizeof(double));
  memcpy(r->bunch_Ib_h, c->bunch_Ib, Nb * sizeof(double));
  r->Ib_frac_h = (double *)malloc(Nb * sizeof(double));
  memcpy(r->Ib_frac_h, c->Ib_frac, Nb * sizeof(double));
  r->bunch_out_h = (int *)malloc((c->Nbunch_out > 0 ? c->Nbunch_out : 1) * sizeof(int));
  if (c->Nbunch_out > 0) memcpy(r->bunch_out_h, c->bunch_out, c->Nbunch_out * sizeof(int));
  r->cfg.nfFill = r->nfFill_h;
  r->cfg.Ibunch = r->Ibunch_h;
  r->cfg.bucket = r->bucket_h;
  r->cfg.bunch_Ib = r->bunch_Ib_h;
  r->cfg.Ib_frac = r->Ib_frac_h;
  r->cfg.bunch_out = r->bunch_out_h;
}

extern "C" void mbt_ring_destroy(mbt_ring *r)
{
  if (!r) return;
  hipSetDevice(r->device);
  void *ptrs[] = {r->C.tau, r->C.del, r->C.x, r->C.xp, r->C.z, r->C.zp, r->wg_first, r->wg_end, r->wg_bunch,
                  r->bunch_wg0, r->bucket, r->np, r->nfFill, r->partial, r->shift, r->H.cnt, r->H.sx, r->H.sz,
                  r->H.tcnt, r->H.tsx, r->H.tsz, r->H.bad, r->sfL, r->sfV, r->sfH, r->lr_wake, r->fnp_ring, r->lrc,
                  r->Gl, r->Gv, r->Gh, r->pending, r->rwhist, r->sqrt_tt, r->Ibunch, r->bunch_Ib, r->stats, r->amp,
                  r->rfbuf, r->fnp_out, r->dipV_out, r->dipH_out, r->trash_tot, r->trash_pend, r->ts, r->lc,
                  r->phasor_end, r->fb_hist};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
    if (ptrs[i]) hipFree(ptrs[i]);
  for (int i = 0; i < 8; i++)
    if (r->ev[i]) hipEventDestroy(r->ev[i]);
  if (r->evpool) {
    for (int i = 0; i < 6 * STATS_TURNS; i++) hipEventDestroy(r->evpool[i]);
    free(r->evpool);
  }
  if (r->stream) hipStreamDestroy(r->stream);
  free(r->nfFill_h);
  free(r->Ibunch_h);
  free(r->bucket_h);
  free(r->bunch_Ib_h);
  free(r->Ib_frac_h);
  free(r->bunch_out_h);
  free(r);
}

extern "C" int mbt_ring_create(const mbt_config *c, int device, int flags, mbt_ring **out)
{
  *out = NULL;
  if (c->bunch_model != 0) {
    mbt_set_error("BunchModel STRONG is handled by the weak-strong driver, not the macroparticle ring");
    return MBT_EUNSUPPORTED;
  }
  if (c->Nbunch <= 0 || c->Np <= 0) {
    mbt_set_error("empty ring (Nbunch=%d Np=%d)", c->Nbunch, c->Np);
    return MBT_EINVAL;
  }
  if (c->NrevMon <= 0) {
    mbt_set_error("NrevMon must be positive");
    return MBT_EINVAL;
  }
  if (c->has_rf_feedback && c->n_lr[0] == 0) {
    mbt_set_error("[rf_feedback] needs a longitudinal long-range resonator");
    return MBT_EINVAL;
  }
  if ((c->n_lr[1] > 0 && !c->TrackPlane[1]) || (c->n_lr[2] > 0 && !c->TrackPlane[2])) {
    mbt_set_error("transverse long-range resonators need their plane tracked");
    return MBT_EINVAL;
  }
  if (c->has_rf_feedback && (c->rffb_len_average == 0)) {
    mbt_set_error("averaging_length must be nonzero");
    return MBT_EINVAL;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
    mbt_set_error("no HIP device %d", device);
    return MBT_EDEVICE;
  }
  mbt_ring *r = (mbt_ring *)calloc(1, sizeof(mbt_ring));
  copy_cfg(r, c);
  r->device = device;
  CK(hipSetDevice(device));
  CK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
  for (int i = 0; i < 8; i++) CK(hipEventCreate(&r->ev[i]));
  fill_params(r);
  const mbt_config *cf = &r->cfg;
  const int Nb = cf->Nbunch, Nc = cf->Ncell, h = (int)cf->Nharm;
  r->Nb = Nb;
  r->Nc = Nc;
  r->h = h;
  const long long Np = cf->Np;
  r->Ntot = Np * Nb;
  // ---- green functions
  double *Gl = (double *)calloc(Nc > 0 ? Nc : 1, sizeof(double));
  double *Gv = (double *)calloc(Nc > 0 ? Nc : 1, sizeof(double));
  double *Gh = (double *)calloc(Nc > 0 ? Nc : 1, sizeof(double));
  int planes[3];
  int rc = mbt_greens_functions(cf, Gl, Gv, Gh, planes);
  if (rc) {
    free(Gl);
    free(Gv);
    free(Gh);
    mbt_ring_destroy(r);
    return rc;
  }
  int lr_tot = cf->n_lr[0] + cf->n_lr[1] + cf->n_lr[2];
  r->P.sf_active = (planes[0] + planes[1] + planes[2] + lr_tot) > 0 && Nc > 0;
  r->P.kickL = (cf->n_lr[0] + planes[0]) > 0;
  r->P.convL = planes[0] > 0;
  r->P.kickV = planes[2] > 0;
  r->P.kickH = planes[1] > 0;
  r->lr_plane[0] = cf->n_lr[0] > 0;
  r->lr_plane[1] = cf->n_lr[1] > 0 && cf->TrackPlane[1];
  r->lr_plane[2] = cf->n_lr[2] > 0 && cf->TrackPlane[2];
  // ---- workgroup decomposition: contiguous chunks of one bunch
  const long long CH = (long long)WG * CHUNK_ITEMS;
  const long long per_b = (Np + CH - 1) / CH;
  r->nwg = (int)(per_b * Nb);
  long long *wf = (long long *)malloc(r->nwg * sizeof(long long)), *we = (long long *)malloc(r->nwg * sizeof(long long));
  int *wb = (int *)malloc(r->nwg * sizeof(int)), *b0 = (int *)malloc((Nb + 1) * sizeof(int));
  int *nph = (int *)malloc(Nb * sizeof(int));
  for (int b = 0, w = 0; b < Nb; b++) {
    b0[b] = w;
    nph[b] = (int)Np;
    for (long long k = 0; k < per_b; k++, w++) {
      wf[w] = b * Np + k * CH;
      we[w] = b * Np + ((k + 1) * CH < Np ? (k + 1) * CH : Np);
      wb[w] = b;
    }
  }
  b0[Nb] = r->nwg;
  DA(r->C.tau, r->Ntot);
  DA(r->C.del, r->Ntot);
  DA(r->C.x, r->Ntot);
  DA(r->C.xp, r->Ntot);
  DA(r->C.z, r->Ntot);
  DA(r->C.zp, r->Ntot);
  DA(r->wg_first, r->nwg);
  DA(r->wg_end, r->nwg);
  DA(r->wg_bunch, r->nwg);
  DA(r->bunch_wg0, Nb + 1);
  DA(r->bucket, Nb);
  DA(r->np, Nb);
  DA(r->nfFill, h);
  DA(r->partial, (size_t)r->nwg * NMOM);
  DA(r->shift, (size_t)Nb * 6);
  int NcA = Nc > 0 ? Nc : 1;
  DA(r->H.cnt, (size_t)Nb * NcA);
  DA(r->H.sx, (size_t)Nb * NcA * 2);
  DA(r->H.sz, (size_t)Nb * NcA * 2);
  DA(r->H.tcnt, (size_t)Nb * 2);
  DA(r->H.tsx, (size_t)Nb * 4);
  DA(r->H.tsz, (size_t)Nb * 4);
  DA(r->H.bad, Nb);
  DA(r->sfL, (size_t)Nb * NcA);
  DA(r->sfV, (size_t)Nb * NcA);
  DA(r->sfH, (size_t)Nb * NcA);
  DA(r->lr_wake, (size_t)3 * h * NcA);
  DA(r->fnp_ring, (size_t)3 * h * NcA);
  DA(r->Gl, NcA);
  DA(r->Gv, NcA);
  DA(r->Gh, NcA);
  DA(r->pending, (size_t)Nb * 2);
  long long M = (long long)(cf->Nmlt + 1) * h;
  DA(r->rwhist, (size_t)(cf->EnableRW_long ? (cf->Nmlt + 2) * h * 2 : 2));
  DA(r->sqrt_tt, (size_t)(cf->EnableRW_long ? M + 1 : 1));
  DA(r->Ibunch, h);
  DA(r->bunch_Ib, Nb);
  DA(r->stats, (size_t)STATS_TURNS * Nb * 12);
  DA(r->amp, (size_t)STATS_TURNS * Nb * 6);
  DA(r->rfbuf, (size_t)STATS_TURNS * 2);
  DA(r->fnp_out, (size_t)Nb * NcA);
  DA(r->dipV_out, (size_t)Nb * NcA);
  DA(r->dipH_out, (size_t)Nb * NcA);
  DA(r->trash_tot, (size_t)Nb * 2);
  DA(r->trash_pend, (size_t)Nb * 2);
  DA(r->ts, 1);
  // ---- long-range resonator constants (host libm, reference formulas)
  LRConst lc[MAXLR];
  int nres = 0;
  for (int p = 0; p < 3; p++) {
    if (!r->lr_plane[p]) continue;
    for (int l = 0; l < cf->n_lr[p]; l++) {
      const mbt_lr_resonator *res = &cf->lr[p][l];
      LRConst &L = lc[nres++];
      L.plane = p;
      L.l = l;
      double alpha = res->wr * 0.5 / res->Qfactor;
      double Amp = 2 * alpha * res->Rs * cf->T0 / cf->E0 / MBT_FGIGA;
      double dTau = cf->dT * cf->sigma_tau;
      double tbucket = (cf->T0 / cf->h - Nc * dTau);
      L.fac = Amp / (unsigned)cf->Np * MBT_FMILLI;
      double C0 = -alpha, C1 = res->wr;
      L.pr0 = exp(C0 * dTau) * cos(C1 * dTau);
      L.pr1 = exp(C0 * dTau) * sin(C1 * dTau);
      L.ge = exp(C0 * tbucket);
      L.gc = cos(C1 * tbucket);
      L.gs = sin(C1 * tbucket);
      double tb = cf->T0 / cf->h;
      L.ee = exp(C0 * tb);
      L.ec = cos(C1 * tb);
      L.es = sin(C1 * tb);
      L.fb = cf->has_rf_feedback && p == 0 && cf->rffb_resonator == l + 1;
      double taubeam = 0, taub2beam = 0;
      if (L.fb) {
        taubeam = cf->Nsigma * cf->sigma_tau + dTau * 1.5;
        taub2beam = -cf->Nsigma * cf->sigma_tau + dTau * 0.5 - tbucket;
      }
      L.pb0 = exp(C0 * taubeam) * cos(C1 * taubeam);
      L.pb1 = exp(C0 * taubeam) * sin(C1 * taubeam);
      L.pa0 = exp(C0 * taub2beam) * cos(C1 * taub2beam);
      L.pa1 = exp(C0 * taub2beam) * sin(C1 * taub2beam);
    }
  }
  r->nres = nres;
  DA(r->lc, nres > 0 ? nres : 1);
  DA(r->lrc, (size_t)(nres > 0 ? nres : 1) * h * NcA);
  DA(r->phasor_end, (size_t)2 * (nres > 0 ? nres : 1));
  r->fb_len = cf->has_rf_feedback ? abs(cf->rffb_len_average) : 1;
  r->fb_off = cf->has_rf_feedback && cf->rffb_len_average < 0;
  DA(r->fb_hist, (size_t)2 * r->fb_len);
  // ---- uploads
  CK(hipMemcpy(r->wg_first, wf, r->nwg * sizeof(long long), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->wg_end, we, r->nwg * sizeof(long long), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->wg_bunch, wb, r->nwg * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->bunch_wg0, b0, (Nb + 1) * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->np, nph, Nb * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->bucket, cf->bucket, Nb * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->nfFill, cf->nfFill, h * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->Ibunch, cf->Ibunch, h * sizeof(double), hipMemcpyHostToDevice));
  CK(hipMemcpy(r->bunch_Ib, cf->bunch_Ib, Nb * sizeof(double), hipMemcpyHostToDevice));
  if (Nc > 0) {
    CK(hipMemcpy(r->Gl, Gl, Nc * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(r->Gv, Gv, Nc * sizeof(double), hipMemcpyHostToDevice));
    CK(hipMemcpy(r->Gh, Gh, Nc * sizeof(double), hipMemcpyHostToDevice));
  }
  if (nres > 0) CK(hipMemcpy(r->lc, lc, nres * sizeof(LRConst), hipMemcpyHostToDevice));
  if (cf->EnableRW_long) {
    double *st = (double *)malloc((M + 1) * sizeof(double));
    double deltab = cf->T0 / cf->h;
    st[0] = 1.0;
    for (long long m = 1; m <= M; m++) st[m] = sqrt(m * deltab);
    CK(hipMemcpy(r->sqrt_tt, st, (M + 1) * sizeof(double), hipMemcpyHostToDevice));
    free(st);
  }
  TurnState ts0;
  memset(&ts0, 0, sizeof ts0);
  ts0.phai0 = cf->phai0;
  ts0.Vrf0 = cf->Vrf0;
  ts0.Iring = cf->Iring;
  ts0.scan_val = cf->scan_start;
  CK(hipMemcpy(r->ts, &ts0, sizeof ts0, hipMemcpyHostToDevice));
  free(wf);
  free(we);
  free(wb);
  free(b0);
  free(nph);
  free(Gl);
  free(Gv);
  free(Gh);
  r->R.stats = r->stats;
  r->R.amp = r->amp;
  r->R.shift = r->shift;
  r->R.rwhist = r->rwhist;
  r->R.fnp_ring = r->fnp_ring;
  r->R.sfL = r->sfL;
  r->R.sfV = r->sfV;
  r->R.sfH = r->sfH;
  r->R.fnp_out = r->fnp_out;
  r->R.dipV_out = r->dipV_out;
  r->R.dipH_out = r->dipH_out;
  r->R.trash_tot = r->trash_tot;
  r->R.trash_pend = r->trash_pend;
  r->R.Gl = r->Gl;
  r->R.Gv = r->Gv;
  r->R.Gh = r->Gh;
  r->R.bunch_Ib = r->bunch_Ib;
  r->R.np = r->np;
  r->lds_turn = turn_lds(Nc);
  r->lds_red = (size_t)3 * NcA * 8 + 64;
  if (r->lds_turn > 64 * 1024) {
    mbt_set_error("Ncell %d too large for the LDS histogram", Nc);
    mbt_ring_destroy(r);
    return MBT_EINVAL;
  }
  // ---- initial distributions (host threads, one glibc stream per bunch)
  if (!(flags & MBT_RING_NO_INIT_DIST)) {
    double *buf = (double *)malloc(sizeof(double) * 6 * (size_t)Np * Nb);
    if (!buf) {
      mbt_ring_destroy(r);
      return MBT_ENOMEM;
    }
    rc = mbt_generate_all(cf, buf);
    for (int b = 0; b < Nb && rc == MBT_OK; b++) rc = mbt_ring_set_bunch(r, b, buf + (size_t)6 * Np * b);
    free(buf);
    if (rc) {
      mbt_ring_destroy(r);
      return rc;
    }
  }
  *out = r;
  return MBT_OK;
}

extern "C" int mbt_ring_set_bunch(mbt_ring *r, int ib, const double *s)
{
  if (ib < 0 || ib >= r->Nb) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  const long long Np = r->cfg.Np;
  double *dst[6] = {r->C.tau, r->C.del, r->C.x, r->C.xp, r->C.z, r->C.zp};
  for (int c = 0; c < 6; c++)
    CK(hipMemcpy(dst[c] + ib * Np, s + c * Np, Np * sizeof(double), hipMemcpyHostToDevice));
  return MBT_OK;
}

extern "C" int mbt_ring_get_bunch(const mbt_ring *r, int ib, double *s)
{
  if (ib < 0 || ib >= r->Nb) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  const long long Np = r->cfg.Np;
  double *src[6] = {r->C.tau, r->C.del, r->C.x, r->C.xp, r->C.z, r->C.zp};
  for (int c = 0; c < 6; c++)
    CK(hipMemcpy(s + c * Np, src[c] + ib * Np, Np * sizeof(double), hipMemcpyDeviceToHost));
  return MBT_OK;
}

typedef void (*turn_kernel_t)(DevParams, Coords, const long long *, const long long *, const int *, const int *,
                              const double *, const double *, const double *, const double *, const double *,
                              const double *, const double *, const double *, double *, HistPtr, const TurnState *);

#define TK(Q, D, G, M) k_turn<MODE_TURN, Q, D, G, M>
static const turn_kernel_t g_turn_kernels[16] = {
    TK(false, false, false, false), TK(false, false, false, true), TK(false, false, true, false),
    TK(false, false, true, true),   TK(false, true, false, false), TK(false, true, false, true),
    TK(false, true, true, false),   TK(false, true, true, true),   TK(true, false, false, false),
    TK(true, false, false, true),   TK(true, false, true, false),  TK(true, false, true, true),
    TK(true, true, false, false),   TK(true, true, false, true),   TK(true, true, true, false),
    TK(true, true, true, true)};
#undef TK

static int launch_turn_pass(mbt_ring *r, int mode)
{
  turn_kernel_t k;
  if (mode == MODE_STATS) k = k_turn<MODE_STATS, false, true, true, false>;
  else if (mode == MODE_FINAL) k = k_turn<MODE_FINAL, false, true, true, false>;
  else {
    const bool mon = ((r->rev_host + 1) % r->cfg.NrevMon) == 0;
    const bool dip = r->P.kickV || r->P.kickH || r->lr_plane[1] || r->lr_plane[2];
    const bool gen = r->P.ideal_hc || r->P.n_ahc > 0 || r->P.amp_dq;
    k = g_turn_kernels[(r->P.qe ? 8 : 0) + (dip ? 4 : 0) + (gen ? 2 : 0) + (mon ? 1 : 0)];
  }
  const size_t NcA = r->Nc > 0 ? r->Nc : 1;
  hipLaunchKernelGGL(k, dim3(r->nwg), dim3(WG), r->lds_turn, r->stream, r->P, r->C, r->wg_first, r->wg_end,
                     r->wg_bunch, r->bucket, r->shift, r->sfL, r->sfV, r->sfH, r->lr_wake,
                     r->lr_wake + (size_t)r->h * NcA, r->lr_wake + (size_t)2 * r->h * NcA, r->pending, r->partial,
                     r->H, r->ts);
  CK(hipGetLastError());
  return MBT_OK;
}

static int launch_reduce(mbt_ring *r, int phase)
{
  hipLaunchKernelGGL(k_reduce, dim3(r->Nb), dim3(WG), r->lds_red, r->stream, r->P, r->partial, r->bunch_wg0, r->bucket,
                     r->H, r->R, r->ts, phase, r->lr_plane[0], r->lr_plane[1], r->lr_plane[2]);
  CK(hipGetLastError());
  return MBT_OK;
}

static int launch_lr(mbt_ring *r, int init_turns)
{
  if (r->nres == 0) return MBT_OK;
  hipLaunchKernelGGL(k_lr, dim3(1), dim3(64), 0, r->stream, r->nres, r->lc, r->fnp_ring, r->Ibunch, r->nfFill, r->h,
                     r->Nc, r->phasor_end, r->lrc, r->fb_hist, r->fb_len, r->ts, init_turns);
  CK(hipGetLastError());
  if (!init_turns) {
    long long tot = 3LL * r->h * r->Nc;
    hipLaunchKernelGGL(k_lr_sum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, r->stream, r->nres, r->lc, r->lrc,
                       r->nfFill, r->h, r->Nc, r->lr_wake);
    CK(hipGetLastError());
  }
  return MBT_OK;
}

static int launch_fb_solve(mbt_ring *r, int switch_off)
{
  if (!r->cfg.has_rf_feedback || r->nres == 0) return MBT_OK;
  hipLaunchKernelGGL(k_fb_solve, dim3(1), dim3(64), 0, r->stream, r->ts, r->fb_hist, r->fb_len, r->cfg.E0,
                     r->cfg.rffb_vrf_design, r->cfg.rffb_phi0_design, switch_off);
  CK(hipGetLastError());
  return MBT_OK;
}

// Long-range recurrence and RF-feedback solve that precede a turn.
static int pre_turn(mbt_ring *r)
{
  int rc;
  if (r->P.sf_active && (rc = launch_lr(r, 0))) return rc;
  hipLaunchKernelGGL(k_fb_solve, dim3(1), dim3(64), 0, r->stream, r->ts, r->fb_hist, r->fb_len, r->cfg.E0,
                     r->cfg.rffb_vrf_design, r->cfg.rffb_phi0_design, 0);
  CK(hipGetLastError());
  return MBT_OK;
}

static int download_slots(mbt_ring *r, long long rev0, long n, double *stats, double *ampinv, double *rf)
{
  CK(hipStreamSynchronize(r->stream));
  const int Nb = r->Nb;
  for (long t = 0; t < n; t++) {
    int slot = (int)((rev0 + t) % STATS_TURNS);
    if (stats)
      CK(hipMemcpy(stats + (size_t)t * Nb * 12, r->stats + (size_t)slot * Nb * 12, Nb * 12 * sizeof(double),
                   hipMemcpyDeviceToHost));
    if (ampinv && ((rev0 + t + 1) % r->cfg.NrevMon == 0))
      CK(hipMemcpy(ampinv + (size_t)t * Nb * 6, r->amp + (size_t)slot * Nb * 6, Nb * 6 * sizeof(double),
                   hipMemcpyDeviceToHost));
    if (rf) CK(hipMemcpy(rf + 2 * t, r->rfbuf + 2 * slot, 2 * sizeof(double), hipMemcpyDeviceToHost));
  }
  return MBT_OK;
}

extern "C" int mbt_ring_prepare(mbt_ring *r, double *init_stats)
{
  CK(hipSetDevice(r->device));
  int rc;
  // two stats passes: means (shift 0), then moments about the means + histograms
  CK(hipMemsetAsync(r->shift, 0, (size_t)r->Nb * 6 * sizeof(double), r->stream));
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  if ((rc = launch_reduce(r, 0))) return rc;
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  // the initial statistics land in slot (rev=0) and the RW history rows 0..Nmlt-1
  if ((rc = launch_reduce(r, 1))) return rc;
  if (init_stats) {
    CK(hipStreamSynchronize(r->stream));
    CK(hipMemcpy(init_stats, r->stats, (size_t)r->Nb * 12 * sizeof(double), hipMemcpyDeviceToHost));
  }
  // LR resonators: wake_phasor_init replays, then the RF-feedback design solve
  if (r->nres > 0) {
    TurnState ts;
    CK(hipStreamSynchronize(r->stream));
    CK(hipMemcpy(&ts, r->ts, sizeof ts, hipMemcpyDeviceToHost));
    ts.has_fb = r->cfg.has_rf_feedback ? 1 : 0;
    ts.fb_count = 0;
    CK(hipMemcpy(r->ts, &ts, sizeof ts, hipMemcpyHostToDevice));
    int nturns = (int)(r->cfg.Nbumax / r->cfg.h);
    if ((rc = launch_lr(r, nturns > 0 ? nturns : 0))) return rc;
    if (nturns <= 0) {
      // no replay: phasors stay at zero (Nbumax < h)
    }
    if ((rc = launch_fb_solve(r, r->fb_off))) return rc;
  }
  if ((rc = pre_turn(r))) return rc;
  CK(hipStreamSynchronize(r->stream));
  r->rev_host = 0;
  return MBT_OK;
}

extern "C" int mbt_ring_track(mbt_ring *r, long nturns, double *stats, double *ampinv, double *rf)
{
  CK(hipSetDevice(r->device));
  int rc;
  for (int i = 0; i < 4; i++) {
    r->kt_ms[i] = 0;
    r->kt_launches[i] = 0;
  }
  long done = 0;
  while (done < nturns) {
    long n = nturns - done;
    if (n > STATS_TURNS) n = STATS_TURNS;
    long long rev0 = r->rev_host;
    const bool timed = r->timing != 0;
    for (long t = 0; t < n; t++) {
      hipEvent_t *E = timed ? r->evpool + 6 * t : NULL;
      hipLaunchKernelGGL(k_record_rf, dim3(1), dim3(64), 0, r->stream, r->ts, r->rfbuf);
      if (timed) CK(hipEventRecord(E[1], r->stream));
      if ((rc = launch_turn_pass(r, MODE_TURN))) return rc;
      if (timed) CK(hipEventRecord(E[2], r->stream));
      if ((rc = launch_reduce(r, 2))) return rc;
      if (timed) CK(hipEventRecord(E[3], r->stream));
      if (r->P.rw) {
        hipLaunchKernelGGL(k_rw, dim3(r->Nb), dim3(WG), 0, r->stream, r->P, r->rwhist, r->sqrt_tt, r->Ibunch,
                           r->nfFill, r->bucket, r->pending, r->ts);
        CK(hipGetLastError());
      }
      if (timed) CK(hipEventRecord(E[4], r->stream));
      hipLaunchKernelGGL(k_end, dim3(1), dim3(64), 0, r->stream, r->P, r->ts, r->bunch_Ib, r->Ibunch, r->nfFill,
                         r->cfg.scan, (long long)r->cfg.NrevScan, r->cfg.scan_step);
      CK(hipGetLastError());
      // the next turn's long-range wakes and RF-feedback solve (the reference runs
      // them at the start of the turn; nothing changes the inputs in between)
      if (timed) CK(hipEventRecord(E[0], r->stream));
      if ((rc = pre_turn(r))) return rc;
      if (timed) CK(hipEventRecord(E[5], r->stream));
      r->rev_host++;
    }
    if (timed) {
      CK(hipStreamSynchronize(r->stream));
      for (long t = 0; t < n; t++) {
        hipEvent_t *E = r->evpool + 6 * t;
        float a, b, c, d;
        CK(hipEventElapsedTime(&a, E[0], E[5]));
        CK(hipEventElapsedTime(&b, E[1], E[2]));
        CK(hipEventElapsedTime(&c, E[2], E[3]));
        CK(hipEventElapsedTime(&d, E[3], E[4]));
        (void)0;
        r->kt_ms[2] += a;
        r->kt_ms[0] += b;
        r->kt_ms[1] += c;
        r->kt_ms[3] += d;
        for (int i = 0; i < 4; i++) r->kt_launches[i]++;
      }
    }
    rc = download_slots(r, rev0, n, stats ? stats + (size_t)done * r->Nb * 12 : NULL,
                        ampinv ? ampinv + (size_t)done * r->Nb * 6 : NULL, rf ? rf + 2 * done : NULL);
    if (rc) return rc;
    done += n;
  }
  CK(hipStreamSynchronize(r->stream));
  return MBT_OK;
}

extern "C" int mbt_ring_final_stats(mbt_ring *r, double *stats)
{
  CK(hipSetDevice(r->device));
  int rc;
  // apply the pending RW kick, then two exact passes for the statistics
  if ((rc = launch_turn_pass(r, MODE_FINAL))) return rc;
  CK(hipMemsetAsync(r->pending, 0, (size_t)r->Nb * 2 * sizeof(double), r->stream));
  CK(hipMemsetAsync(r->shift, 0, (size_t)r->Nb * 6 * sizeof(double), r->stream));
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  if ((rc = launch_reduce(r, 0))) return rc;
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  // stats land in the slot of the current turn counter; histograms stay for the next turn
  if ((rc = launch_reduce(r, 3))) return rc;
  CK(hipStreamSynchronize(r->stream));
  TurnState ts;
  CK(hipMemcpy(&ts, r->ts, sizeof ts, hipMemcpyDeviceToHost));
  int slot = (int)(ts.rev % STATS_TURNS);
  CK(hipMemcpy(stats, r->stats + (size_t)slot * r->Nb * 12, (size_t)r->Nb * 12 * sizeof(double),
               hipMemcpyDeviceToHost));
  return MBT_OK;
}

extern "C" int mbt_ring_histograms(const mbt_ring *r, double *counts, int *trash)
{
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  if (r->Nc > 0) CK(hipMemcpy(counts, r->fnp_out, (size_t)r->Nb * r->Nc * sizeof(double), hipMemcpyDeviceToHost));
  int *tt = (int *)malloc((size_t)r->Nb * 2 * sizeof(int)), *tp = (int *)malloc((size_t)r->Nb * 2 * sizeof(int));
  hipMemcpy(tt, r->trash_tot, (size_t)r->Nb * 2 * sizeof(int), hipMemcpyDeviceToHost);
  hipMemcpy(tp, r->trash_pend, (size_t)r->Nb * 2 * sizeof(int), hipMemcpyDeviceToHost);
  for (int i = 0; i < 2 * r->Nb; i++) trash[i] = tt[i] + (r->P.sf_active ? 0 : 0);
  free(tt);
  free(tp);
  return MBT_OK;
}

extern "C" int mbt_ring_potentials(const mbt_ring *r, int ib, double *out, double *rfprev)
{
  if (ib < 0 || ib >= r->Nb || r->Nc <= 0) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  const size_t Nc = r->Nc, off = (size_t)ib * Nc, loff = (size_t)r->cfg.bucket[ib] * Nc;
  const double *src[7] = {r->sfL + off, r->lr_wake + loff, r->fnp_out + off, r->sfV + off, r->dipV_out + off,
                          r->sfH + off, r->dipH_out + off};
  for (int k = 0; k < 7; k++) CK(hipMemcpy(out + k * Nc, src[k], Nc * sizeof(double), hipMemcpyDeviceToHost));
  TurnState ts;
  CK(hipMemcpy(&ts, r->ts, sizeof ts, hipMemcpyDeviceToHost));
  if (rfprev) {
    rfprev[0] = ts.phai0_prev;
    rfprev[1] = ts.Vrf0_prev;
    rfprev[2] = ts.scan_val;
  }
  return MBT_OK;
}

extern "C" int mbt_ring_rw_history(const mbt_ring *r, double *out)
{
  if (!r->cfg.EnableRW_long) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  CK(hipMemcpy(out, r->rwhist, (size_t)(r->cfg.Nmlt + 2) * r->h * 2 * sizeof(double), hipMemcpyDeviceToHost));
  return MBT_OK;
}

extern "C" int mbt_ring_kernel_times(const mbt_ring *r, double *ms, long *launches)
{
  for (int i = 0; i < 4; i++) {
    ms[i] = r->kt_ms[i];
    launches[i] = r->kt_launches[i];
  }
  return MBT_OK;
}

extern "C" double mbt_ring_bytes_per_turn(const mbt_ring *r) { return 96.0 * (double)r->Ntot; }
extern "C" long mbt_ring_turn(const mbt_ring *r) { return (long)r->rev_host; }

extern "C" int mbt_ring_set_timing(mbt_ring *r, int on)
{
  CK(hipSetDevice(r->device));
  if (on && !r->evpool) {
    r->evpool = (hipEvent_t *)calloc(6 * STATS_TURNS, sizeof(hipEvent_t));
    for (int i = 0; i < 6 * STATS_TURNS; i++) CK(hipEventCreate(&r->evpool[i]));
  }
  r->timing = on;
  return MBT_OK;
}
