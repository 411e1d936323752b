// The multi-bunch tracking turn on one MI355X (gfx950).
//
// Every bunch is resident in HBM as six FP64 coordinates (tau, delta, x, x', z,
// z') in tiles of 4,096 slots: a tile holds one run of 4,096 values per
// coordinate, each bunch starts on a tile and bunches follow one another.  The unit is an ensemble of independent
// rings (one ring is an ensemble of one): per-ring constants sit in a DevParams
// table indexed through bunch -> ring, so every launch covers all rings and
// rings of uneven bunch and macroparticle counts share each kernel.  One turn
// is a few kernels on one stream (replayed as a HIP graph per block of turns):
//
//   k_turn    one pass over every macroparticle: pending RW kick, selffield
//             kick (src/transform_weak.c:697-731), one-turn map
//             (src/transform_weak.c:21-192), shifted moment sums and the next
//             turn's slice histogram with LDS-privatised exact integer sums;
//   k_reduce  per bunch: statistics (src/statistics.c:91-195), RW dipole history,
//             LR histograms (src/transform_weak.c:839-865) and the selffield
//             convolutions for the next turn (src/transform_weak.c:633-731);
//   k_rw_fresh, k_end (large rings) or k_rw (small rings)
//             per bunch: resistive-wall long-range sum over (Nmlt+1) h past
//             buckets (src/transform_weak.c:780-833), applied as a pending kick
//             at the start of the next pass; for large rings the taps of earlier
//             turns' rows are summed by k_turn's leading blocks during the map
//             and k_rw_fresh adds this turn's row after k_reduce; then the turn
//             end.  For ensembles of rings sharing one tap table the sum is one
//             MFMA GEMM instead (k_rwg_tile, k_rwg_sum, then k_end);
//   k_lr_par  with long-range resonators: the phasor recurrence over all buckets
//             and the RF-feedback samples (construct_wake_phasor
//             src/transform_weak.c:961-1048), linear, so one block per resonator
//             with the buckets in parallel and one short chain over bucket start
//             phasors (also for the replayed turns of wake_phasor_init); k_lr
//             keeps the one-lane recurrence (the reference's arithmetic) for
//             h > 1365; then k_lr_sum and k_fb_solve, the RF-feedback solve
//             (src/worker_weakweak.c:268-278).
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
#ifndef CHUNK_ITEMS
#define CHUNK_ITEMS 8
#endif
#define NMOM 15
// threads per bunch of the one-pass resistive-wall sum of small rings
#define RW_WG 1024
// macroparticles from which the RW taps of earlier turns' rows run in k_turn's
// leading blocks (below, the map is too short to hide them and the two-kernel
// tail costs more than the one-pass sum)
#define RW_SPLIT_MIN_MP (1LL << 22)
// coordinate footprints from which the set-up reserves and releases the whole
// footprint before allocating the tiled slab
#define COORD_RESERVE_MIN (1ULL << 30)
// consecutive taps per RW tile of k_turn's leading blocks, and its LDS budget
// (the slots of the tile's bunches and taps; k_turn's dynamic LDS grows to it)
#define RW_TILE_TAPS 256
#define RW_OLD_LDS_MAX (48 * 1024)
#ifndef TURN_WAVES_PER_EU
#define TURN_WAVES_PER_EU 4
#endif
#define STATS_TURNS 256
#define MIN_CHUNK_ITEMS 4
// rings below this many macroparticles may use 1 or 2 items per lane (their
// pass is latency-bound: fewer items per lane, more work-groups in flight);
// larger rings keep the buffers of 4 items per lane
#define SMALL_RING_MP (1LL << 22)
// work-groups the default decomposition aims for (eight per CU)
#define DEFAULT_WG_TARGET 2048
#define RWG_MIN_RINGS 8
#define AUTOTUNE_CANDIDATES 7
#define AUTOTUNE_ROUNDS 3
#define AUTOTUNE_REPS 2
#define MAX_CHUNK_ITEMS 64
#define LR_PREFETCH 16
// dynamic LDS of the bucket-parallel LR recurrence (6 doubles per bucket): up
// to h = 1365; larger rings use the one-lane recurrence
#define LR_PAR_LDS_MAX (64 * 1024)
// Coordinates in tiles of TILE slots: tile t holds coordinate k (tau, delta,
// x, x', z, z') of slots t TILE .. t TILE + TILE - 1 at (6 t + k) TILE, so a
// work-group's chunk reads and writes six 32 KB runs of one 192 KB region.
// Measured on the in-place stream of the production ring's working set (six
// FP64 values per slot): 6.2 TB/s in tiles of 4,096 slots against 5.4 TB/s in
// six separate arrays (5.6-5.9 in tiles of 1,024-2,048, 5.2 in 8,192).  Every
// bunch starts on a tile (its slots padded to whole tiles), so chunks of up to
// TILE slots never cross one.
#define TILE_SHIFT 12
#define TILE (1 << TILE_SHIFT)
#define BUNCH_ALIGN TILE
#ifndef COORD_NONTEMPORAL
#define COORD_NONTEMPORAL 1
#endif
#if COORD_NONTEMPORAL
#define LDC(p) __builtin_nontemporal_load(p)
#define STC(v, p) __builtin_nontemporal_store((v), (p))
#else
#define LDC(p) (*(p))
#define STC(v, p) (*(p) = (v))
#endif
#define MAXLR (3 * MBT_MAX_LR)

#define CK(x)                                                                                    \
  do {                                                                                           \
    hipError_t e_ = (x);                                                                         \
    if (e_ != hipSuccess) {                                                                      \
      mbt_set_error("%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), __FILE__, __LINE__);     \
      return MBT_EDEVICE;                                                                        \
    }                                                                                            \
  } while (0)

// Per-ring constants of the turn (one entry per ring of an ensemble; a single
// ring is an ensemble of one).  Derived on the host with glibc libm from the
// reference's formulas so that every constant equals the reference's.
struct DevParams {
  int ring, Nb, h, Nc, Nmlt;
  // bound on |x| and |z| of a particle whose dipole enters the exact fixed-point
  // slice sums: min(2^20, 2^27 / Np), so that the sum of a whole bunch stays
  // inside the 128-bit range (|sum| < 2^27 at the quantum 2^-100); a particle
  // beyond it marks the bunch's sums NaN
  double fx_lim;
  unsigned long long zero_bits; // always 0: the timing pass's stores xor with it
  int mixed;                    // compensated FP32 wake convolutions (MBT_RING_MIXED)
  int bunch_off;  // first global bunch of the ring
  int bucket_off; // first global bucket of the ring
  long long rw_off, st_off; // offsets into the RW history and sqrt table
  int track_h, track_v;
  int qe, ideal_hc, amp_dq;
  int kickL, kickV, kickH; // selffield kick of the plane is applied
  int convL;               // longitudinal Green function present (PlaneL > 0)
  int sf_active;
  int rw;
  int n_ahc;
  int monitor_every;
  int scan;
  long long NrevScan;
  double scan_step;
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
  // RF feedback (src/feedback_rf.c, src/input.c:1678-1690)
  int fb_len, fb_switch_off;
  long long fbh_off;
  double vrf_design, phi0_design;
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
  int ring, plane, l, fb;
  int h, bucket_off;
  int init_turns; // wake_phasor_init replay length Nbumax / h
  long long lrc_off;
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

#include "fixedpoint.h"

__device__ __forceinline__ void fx_atomic_add(unsigned long long *plo, unsigned long long *phi, unsigned long long lo,
                                              unsigned long long hi)
{
  unsigned long long old = atomicAdd(plo, lo);
  unsigned long long carry = (old + lo) < old ? 1ULL : 0ULL;
  atomicAdd(phi, hi + carry);
}

// ------------------------------------------------------------------ k_turn

// MODE_TUNE: the turn's full work for timing work decompositions -- the map's
// results feed the moment sums and the histogram, the coordinates are written
// back unchanged (every store and atomic target is scratch the next pass
// overwrites or the caller clears)
enum { MODE_TURN = 0, MODE_STATS = 1, MODE_FINAL = 2, MODE_TUNE = 3 };

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
  double *s; // the tiled slab: Nalloc / TILE tiles of 6 TILE doubles
};

// element of slot j's coordinate k in the tiled slab
__host__ __device__ __forceinline__ long long coord_idx(long long j, int k)
{
  return (((j >> TILE_SHIFT) * 6 + k) << TILE_SHIFT) + (j & (TILE - 1));
}

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

// fma(a, b, c) as the three-address v_fma_f64.  The compiler turns a
// polynomial step fma(z, p, coefficient) into the two-address v_fmac_f64,
// whose addend is also its result, and so copies the loop-invariant
// coefficient into a fresh register before every step (v_mov_b64 + v_fmac_f64,
// measured in the gfx950 assembly of k_turn); the same IEEE fused multiply-add
// in one instruction.  Register operands only, no memory access.
__device__ __forceinline__ double fma3(double a, double b, double c)
{
  double r;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
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
  double ps = fma3(z, 2.8114572543455206e-15, -7.647163731819816e-13);
  ps = fma3(z, ps, 1.6059043836821613e-10);
  ps = fma3(z, ps, -2.505210838544172e-08);
  ps = fma3(z, ps, 2.7557319223985893e-06);
  ps = fma3(z, ps, -0.0001984126984126984);
  ps = fma3(z, ps, 0.008333333333333333);
  ps = fma3(z, ps, -0.16666666666666666);
  const double sn = fma(r * z, ps, r);
  double pc = fma3(z, -1.5619206968586225e-16, 4.779477332387385e-14);
  pc = fma3(z, pc, -1.1470745597729725e-11);
  pc = fma3(z, pc, 2.08767569878681e-09);
  pc = fma3(z, pc, -2.755731922398589e-07);
  pc = fma3(z, pc, 2.48015873015873e-05);
  pc = fma3(z, pc, -0.001388888888888889);
  pc = fma3(z, pc, 0.041666666666666664);
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

// Arguments of the resistive-wall tiles k_turn's leading blocks run (nlead = 0:
// none): the taps of the history rows written in earlier turns.
struct RwOld {
  const int4 *tiles; // ring, first bunch, bunches, tap tile
  int ntile, nlead;  // tiles; leading blocks (ntile rounded up to whole rounds of the 8 XCDs)
  const double *rwhist, *rsqrt_tt, *Ibunch; // rsqrt_tt: 1 / sqrt(m T0 / h) per tap
  const int *nfFill;
  double *rw_part; // [tap tiles][G][2]
  int G;
};

// One tile: up to WG bunches of one ring (one lane each) times RW_TILE_TAPS
// consecutive taps of the resistive-wall sum (src/transform_weak.c:780-833).
// Tap m of the bunch in bucket kb reads history slot u = kb + m (slot
// (-rev h + u) mod n of the cyclic buffer, bucket u mod h); only m < h - kb
// reads the row this turn's k_reduce writes, so the taps m >= h - kb can be
// summed during the turn's map.  The tile's bunches share the span of slots
// [kb_lo + m0, kb_hi + m1], staged once in LDS, so the loads per turn are the
// history once per tile instead of once per bunch.  Each term is fNp * dip * I
// times the tabulated 1 / sqrt(t) (the reference divides by sqrt(t): one
// rounding apart, and a multiply instead of a division keeps these blocks short
// beside the map's); one plane per sweep keeps the register count low.
__device__ void rw_old_tile(const DevParams *__restrict__ Pr, const RwOld &RO, int t, const int *bucket,
                            const TurnState *__restrict__ tsr, double *sm)
{
  const int4 T = RO.tiles[t];
  const DevParams &P = Pr[T.x];
  const int tid = threadIdx.x;
  const int h = P.h, n = (P.Nmlt + 2) * h, M = (P.Nmlt + 1) * h;
  const long long rev = tsr[T.x].rev;
  const int m0 = 1 + T.w * RW_TILE_TAPS;
  const int m1 = min(m0 + RW_TILE_TAPS - 1, M);
  const int kb_lo = bucket[T.y], kb_hi = bucket[T.y + T.z - 1];
  const int u0 = kb_lo + m0, span = kb_hi - kb_lo + m1 - m0 + 1;
  double *lx = sm, *lz = lx + span, *lI = lz + span, *lst = lI + span;
  int *lf = (int *)(lst + RW_TILE_TAPS);
  const double *hist = RO.rwhist + P.rw_off;
  const double *Ib = RO.Ibunch + P.bucket_off;
  const int *fill = RO.nfFill + P.bucket_off;
  const long long k0 = (-rev * h + u0) % n;
  const int base = (int)(k0 < 0 ? k0 + n : k0);
  for (int e = tid; e < span; e += WG) {
    int idx = base + e;
    if (idx >= n) idx -= n;
    const int bk = (u0 + e) % h;
    lx[e] = hist[2 * idx];
    lz[e] = hist[2 * idx + 1];
    lI[e] = Ib[bk];
    lf[e] = fill[bk];
  }
  for (int i = tid; i <= m1 - m0; i += WG) lst[i] = RO.rsqrt_tt[P.st_off + m0 + i];
  __syncthreads();
  if (tid >= T.z) return;
  const int b = T.y + tid, kb = bucket[b];
  const int mlo = max(m0, h - kb); // earlier rows only: m >= h - kb
  const double fNp = P.fNp;
#pragma unroll 1
  for (int pl = 0; pl < 2; pl++) {
    const double *lv = pl ? lz : lx;
    double a = 0.0;
#pragma unroll 1
    for (int m = mlo; m <= m1; m++) {
      const int e = kb + m - u0;
      const double tm = fNp * lv[e] * lI[e] * lst[m - m0];
      a += lf[e] ? tm : 0.0;
    }
    RO.rw_part[((size_t)T.w * RO.G + b) * 2 + pl] = a;
  }
}

template <int MODE, bool QE, bool DIP, bool GEN, bool MON>
__global__ __launch_bounds__(WG, TURN_WAVES_PER_EU) void k_turn(const DevParams *__restrict__ Pr, Coords C,
                                                                const long long *wg_first, const long long *wg_end,
                                                                const int *wg_bunch, const int *bunch_ring,
                                                                const int *bucket, const long long *bunch_first,
                                                                const double *shift, const double *sfL,
                                                                const double *sfV, const double *sfH,
                                                                const double *lr_wake, long long lr_plane_stride,
                                                                const double *pending, double *partial, HistPtr H,
                                                                const TurnState *__restrict__ tsr, RwOld RO)
{
  extern __shared__ unsigned long long smem[];
  if (MODE == MODE_TURN && (int)blockIdx.x < RO.nlead) {
    // leading blocks: the resistive-wall taps of earlier turns' rows, dispatched
    // first so that they finish early in the pass
    if ((int)blockIdx.x < RO.ntile) rw_old_tile(Pr, RO, blockIdx.x, bucket, tsr, (double *)smem);
    return;
  }
  const int wg = xcd_remap(blockIdx.x - RO.nlead, gridDim.x - RO.nlead);
  const int b = wg_bunch[wg];
  const int ring = bunch_ring[b];
  const DevParams &P = Pr[ring];
  const TurnState *ts = tsr + ring;
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

  const int kb = bucket[b];
  const long long j0 = wg_first[wg], j1 = wg_end[wg];
  const int tid = threadIdx.x;
  const long long rev = ts->rev;
  const bool do_hist = Nc > 0 && MODE != MODE_FINAL;
  const double *lrL = lr_wake + (size_t)(P.bucket_off + kb) * Nc;
  const double *lrH = lrL + lr_plane_stride;
  const double *lrV = lrH + lr_plane_stride;

  if (DIP)
    for (int i = tid; i < 2 * Nc; i += WG) {
      lsx[i] = 0;
      lsz[i] = 0;
    }
  for (int i = tid; i < Nc; i += WG) {
    lcnt[i] = 0;
    if (MODE == MODE_TURN || MODE == MODE_TUNE) {
      if (P.kickL) kL[i] = sfL[(size_t)b * Nc + i] + lrL[i];
      if (P.kickV) kV[i] = sfV[(size_t)b * Nc + i] + lrV[i];
      if (P.kickH) kH[i] = sfH[(size_t)b * Nc + i] + lrH[i];
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
  const bool qe = QE && P.qe;

  double s1[6] = {0, 0, 0, 0, 0, 0}, s2[6] = {0, 0, 0, 0, 0, 0}, am[3] = {0, 0, 0};

  // block-uniform base pointers (the six runs of the chunk's first tile) +
  // one 32-bit lane offset for all six (scalar-base global loads)
  double *__restrict__ pt = C.s + ((j0 >> TILE_SHIFT) * 6 << TILE_SHIFT);
  double *__restrict__ pd = pt + TILE;
  double *__restrict__ px = pd + TILE;
  double *__restrict__ pxp = px + TILE;
  double *__restrict__ pz = pxp + TILE;
  double *__restrict__ pzp = pz + TILE;
  const unsigned cnt = (unsigned)(j1 - j0), o0 = (unsigned)(j0 & (TILE - 1));
  const unsigned jloc0 = (unsigned)(j0 - bunch_first[b]);

  for (unsigned ii = tid; ii < cnt; ii += WG) {
    // offset of the slot in the tiles from the chunk's first (chunks longer
    // than a tile cross into the next ones)
    const unsigned u = o0 + ii, i = ((u >> TILE_SHIFT) * 6u << TILE_SHIFT) + (u & (TILE - 1));
    // coordinates stream through once per turn: nontemporal loads and stores
    double tau = LDC(pt + i), del = LDC(pd + i), x = LDC(px + i), xp = LDC(pxp + i), z = LDC(pz + i),
           zp = LDC(pzp + i);
    const double in[6] = {tau, del, x, xp, z, zp}; // kept only by the timing pass
    if (rw_h) xp += rwx;
    if (rw_v) zp += rwz;
    if (MODE == MODE_TURN || MODE == MODE_TUNE) {
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
      double nq[5] = {0, 0, 0, 0, 0}, rff = 0.0;
      if (qe) {
        qe_normals(P.seed, (uint32_t)kb, (unsigned long long)rev, jloc0 + ii, nq);
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
        if (qe) {
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
        if (qe) {
          x = x + P.qHp * nq[3];
          xp = rff * xp + P.qHs * nq[4];
        }
      }
    }
    if (MODE == MODE_TUNE) {
      // a value the compiler cannot prove equal to the loaded one, so the
      // stores (and their traffic) stay; the bits written are the bits read
      const unsigned long long zb = P.zero_bits;
      STC(__longlong_as_double(__double_as_longlong(in[0]) ^ zb), pt + i);
      STC(__longlong_as_double(__double_as_longlong(in[1]) ^ zb), pd + i);
      STC(__longlong_as_double(__double_as_longlong(in[2]) ^ zb), px + i);
      STC(__longlong_as_double(__double_as_longlong(in[3]) ^ zb), pxp + i);
      STC(__longlong_as_double(__double_as_longlong(in[4]) ^ zb), pz + i);
      STC(__longlong_as_double(__double_as_longlong(in[5]) ^ zb), pzp + i);
    } else if (MODE != MODE_STATS) {
      STC(tau, pt + i);
      STC(del, pd + i);
      STC(x, px + i);
      STC(xp, pxp + i);
      STC(z, pz + i);
      STC(zp, pzp + i);
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
        if (!(fabs(x) < P.fx_lim) || !(fabs(z) < P.fx_lim)) {
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

// Wake convolution term by term in FP32 with hi/lo compensation (mixed
// precision): each FP64 operand split into a high and a low float, the product
// of the high parts exact through fmaf, the cross terms added to the low part,
// and the running sum kept as an unevaluated float pair (two-sum), so the
// cancellations of an oscillating Green function do not amplify FP32 rounding.
struct F2 {
  float hi, lo;
};
__device__ __forceinline__ void f2_add_prod(F2 &s, double a, double b)
{
  const float ah = (float)a, al = (float)(a - (double)ah);
  const float bh = (float)b, bl = (float)(b - (double)bh);
  const float p = ah * bh;
  const float pe = fmaf(ah, bh, -p) + (ah * bl + al * bh);
  const float t = s.hi + p;
  const float bb = t - s.hi;
  const float e = (s.hi - (t - bb)) + (p - bb);
  s.hi = t;
  s.lo = s.lo + (e + pe);
}

struct RedOut {
  double *stats;     // [STATS_TURNS][G][12]
  double *amp;       // [STATS_TURNS][G][6]
  double *shift;     // [G][6]
  double *rwhist;    // per ring [(Nmlt+2)h][2] at rw_off
  double *fnp_ring;  // [3][H*Nc]
  long long plane_stride; // H*Nc
  double *sfL, *sfV, *sfH;
  double *fnp_out, *dipV_out, *dipH_out; // last clamped histograms [G*Nc]
  int *trash_tot;    // [G*2] cumulative (consumed)
  int *trash_pend;   // [G*2]
  const double *G;   // Green functions [R][3][Nc]: Gl, Gv, Gh
  const double *bunch_Ib;
  const int *np;
  int Gtot;
  double *rfbuf;     // [STATS_TURNS][R][2] phase and voltage the turn's map used
  int R;
};

// phase: 0 = means only (shift <- mean), 1 = initial statistics + RW history
// fill, 2 = turn, 3 = final statistics
__global__ __launch_bounds__(WG) void k_reduce(const DevParams *__restrict__ Pr, const double *partial,
                                               const int *bunch_wg0, const int *bunch_ring, const int *bucket,
                                               HistPtr H, RedOut R, const TurnState *__restrict__ tsr, int phase,
                                               const int *lr_planes, int dip)
{
  extern __shared__ double sm[];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int ring = bunch_ring[b];
  const DevParams &P = Pr[ring];
  const TurnState *ts = tsr + ring;
  const int kb = bucket[b];
  const int Nc = P.Nc;
  double *fnp = sm;
  double *dV = fnp + Nc;
  double *dH = dV + Nc;
  double *Gs = dH + Nc;  // the ring's three Green's functions, [plane][Nc]
  __shared__ double mom[NMOM];
  __shared__ int icmin_s, icmax_s;
  const long long rev = ts->rev;
  const int w0 = bunch_wg0[b], w1 = bunch_wg0[b + 1];
  const double Np = (double)R.np[b];
  {
    // moment sums over the bunch's workgroups: 16 lanes per moment, each summing
    // every 16th workgroup in order, then a fixed butterfly over the 16 lanes
    const int q = tid >> 4, j = tid & 15;
    double s = 0.0;
    if (q < NMOM)
      for (int w = w0 + j; w < w1; w += 16) s += partial[(size_t)w * NMOM + q];
    for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 16);
    if (q < NMOM && j == 0) mom[q] = s;
  }
  if (phase != 0 && Nc > 0) {
    const double *Gr = R.G + (size_t)ring * 3 * Nc;
    for (int i = tid; i < 3 * Nc; i += WG) Gs[i] = Gr[i];
  }
  if (tid == 0) {
    icmin_s = Nc - 1;
    icmax_s = 0;
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
      double *o = R.stats + ((size_t)slot * R.Gtot + b) * 12;
      for (int i = 0; i < 12; i++) o[i] = st[i];
      for (int pl = 0; pl < 3; pl++) {
        R.shift[b * 6 + 2 * pl] = st[pl];
        R.shift[b * 6 + 2 * pl + 1] = st[6 + pl];
      }
      if (phase == 2 && b == P.bunch_off) {
        // nothing changes phai0 / Vrf0 between the map and this point of the turn
        R.rfbuf[((size_t)slot * R.R + ring) * 2] = ts->phai0;
        R.rfbuf[((size_t)slot * R.R + ring) * 2 + 1] = ts->Vrf0;
      }
      if (phase == 2 && (rev + 1) % P.monitor_every == 0) {
        double *a = R.amp + ((size_t)slot * R.Gtot + b) * 6;
        for (int pl = 0; pl < 3; pl++) {
          a[pl] = track[pl] ? mom[12 + pl] / Np : 0.0;
          a[3 + pl] = track[pl] ? P.gamma1[pl] * (st[pl] * st[pl]) + 2.0 * P.alpha1[pl] * st[pl] * st[6 + pl] +
                                      P.beta1[pl] * (st[6 + pl] * st[6 + pl])
                                : 0.0;
        }
      }
      if (P.rw && phase != 3) {
        double *hist = R.rwhist + P.rw_off;
        long long n = (long long)(P.Nmlt + 2) * P.h;
        long long k = (phase == 1 ? 0 : -rev * P.h) + kb;
        long long idx = ((k % n) + n) % n;
        hist[2 * idx] = st[1];
        hist[2 * idx + 1] = st[2];
        if (phase == 1)
          for (int jj = 1; jj < P.Nmlt; jj++) {
            long long i2 = ((long long)jj * P.h + kb) % n;
            hist[2 * i2] = st[1];
            hist[2 * i2 + 1] = st[2];
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
  // dip = 0: the pass before (a turn pass without dipole histograms) left the
  // dipole sums at the zeros the last clear wrote, so they are neither read
  // nor cleared here (the outputs get the same zeros)
  const bool bad = H.bad[b] != 0;
  const size_t lrrow = (size_t)(P.bucket_off + kb) * Nc;
  for (int i = tid; i < Nc; i += WG) {
    size_t g = (size_t)b * Nc + i;
    double cnt = (double)H.cnt[g];
    unsigned long long xl = 0, xh = 0, zl = 0, zh = 0;
    if (dip) {
      xl = H.sx[2 * g];
      xh = H.sx[2 * g + 1];
      zl = H.sz[2 * g];
      zh = H.sz[2 * g + 1];
    }
    double sx = bad ? NAN : fx_to_double(xl, xh);
    double sz = bad ? NAN : fx_to_double(zl, zh);
    if (lr_planes[0]) R.fnp_ring[lrrow + i] = cnt;
    if (lr_planes[1]) R.fnp_ring[R.plane_stride + lrrow + i] = sx;
    if (lr_planes[2]) R.fnp_ring[2 * R.plane_stride + lrrow + i] = sz;
    // clamped: out-of-range particles join the edge bins
    unsigned int c = H.cnt[g];
    for (int tt = 0; tt < 2; tt++) {
      if (!((tt == 0 && i == 0) || (tt == 1 && i == Nc - 1))) continue;
      c += H.tcnt[2 * b + tt];
      if (!dip) continue;
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
    fnp[i] = (double)c;
    // icellmin / icellmax of src/transform_weak.c:633-651: the first and last
    // occupied slice (Nc - 1 and 0 when there is none)
    if (c > 0) {
      atomicMin(&icmin_s, i);
      atomicMax(&icmax_s, i);
    }
    dV[i] = bad ? NAN : fx_to_double(zl, zh);
    dH[i] = bad ? NAN : fx_to_double(xl, xh);
    R.fnp_out[g] = fnp[i];
    R.dipV_out[g] = dV[i];
    R.dipH_out[g] = dH[i];
  }
  if (tid == 0) {
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
    if (dip) {
      H.sx[2 * g] = H.sx[2 * g + 1] = 0;
      H.sz[2 * g] = H.sz[2 * g + 1] = 0;
    }
  }
  if (tid < 2) {
    H.tcnt[2 * b + tid] = 0;
    if (dip) {
      H.tsx[4 * b + 2 * tid] = H.tsx[4 * b + 2 * tid + 1] = 0;
      H.tsz[4 * b + 2 * tid] = H.tsz[4 * b + 2 * tid + 1] = 0;
    }
  }
  if (tid == 0) H.bad[b] = 0;
  const int icmin = icmin_s, icmax = icmax_s;
  // the next turn's kick uses the current after this turn's scan step (src/tracking.c:43-59)
  double Ib = R.bunch_Ib[b];
  if (phase == 2 && P.scan == 1 && (rev + 1) % P.NrevScan == 0) Ib += P.scan_step / (((double)P.Nb) * 1.0e3);
  const double factG = -P.T0 * Ib / (R.np[b] * P.E0 * 1.0e9);
  const double sgm = P.sigma_tau, sgm2 = sgm * sgm, dT2 = P.dT * P.dT;
  // one (plane, slice) convolution per task; tasks run from the last slice down
  // so that the second sweep of the block gets the short ones
  for (int t = tid; t < 3 * Nc; t += WG) {
    const int pl = t % 3, i = Nc - 1 - t / 3;
    const double *Gp = Gs + pl * Nc;
    const double *src = pl == 0 ? fnp : (pl == 1 ? dV : dH);
    double val = 0.0;
    if (pl == 0) {
      if (P.convL && i >= icmin) {
        double Gn = 0.0;
        if (P.mixed) {
          F2 g = {0.0f, 0.0f};
          for (int k = i; k >= icmin; k--) f2_add_prod(g, src[k], Gp[i - k]);
          Gn = (double)g.hi + (double)g.lo;
        } else {
#pragma unroll 4
          for (int k = i; k >= icmin; k--) Gn += src[k] * Gp[i - k];
        }
        double GL1 = 1 * Gn;
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
        val = GL1;
      }
      R.sfL[(size_t)b * Nc + i] = factG * val;
    } else {
      if (pl == 1 ? P.kickV : (P.kickH && i >= 1)) {
        double Gn = 0.0;
        if (P.mixed) {
          F2 g = {0.0f, 0.0f};
          for (int k = i; k >= 0; k--) f2_add_prod(g, src[k], Gp[i - k]);
          Gn = (double)g.hi + (double)g.lo;
        } else {
#pragma unroll 4
          for (int k = i; k >= 0; k--) Gn += src[k] * Gp[i - k];
        }
        val = 1 * Gn;
      }
      // a plane without a kick keeps the +0.0 of its allocation (its value
      // here, -factG * 0 with factG <= 0), so its table is not rewritten
      if (pl == 1 ? P.kickV : P.kickH) (pl == 1 ? R.sfV : R.sfH)[(size_t)b * Nc + i] = -factG * val;
    }
  }
}

// ------------------------------------------------------------- k_rw_fresh, k_rw

// Resistive-wall long-range kick of bunch b (src/transform_weak.c:780-833): the
// sum over the taps m = 1 .. (Nmlt+1) h of the ring's dipole history, tap m
// reading the bucket m places after b's.  Only the taps m < h - kb read the row
// this turn's k_reduce writes; the others read rows of earlier turns.  So for
// large rings the sum is split: k_turn's leading blocks add the earlier rows'
// taps while the map runs (rw_old_tile), k_rw_fresh adds the fresh row's taps
// after k_reduce and stores the pending kick the next pass applies.
//
// rw_taps: NT lanes sum taps [mlo, mhi]; the block total ends in (sx, sz) of lane 0.
template <int NT, int UNR>
__device__ void rw_taps(const DevParams &P, int kb, long long rev, const double *rwhist, const double *sqrt_tt,
                        const double *Ibunch, const int *nfFill, long long mlo, long long mhi, double (*red)[2],
                        double &sx, double &sz)
{
  const int tid = threadIdx.x;
  const int h = P.h;
  const double *hist = rwhist + P.rw_off;
  const double *stt = sqrt_tt + P.st_off;
  const double *Ib = Ibunch + P.bucket_off;
  const int *fill = nfFill + P.bucket_off;
  const long long n = (long long)(P.Nmlt + 2) * h;
  // index of tap m in the cyclic history: (-rev h + kb + m) mod n, bucket (kb + m) mod h,
  // advanced incrementally (stride NT) instead of by 64-bit divisions per tap.
  // Every load of a tap is unconditional (all indices are in range) so that the
  // unrolled loop keeps several taps' loads in flight; an empty bucket adds +0.
  const long long m0 = mlo + tid;
  long long k0 = (-rev * h + kb + m0) % n;
  long long idx = k0 < 0 ? k0 + n : k0;
  int in_mod = (int)((kb + m0) % h);
  const long long step_n = NT % n;
  const int step_h = NT % h;
  double ax = 0.0, az = 0.0;
#pragma unroll UNR
  for (long long m = m0; m <= mhi; m += NT) {
    const int f = fill[in_mod];
    const double I = Ib[in_mod];
    const double st = stt[m];
    const double hx = hist[2 * idx], hz = hist[2 * idx + 1];
    const double tx = P.fNp * hx * I / st, tz = P.fNp * hz * I / st;
    ax += f ? tx : 0.0;
    az += f ? tz : 0.0;
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
  sx = 0.0;
  sz = 0.0;
  if (tid == 0)
    for (int w = 0; w < NT / 64; w++) {
      sx += red[w][0];
      sz += red[w][1];
    }
}

// Turn end of one ring: RW trigger counter, current scan (src/tracking.c:43-59),
// turn counter.
__device__ void end_turn(const DevParams &P, TurnState *ts, double *bunch_Ib, double *Ibunch, const int *nfFill)
{
  long long rev = ts->rev;
  if (P.rw) ts->m++;
  if ((rev + 1) % P.NrevScan == 0) {
    if (P.scan == 1) {
      ts->Iring += P.scan_step;
      ts->scan_val = ts->Iring;
      for (int k = 0; k < P.Nb; k++) bunch_Ib[P.bunch_off + k] += P.scan_step / (((double)P.Nb) * 1.0e3);
      for (int i = 0; i < P.h; i++)
        if (nfFill[P.bucket_off + i]) Ibunch[P.bucket_off + i] += P.scan_step / ((double)P.Nb);
    }
    ts->m = 0;
  }
  ts->rev = rev + 1;
}

// RF feedback of one ring: rffb_calc_mean_voltage_phase (src/feedback_rf.c:21-34)
// and the solve of src/worker_weakweak.c:268-278; the phase and voltage before
// the solve are kept for the potentials output.
// The solve from the sums of the sampled voltages vs and phases ps.
__device__ void fb_apply(const DevParams &P, TurnState *ts, double vs, double ps, int init)
{
  const int fb_len = P.fb_len;
  double vmbar = vs / fb_len, phibar = ps / fb_len;
  vmbar = vmbar * P.E0 * 1.0e3;
  double num = P.vrf_design * sin(P.phi0_design) - vmbar * sin(phibar);
  double ph = atan2(num, P.vrf_design * cos(P.phi0_design) - vmbar * cos(phibar));
  ts->phai0 = ph;
  ts->Vrf0 = num / sin(ph);
  if (init && P.fb_switch_off) ts->has_fb = 0;
}

__device__ void fb_solve(const DevParams &P, TurnState *ts, const double *fb_hist, int init)
{
  ts->phai0_prev = ts->phai0;
  ts->Vrf0_prev = ts->Vrf0;
  if (!ts->has_fb) return;
  const double *fbh = fb_hist + P.fbh_off;
  double vs = 0, ps = 0;
  for (int m = 0; m < P.fb_len; m++) {
    vs += fbh[2 * m];
    ps += fbh[2 * m + 1];
  }
  fb_apply(P, ts, vs, ps, init);
}

static size_t rw_old_lds(int span) { return (size_t)span * (3 * 8 + 4) + RW_TILE_TAPS * 8; }

// Per bunch, after k_reduce: the taps of this turn's row (m < h - kb) plus the
// earlier rows' tile partials (in tile order), times the RW constant of the
// trigger state.
__global__ __launch_bounds__(WG) void k_rw_fresh(const DevParams *__restrict__ Pr, const int *bunch_ring,
                                                 const double *rwhist, const double *sqrt_tt, const double *Ibunch,
                                                 const int *nfFill, const int *bucket,
                                                 const TurnState *__restrict__ tsr, const double *rw_part, int G,
                                                 double *pending)
{
  __shared__ double red[WG / 64][2];
  const int b = blockIdx.x, tid = threadIdx.x;
  const int ring = bunch_ring[b];
  const DevParams &P = Pr[ring];
  if (!P.rw) return;
  const int kb = bucket[b];
  const TurnState *ts = tsr + ring;
  double fx, fz;
  rw_taps<WG, 4>(P, kb, ts->rev, rwhist, sqrt_tt, Ibunch, nfFill, 1, (long long)(P.h - kb - 1), red, fx, fz);
  __syncthreads();
  // earlier rows: one tile partial per lane (fixed order)
  const int ntile = ((P.Nmlt + 1) * P.h + RW_TILE_TAPS - 1) / RW_TILE_TAPS;
  double ox = 0.0, oz = 0.0;
  for (int c = tid; c < ntile; c += WG) {
    ox += rw_part[((size_t)c * G + b) * 2];
    oz += rw_part[((size_t)c * G + b) * 2 + 1];
  }
  ox = wave_sum(ox);
  oz = wave_sum(oz);
  if ((tid & 63) == 0) {
    red[tid >> 6][0] = ox;
    red[tid >> 6][1] = oz;
  }
  __syncthreads();
  if (tid == 0) {
    double sx = 0.0, sz = 0.0;
    for (int w = 0; w < WG / 64; w++) {
      sx += red[w][0];
      sz += red[w][1];
    }
    const int trig = ts->m < 50 ? 1 : 0;
    pending[2 * b] = P.track_h ? (sx + fx) * P.rwconst[0][trig] : 0.0;
    pending[2 * b + 1] = P.track_v ? (sz + fz) * P.rwconst[1][trig] : 0.0;
  }
}

// Small rings (the map too short to hide the earlier rows' taps): every tap in one pass, and
// the ring's last block to finish ends the turn (every block of the ring has read
// the turn state and the bucket currents before its increment).
__global__ __launch_bounds__(RW_WG) void k_rw(const DevParams *__restrict__ Pr, const int *bunch_ring,
                                              const double *rwhist, const double *sqrt_tt, double *Ibunch,
                                              const int *nfFill, const int *bucket, double *pending,
                                              TurnState *tsr, unsigned *ring_done, double *bunch_Ib, int fb_here)
{
  __shared__ double red[RW_WG / 64][2];
  __shared__ int last;
  const int b = blockIdx.x, tid = threadIdx.x;
  const int ring = bunch_ring[b];
  const DevParams &P = Pr[ring];
  TurnState *ts = tsr + ring;
  if (P.rw) {
    const int kb = bucket[b];
    double sx, sz;
    rw_taps<RW_WG, 4>(P, kb, ts->rev, rwhist, sqrt_tt, Ibunch, nfFill, 1, (long long)(P.Nmlt + 1) * P.h, red, sx,
                      sz);
    if (tid == 0) {
      const int trig = ts->m < 50 ? 1 : 0;
      pending[2 * b] = P.track_h ? sx * P.rwconst[0][trig] : 0.0;
      pending[2 * b + 1] = P.track_v ? sz * P.rwconst[1][trig] : 0.0;
    }
  }
  if (tid == 0) {
    __threadfence();
    last = atomicAdd(&ring_done[ring], 1u) == (unsigned)(P.Nb - 1);
  }
  __syncthreads();
  if (last && tid == 0) {
    __threadfence();
    ring_done[ring] = 0;
    end_turn(P, ts, bunch_Ib, Ibunch, nfFill);
    if (fb_here) fb_solve(P, ts, nullptr, 0);
  }
}

// ------------------------------------------------------- RW sum as one GEMM

// Resistive-wall sum of many rings at once (rings of one ensemble with equal h,
// Nmlt and T0 share the tap weights c[m] = 1 / sqrt(m T0 / h)):
//   out[kb][col] = sum_a T[kb][a] Y[a][col],  T[kb][a] = c[a - kb] (1 <= a - kb <= M)
// a Toeplitz tap matrix times one column per (ring, plane) of
//   Y[a][col] = hist[(-rev h + a) mod n][plane] * Ibunch[a mod h]   (filled buckets)
// -- the sum of src/transform_weak.c:806-824 for every bunch of every ring; the
// ring's fNp and RWconst scale the result.
//
// k_rwg_tile: one block per (16-column tile, K slice of `per` steps of 4); the
// block forms its slice of Y in LDS straight from the rings' cyclic dipole
// histories and bucket currents (Y is never stored) and stages the tap weights
// its rows need, then each wave multiplies up to RWG_RT row tiles per pass out of LDS
// (MFMA 16x16x4: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]) and the
// block writes its K slice's partial sums.  k_rwg_sum adds the slices in a fixed
// order and stores the pending kicks.
// FP64: v_mfma_f64_16x16x4_f64.  F32 (mixed precision): both operands split into
// high and low floats, hi*hi + hi*lo + lo*hi with v_mfma_f32_16x16x4_f32, summed
// in FP32 over RWG32_CHUNK steps and then in FP64.

typedef double mfma_d4 __attribute__((ext_vector_type(4)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
#define RWG_RT 8
#define RWG32_CHUNK 64
#define RWG_LDS_MAX (64 * 1024)
#ifndef RWG_PER
#define RWG_PER 48 // K steps (of 4) per slice (48 measured best of 24, 48, 96)
#endif

struct RWGemm {
  int K;        // rows of Y (a = 0 .. K-1), multiple of 16
  int N;        // columns: 2 per RW ring
  int M;        // taps (Nmlt + 1) h
  int h, n;     // buckets, history length (Nmlt + 2) h
  int HP;       // h rounded up to 16 (rows of a partial)
  int per, KS;  // K steps (of 4) per slice, slices
  const int *col_ring;     // [N/2] ring of each column pair
  const double *c;         // [M + 1] tap weights, c[0] unused
  double *part;            // [KS][HP][N] K-slice partial sums
  const int *bucket_bunch; // [H] global bucket -> bunch (-1 empty)
};

// LDS bytes of one k_rwg_tile block: Y slice [4 per][16] + taps [4 per + HP - 1]
static inline size_t rwg_lds(int per, int HP) { return sizeof(double) * ((size_t)68 * per + HP); }

// The K loop of one wave over NG row tiles rt0, rt0 + 4, ... (all valid):
// A[i][a] = c[a - i] read from the staged taps at al - i + HP - 1, B = the Y
// slice.  Rows i >= h (up to HP) read taps too: their sums are never read.
// F32: operands split into high and low floats, hi*hi + hi*lo + lo*hi in FP32,
// summed in FP32 over RWG32_CHUNK steps and then in FP64.
template <bool F32, int NG>
__device__ __forceinline__ void rwg_rows(const double *Yf, const double *cs, int nA, int HP, int rt0, int kq,
                                         int l16, double *part, int N, int j)
{
  mfma_d4 acc[NG];
  mfma_f4 acc32[NG];
#pragma unroll
  for (int g = 0; g < NG; g++) {
    acc[g] = (mfma_d4){0.0, 0.0, 0.0, 0.0};
    acc32[g] = (mfma_f4){0.0f, 0.0f, 0.0f, 0.0f};
  }
  for (int st = 0; st < nA / 4; st++) {
    const int al = 4 * st + kq;
    const double bv = Yf[al * 16 + l16];
    double avs[NG];
#pragma unroll
    for (int g = 0; g < NG; g++) avs[g] = cs[al - ((rt0 + 4 * g) * 16 + l16) + HP - 1];
    if (F32) {
      const float bh = (float)bv, bl = (float)(bv - (double)bh);
#pragma unroll
      for (int g = 0; g < NG; g++) {
        const float ah = (float)avs[g], alo = (float)(avs[g] - (double)ah);
        acc32[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(alo, bh, acc32[g], 0, 0, 0);
        acc32[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bl, acc32[g], 0, 0, 0);
        acc32[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(ah, bh, acc32[g], 0, 0, 0);
      }
      if ((st + 1) % RWG32_CHUNK == 0 || st == nA / 4 - 1) {
#pragma unroll
        for (int g = 0; g < NG; g++) {
#pragma unroll
          for (int v = 0; v < 4; v++) acc[g][v] += (double)acc32[g][v];
          acc32[g] = (mfma_f4){0.0f, 0.0f, 0.0f, 0.0f};
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < NG; g++) acc[g] = __builtin_amdgcn_mfma_f64_16x16x4f64(avs[g], bv, acc[g], 0, 0, 0);
    }
  }
  // part: this K slice's partial sums [HP][N]; column j of this lane
  if (j < N) {
#pragma unroll
    for (int g = 0; g < NG; g++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        // result layouts: f64 16x16x4 element v of lane l is row l / 16 + 4 v,
        // f32 16x16x4 row 4 (l / 16) + v; column l % 16 in both
        const int row = (rt0 + 4 * g) * 16 + (F32 ? 4 * kq + v : kq + 4 * v);
        part[(size_t)row * N + j] = acc[g][v];
      }
  }
}

template <bool F32>
__global__ __launch_bounds__(256) void k_rwg_tile(RWGemm Gm, const DevParams *__restrict__ Pr,
                                                const TurnState *__restrict__ tsr, const double *rwhist,
                                                const double *Ibunch, const int *nfFill)
{
  extern __shared__ double sm[];
  const int tj = blockIdx.x, ks = blockIdx.y, tid = threadIdx.x;
  // the wave index as a uniform value: the row-tile tests below are then
  // scalar branches, not exec masks around the accumulators
  const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6), kq = lane >> 4, l16 = lane & 15;
  const int h = Gm.h, M = Gm.M, HP = Gm.HP;
  const int ksteps = Gm.K / 4;
  const int s0 = ks * Gm.per, s1 = s0 + Gm.per < ksteps ? s0 + Gm.per : ksteps;
  const int A0 = 4 * s0, nA = 4 * (s1 - s0);
  double *Yf = sm;                        // [nA][16]
  double *cs = sm + (size_t)64 * Gm.per;  // taps m = A0 - HP + 1 + e
  __shared__ long long ybase[16], yroff[16]; // history slot of row A0 and the ring's history offset, per column
  __shared__ int yring[16], yboff[16];       // ring and its bucket offset, per column
  // ---- stage the Y slice (formed here from the history: Y[a][col] = dipole of
  // history slot (-rev h + a) mod n times the current of bucket a mod h, zero
  // for empty buckets and rows outside 1 <= a < h + M) and the taps
  if (tid < 16) {
    const int j = tj * 16 + tid;
    const int ring = j < Gm.N ? Gm.col_ring[j >> 1] : -1;
    yring[tid] = ring;
    if (ring >= 0) {
      long long k0 = (-tsr[ring].rev * h + A0) % Gm.n;
      ybase[tid] = k0 < 0 ? k0 + Gm.n : k0;
      yboff[tid] = Pr[ring].bucket_off;
      yroff[tid] = Pr[ring].rw_off;
    }
  }
  __syncthreads();
  const int q0 = A0 % h;
  for (int e = tid; e < nA * 16; e += 256) {
    const int t = e >> 4, jl = e & 15, ring = yring[jl];
    const int a = A0 + t;
    double v = 0.0;
    if (ring >= 0 && a >= 1 && a < h + M) {
      int q = q0 + t; // bucket a mod h (q0 < h, t < 4 per)
      while (q >= h) q -= h;
      if (nfFill[yboff[jl] + q]) {
        long long idx = ybase[jl] + t;
        if (idx >= Gm.n) idx %= Gm.n; // t < 4 per may exceed n when h is small
        v = rwhist[yroff[jl] + 2 * idx + (jl & 1)] * Ibunch[yboff[jl] + q];
      }
    }
    Yf[e] = v;
  }
  const int mlo = A0 - HP + 1;
  for (int e = tid; e < nA + HP - 1; e += 256) {
    const int m = mlo + e;
    cs[e] = (m >= 1 && m <= M) ? Gm.c[m] : 0.0;
  }
  __syncthreads();
  // ---- row tiles out of LDS: wave wv takes tiles wv, wv + 4, ... in groups of
  // up to RWG_RT; the group's tile count is a template argument of the K loop,
  // so that the loop has no conditions and no accumulator moves between
  // register files around its MFMAs
  const int nrt = HP / 16;
  double *part = Gm.part + (size_t)ks * HP * Gm.N;
  const int j = tj * 16 + l16;
  for (int rt0 = wv; rt0 < nrt; rt0 += 4 * RWG_RT) {
    const int ng = min(RWG_RT, (nrt - rt0 + 3) / 4);
    switch (ng) {
    case 1: rwg_rows<F32, 1>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 2: rwg_rows<F32, 2>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 3: rwg_rows<F32, 3>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 4: rwg_rows<F32, 4>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 5: rwg_rows<F32, 5>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 6: rwg_rows<F32, 6>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    case 7: rwg_rows<F32, 7>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    default: rwg_rows<F32, 8>(Yf, cs, nA, HP, rt0, kq, l16, part, Gm.N, j); break;
    }
  }
}

// one lane per (bucket, column): the K slices summed in a fixed order, then the
// pending kick of the bucket's bunch (the ring's fNp and RW constant applied)
__global__ __launch_bounds__(256) void k_rwg_sum(RWGemm Gm, const DevParams *__restrict__ Pr,
                                               const TurnState *__restrict__ tsr, double *pending)
{
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)Gm.h * Gm.N) return;
  const int kb = (int)(e / Gm.N), col = (int)(e % Gm.N);
  const int ring = Gm.col_ring[col >> 1], pl = col & 1;
  const DevParams &P = Pr[ring];
  const int g = Gm.bucket_bunch[P.bucket_off + kb];
  if (g < 0) return;
  double sum = 0.0;
  for (int ks = 0; ks < Gm.KS; ks++) sum += Gm.part[((size_t)ks * Gm.HP + kb) * Gm.N + col];
  const int trig = tsr[ring].m < 50 ? 1 : 0;
  const bool on = pl == 0 ? P.track_h : P.track_v;
  pending[2 * g + pl] = on ? (P.fNp * sum) * P.rwconst[pl][trig] : 0.0;
}

// Turn end, one block per ring, for turns without a resistive-wall pass;
// with fb_here also the RF-feedback step (no long-range resonators anywhere).
__global__ void k_end(const DevParams *__restrict__ Pr, TurnState *tsr, double *bunch_Ib, double *Ibunch,
                      const int *nfFill, int fb_here)
{
  if (threadIdx.x != 0) return;
  const int ring = blockIdx.x;
  TurnState *ts = tsr + ring;
  end_turn(Pr[ring], ts, bunch_Ib, Ibunch, nfFill);
  if (fb_here) fb_solve(Pr[ring], ts, nullptr, 0);
}

// ------------------------------------------------------------------- k_lr

// One lane per long-range resonator (of any ring): the recurrence of
// construct_wake_phasor (src/transform_weak.c:961-1048) or, with init set,
// wake_phasor_init (:869-952).  Constants come from the host (glibc exp/cos/sin),
// the update order is the reference's, so the phasors are bit-identical.
__global__ void k_lr(int nres, const LRConst *lc, const DevParams *__restrict__ Pr, const double *fnp_ring,
                     long long plane_stride, const double *Ibunch, const int *nfFill, int Nc, double *phasor_end,
                     double *lrc, double *fb_hist, TurnState *tsr, int init)
{
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nres) return;
  const LRConst L = lc[r];
  const DevParams &P = Pr[L.ring];
  TurnState *ts = tsr + L.ring;
  const int h = L.h;
  const double *fnp = fnp_ring + (size_t)L.plane * plane_stride + (size_t)L.bucket_off * Nc;
  const double *Ib = Ibunch + L.bucket_off;
  const int *fill = nfFill + L.bucket_off;
  double *__restrict__ out = lrc + L.lrc_off;
  double *__restrict__ fbh = fb_hist + P.fbh_off;
  const int fb_len = P.fb_len;
  const int nturn = init ? L.init_turns : 1;
  if (nturn <= 0) return;
  double Vo0, Vo1, Vn0, Vn1;
  if (init) {
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
  for (int k = 0; k < nturn; k++)
    for (int m = 0; m < h; m++) {
      int i = h - m - 1;
      if (fb) {
        double vr = Vn0 * L.pb0 - Vn1 * L.pb1, vi = Vn0 * L.pb1 + Vn1 * L.pb0;
        unsigned key = ((unsigned)cnt % (unsigned)fb_len) * 2;
        fbh[key] = sqrt(vr * vr + vi * vi);
        fbh[key + 1] = atan2(vr, vi);
        cnt++;
      }
      if (fill[i] == 1) {
        double Ii = Ib[i];
        const double *__restrict__ f = fnp + (size_t)i * Nc;
        double *__restrict__ o = out + (size_t)i * Nc;
        // the recurrence is latency-bound: fetch the bucket's slice values a
        // block ahead of the dependent chain
        for (int j0 = 0; j0 < Nc; j0 += LR_PREFETCH) {
          double fv[LR_PREFETCH];
#pragma unroll
          for (int k2 = 0; k2 < LR_PREFETCH; k2++) fv[k2] = j0 + k2 < Nc ? f[j0 + k2] : 0.0;
#pragma unroll
          for (int k2 = 0; k2 < LR_PREFETCH; k2++) {
            if (j0 + k2 >= Nc) break;
            Vn0 = (Vo0 * L.pr0 - Vo1 * L.pr1) - Ii * fv[k2] * L.fac;
            Vn1 = (Vo0 * L.pr1 + Vo1 * L.pr0);
            if (!init) o[j0 + k2] = Vn0 + 0.5 * Ii * fv[k2] * L.fac;
            Vo0 = Vn0;
            Vo1 = Vn1;
          }
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
        fbh[key] = sqrt(vr * vr + vi * vi);
        fbh[key + 1] = atan2(vr, vi);
        cnt++;
      }
    }
  phasor_end[2 * r] = Vn0;
  phasor_end[2 * r + 1] = Vn1;
  if (fb) ts->fb_count = cnt;
}

// The per-turn recurrence of one resonator with its buckets in parallel (one
// block per resonator).  The recurrence is linear: a filled bucket maps its
// start phasor V to G (p^Nc V + L), with L the bucket's own sum started from
// zero, an empty one to E V.  Phase A: every bucket's L (one lane per bucket,
// the recurrence's own operations); phase B: one lane chains the h start
// phasors; phase C: every bucket replays its slices from its start phasor
// (again the recurrence's own operations) and writes the slice terms and the
// RF-feedback samples.  Only the start phasors are rounded differently from
// the sequential loop (by the closed form of phase B).  Dynamic LDS: 6 h doubles.
__device__ __forceinline__ void fb_sample(double *fbh, int fb_len, long long c, double v0, double v1, double p0,
                                          double p1)
{
  const double vr = v0 * p0 - v1 * p1, vi = v0 * p1 + v1 * p0;
  const unsigned key = ((unsigned)c % (unsigned)fb_len) * 2;
  fbh[key] = sqrt(vr * vr + vi * vi);
  fbh[key + 1] = atan2(vr, vi);
}

__global__ __launch_bounds__(256) void k_lr_par(const LRConst *lc, const DevParams *__restrict__ Pr,
                                                const double *fnp_ring, long long plane_stride, const double *Ibunch,
                                                const int *nfFill, int Nc, double *phasor_end, double *lrc,
                                                double *fb_hist, TurnState *tsr, int init)
{
  extern __shared__ double lsm[];
  __shared__ double Pn[2];
  const int r = blockIdx.x, tid = threadIdx.x, nth = blockDim.x;
  const LRConst L = lc[r];
  const DevParams &P = Pr[L.ring];
  TurnState *ts = tsr + L.ring;
  const int h = L.h;
  const double *fnp = fnp_ring + (size_t)L.plane * plane_stride + (size_t)L.bucket_off * Nc;
  const double *Ib = Ibunch + L.bucket_off;
  const int *fill = nfFill + L.bucket_off;
  double *__restrict__ out = lrc + L.lrc_off;
  double *__restrict__ fbh = fb_hist + P.fbh_off;
  const int fb_len = P.fb_len;
  const bool fb = L.fb && ts->has_fb;
  const long long cnt0 = ts->fb_count;
  // wake_phasor_init (init): the same turn replayed init_turns times from a
  // zero phasor without slice terms (src/transform_weak.c:869-952)
  const int nturn = init ? L.init_turns : 1;
  if (nturn <= 0) return;
  double *Lr = lsm, *Li = Lr + h, *Sr = Li + h, *Si = Sr + h, *Ar = Si + h, *Ai = Ar + h;
  // ---- phase A
  for (int i = tid; i < h; i += nth) {
    double V0 = 0.0, V1 = 0.0;
    if (fill[i] == 1) {
      const double Ii = Ib[i];
      const double *f = fnp + (size_t)i * Nc;
      for (int j = 0; j < Nc; j++) {
        const double n0 = (V0 * L.pr0 - V1 * L.pr1) - Ii * f[j] * L.fac;
        const double n1 = (V0 * L.pr1 + V1 * L.pr0);
        V0 = n0;
        V1 = n1;
      }
    }
    Lr[i] = V0;
    Li[i] = V1;
  }
  if (tid == 0) {
    double a = 1.0, b = 0.0;
    for (int j = 0; j < Nc; j++) {
      const double na = a * L.pr0 - b * L.pr1, nb = a * L.pr1 + b * L.pr0;
      a = na;
      b = nb;
    }
    Pn[0] = a;
    Pn[1] = b;
  }
  __syncthreads();
  // ---- phase B, in the recurrence's bucket order (last bucket first); when
  // replaying (init) it also takes the feedback samples that survive in the
  // sample ring, one lane in order
  if (tid == 0) {
    double V0 = init ? 0.0 : phasor_end[2 * r], V1 = init ? 0.0 : phasor_end[2 * r + 1];
    double W0 = phasor_end[2 * r], W1 = phasor_end[2 * r + 1]; // what the first sample reads
    const double g0 = L.ge * L.gc, g1 = L.ge * L.gs, e0 = L.ee * L.ec, e1 = L.ee * L.es;
    const long long keep_init = cnt0 + 2LL * h * nturn - fb_len;
    long long c = cnt0;
    for (int k = 0; k < nturn; k++)
      for (int m = 0; m < h; m++) {
        const int i = h - m - 1;
        if (init && fb) {
          if (c >= keep_init) fb_sample(fbh, fb_len, c, W0, W1, L.pb0, L.pb1);
          c++;
        }
        Sr[i] = V0;
        Si[i] = V1;
        if (fill[i] == 1) {
          const double w0 = (V0 * Pn[0] - V1 * Pn[1]) + Lr[i], w1 = (V0 * Pn[1] + V1 * Pn[0]) + Li[i];
          V0 = w0 * g0 - w1 * g1;
          V1 = w0 * g1 + w1 * g0;
        } else {
          const double n0 = V0 * e0 - V1 * e1, n1 = V0 * e1 + V1 * e0;
          V0 = n0;
          V1 = n1;
        }
        Ar[i] = V0;
        Ai[i] = V1;
        W0 = V0;
        W1 = V1;
        if (init && fb) {
          if (c >= keep_init) fb_sample(fbh, fb_len, c, W0, W1, L.pa0, L.pa1);
          c++;
        }
      }
    phasor_end[2 * r] = V0;
    phasor_end[2 * r + 1] = V1;
  }
  __syncthreads();
  if (init) {
    if (tid == 0 && fb) ts->fb_count = cnt0 + 2LL * h * nturn;
    return;
  }
  // ---- phase C; of the 2 h feedback samples only the last fb_len survive in
  // the sample ring, so only those are written (no two lanes share a slot)
  const long long keep_from = cnt0 + 2LL * h - fb_len;
  for (int i = tid; i < h; i += nth) {
    if (fill[i] == 1) {
      double V0 = Sr[i], V1 = Si[i];
      const double Ii = Ib[i];
      const double *f = fnp + (size_t)i * Nc;
      double *o = out + (size_t)i * Nc;
      for (int j = 0; j < Nc; j++) {
        const double fj = f[j];
        const double n0 = (V0 * L.pr0 - V1 * L.pr1) - Ii * fj * L.fac;
        const double n1 = (V0 * L.pr1 + V1 * L.pr0);
        o[j] = n0 + 0.5 * Ii * fj * L.fac;
        V0 = n0;
        V1 = n1;
      }
    }
    if (fb) {
      const long long c = cnt0 + 2LL * (h - i - 1);
      if (c >= keep_from) fb_sample(fbh, fb_len, c, Sr[i], Si[i], L.pb0, L.pb1);
      if (c + 1 >= keep_from) fb_sample(fbh, fb_len, c + 1, Ar[i], Ai[i], L.pa0, L.pa1);
    }
  }
  __syncthreads(); // every lane has read the sample counter
  if (tid == 0 && fb) ts->fb_count = cnt0 + 2LL * h;
}

// lr_wake[plane][global bucket][j] = sum over the ring's resonators of that
// plane, in order, of their terms
__global__ void k_lr_sum(const int *lr_first, const LRConst *lc, const double *lrc, const DevParams *__restrict__ Pr,
                         const int *bunch_ring, const int *bucket, int G, int Nc, int4 planes,
                         long long plane_stride, double *lr_wake)
{
  // (plane among the planes.x active ones, bunch, slice): only the buckets of
  // bunches are read by the kick; the rows of empty buckets and of planes
  // without resonators stay zero from the allocation
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)G * Nc;
  if (e >= planes.x * per) return;
  const int pi = (int)(e / per);
  const int plane = pi == 0 ? planes.y : (pi == 1 ? planes.z : planes.w);
  const long long bj = e % per;
  const int b = (int)(bj / Nc), j = (int)(bj % Nc);
  const int ring = bunch_ring[b];
  const int q = Pr[ring].bucket_off + bucket[b];
  double s = 0.0;
  // the ring's resonators are contiguous in lc (ring-major, then plane, then index)
  for (int r = lr_first[ring]; r < lr_first[ring + 1]; r++)
    if (lc[r].plane == plane) s += lrc[lc[r].lrc_off + (size_t)(q - lc[r].bucket_off) * Nc + j];
  lr_wake[plane * plane_stride + (long long)q * Nc + j] = s;
}

// RF feedback, one lane per ring: rffb_calc_mean_voltage_phase
// (src/feedback_rf.c:21-34) and the solve of src/worker_weakweak.c:268-278.
// One wave per ring: the samples summed by 64 lanes (fixed order), then the solve.
__global__ __launch_bounds__(64) void k_fb_solve(int nrings, const DevParams *__restrict__ Pr, TurnState *tsr,
                                                 const double *fb_hist, int init)
{
  const int ring = blockIdx.x, lane = threadIdx.x;
  if (ring >= nrings) return;
  const DevParams &P = Pr[ring];
  TurnState *ts = tsr + ring;
  const int has = ts->has_fb;
  if (lane == 0) {
    ts->phai0_prev = ts->phai0;
    ts->Vrf0_prev = ts->Vrf0;
  }
  if (!has) return;
  const double *fbh = fb_hist + P.fbh_off;
  double vs = 0.0, ps = 0.0;
  for (int m = lane; m < P.fb_len; m += 64) {
    vs += fbh[2 * m];
    ps += fbh[2 * m + 1];
  }
  vs = wave_sum(vs);
  ps = wave_sum(ps);
  if (lane == 0) fb_apply(P, ts, vs, ps, init);
}

// =============================================================== host side

struct RingHost {
  mbt_config cfg; // arrays below are owned copies
  int *nfFill, *bucket, *bunch_out;
  double *Ibunch, *bunch_Ib, *Ib_frac;
  int planes[3];
  int lr_plane[3];
};

#define NGRAPH 4
struct GraphEntry {
  hipGraphExec_t exec;
  long mon; // offset of the block's first monitor turn, -1 when it has none
  long n;   // turns
};

struct mbt_ring {
  int R;
  RingHost *rh;
  DevParams *Ph; // host copies of the per-ring parameters
  int device;
  hipStream_t stream;
  int G, H, Nc;    // bunches, buckets (all rings), slices
  long long Ntot;   // macroparticles
  long long Nalloc; // coordinate slots: bunches padded to BUNCH_ALIGN
  int rwg_on;        // RW sum as one GEMM over the ensemble's RW rings
  int force_gemm;    // MBT_RING_RW_GEMM: the GEMM form from one RW ring on
  int mixed;         // MBT_RING_MIXED: FP32 wake convolutions, FP32 hi/lo RW GEMM
  RWGemm rwg;
  int *rwg_col_ring, *rwg_bucket_bunch;
  double *rwg_c, *rwg_part;
  int nwg, nwg_cap;  // work-groups of the current decomposition, buffer capacity
  int chunk_items;   // macroparticles per lane of one work-group's chunk
  int min_items;     // smallest chunk the buffers allow (1 for small rings, else 4)
  int any_qe, any_dip, any_gen, nrev_mon;
  int sf_any;
  Coords C;
  long long *wg_first, *wg_end, *bunch_first;
  int *wg_bunch, *bunch_wg0, *bucket, *bunch_ring, *bucket_ring, *np, *nfFill, *lr_planes;
  DevParams *P;
  double *partial, *shift;
  HistPtr Hs;
  RedOut Ro;
  double *sfL, *sfV, *sfH, *lr_wake, *fnp_ring, *lrc;
  double *Gf;
  double *pending, *rwhist, *sqrt_tt, *rsqrt_tt, *Ibunch, *bunch_Ib;
  double *peek; // one bunch's six coordinates in SoA order, for mbt_ensemble_get_bunch
  double *stats, *amp, *rfbuf;
  double *fnp_out, *dipV_out, *dipH_out;
  int *trash_tot, *trash_pend;
  TurnState *ts;
  int4 *rw_tiles;      // RW tiles of k_turn's leading blocks: ring, first bunch, bunches, tap tile
  double *rw_part;     // [tap tiles][G][2] partial sums of earlier turns' rows
  int rw_ntile;
  size_t rw_lds;
  unsigned *ring_done; // [R] blocks of the ring done with the turn's fused RW pass
  int rw_split;        // RW sum split: earlier rows' taps in k_turn's leading blocks, then k_rw_fresh
  GraphEntry graphs[NGRAPH];
  int graph_next, use_graphs;
  long graph_builds; // turn graphs captured and instantiated so far
  int place_candidates; // coordinate allocations tried at set-up
  double place_gbps;    // k_stream rate of the chosen one (0 when not searched)
  LRConst *lc;
  int *lr_first; // [R+1] first resonator of each ring in lc
  int nres;
  int lr_hmax; // largest h among the resonators' rings
  int4 lr_planes_active; // x: planes with a resonator in some ring, y..w: those planes
  double *phasor_end, *fb_hist;
  long long rw_total, st_total, fbh_total;
  size_t lds_turn, lds_red;
  long long rev_host;
  double *pot_host;    // pinned staging of mbt_ring_potentials: 7 Nc doubles, then a TurnState
  long async_n;        // turns enqueued by mbt_ring_track_async, not yet collected (0: none)
  long long async_rev0; // their first turn
  hipEvent_t *evpool; // 6 per turn of a stats chunk when timing
  double kt_ms[4];
  long kt_launches[4];
  int timing;
};

// Calls that track or change the ring's state refuse to run while an
// asynchronous block is pending (its statistics would be overwritten).
static int no_pending(const mbt_ring *r)
{
  if (r->async_n) {
    mbt_set_error("%ld turns enqueued by mbt_ring_track_async are pending: collect them with mbt_ring_track_wait",
                  r->async_n);
    return MBT_EINVAL;
  }
  return MBT_OK;
}

static size_t turn_lds(int Nc)
{
  return (size_t)(4 * Nc + 8) * 8 + (size_t)3 * Nc * 8 + (size_t)(Nc + 2) * 4 + 16;
}

static void fill_params(const mbt_config *c, const int planes[3], DevParams &P)
{
  memset(&P, 0, sizeof P);
  P.fx_lim = c->Np > 128 ? 134217728.0 / (double)c->Np : 1048576.0;
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
  P.scan = c->scan;
  P.NrevScan = c->NrevScan;
  P.scan_step = c->scan_step;
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
  int lr_tot = c->n_lr[0] + c->n_lr[1] + c->n_lr[2];
  P.sf_active = (planes[0] + planes[1] + planes[2] + lr_tot) > 0 && c->Ncell > 0;
  P.kickL = (c->n_lr[0] + planes[0]) > 0;
  P.convL = planes[0] > 0;
  P.kickV = planes[2] > 0;
  P.kickH = planes[1] > 0;
  P.fb_len = c->has_rf_feedback ? abs(c->rffb_len_average) : 1;
  P.fb_switch_off = c->has_rf_feedback && c->rffb_len_average < 0;
  P.vrf_design = c->rffb_vrf_design;
  P.phi0_design = c->rffb_phi0_design;
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

// uploads are ordered on the ring's stream (the host buffers live until the
// stream is synchronised below)
#define UP(dst, src, n) CK(hipMemcpyAsync((dst), (src), (size_t)(n) * sizeof(*(src)), hipMemcpyHostToDevice, r->stream))

static void copy_cfg(RingHost *h, const mbt_config *c)
{
  h->cfg = *c;
  int H = (int)c->Nharm, Nb = c->Nbunch;
  h->nfFill = (int *)malloc(H * sizeof(int));
  memcpy(h->nfFill, c->nfFill, H * sizeof(int));
  h->Ibunch = (double *)malloc(H * sizeof(double));
  memcpy(h->Ibunch, c->Ibunch, H * sizeof(double));
  h->bucket = (int *)malloc(Nb * sizeof(int));
  memcpy(h->bucket, c->bucket, Nb * sizeof(int));
  h->bunch_Ib = (double *)malloc(Nb * sizeof(double));
  memcpy(h->bunch_Ib, c->bunch_Ib, Nb * sizeof(double));
  h->Ib_frac = (double *)malloc(Nb * sizeof(double));
  memcpy(h->Ib_frac, c->Ib_frac, Nb * sizeof(double));
  h->bunch_out = (int *)malloc((c->Nbunch_out > 0 ? c->Nbunch_out : 1) * sizeof(int));
  if (c->Nbunch_out > 0) memcpy(h->bunch_out, c->bunch_out, c->Nbunch_out * sizeof(int));
  h->cfg.nfFill = h->nfFill;
  h->cfg.Ibunch = h->Ibunch;
  h->cfg.bucket = h->bucket;
  h->cfg.bunch_Ib = h->bunch_Ib;
  h->cfg.Ib_frac = h->Ib_frac;
  h->cfg.bunch_out = h->bunch_out;
}

extern "C" void mbt_ring_destroy(mbt_ring *r)
{
  if (!r) return;
  (void)hipSetDevice(r->device);
  if (r->stream) (void)hipStreamSynchronize(r->stream); // turns of mbt_ring_track_async may still run
  if (r->pot_host) (void)hipHostFree(r->pot_host);
  void *ptrs[] = {r->C.s, r->wg_first, r->wg_end, r->bunch_first,
                  r->wg_bunch, r->bunch_wg0, r->bucket, r->bunch_ring, r->bucket_ring, r->np, r->nfFill,
                  r->lr_planes, r->P, r->partial, r->shift, r->Hs.cnt, r->Hs.sx, r->Hs.sz, r->Hs.tcnt, r->Hs.tsx,
                  r->Hs.tsz, r->Hs.bad, r->sfL, r->sfV, r->sfH, r->lr_wake, r->fnp_ring, r->lrc, r->Gf, r->pending, r->peek,
                  r->rwhist, r->sqrt_tt, r->rsqrt_tt, r->Ibunch, r->bunch_Ib, r->stats, r->amp, r->rfbuf, r->fnp_out,
                  r->dipV_out, r->dipH_out, r->trash_tot, r->trash_pend, r->ts, r->rw_tiles, r->rw_part, r->ring_done, r->lc, r->lr_first, r->phasor_end, r->fb_hist,
                  r->rwg_col_ring, r->rwg_bucket_bunch, r->rwg_c, r->rwg_part};
  for (size_t i = 0; i < sizeof ptrs / sizeof ptrs[0]; i++)
    if (ptrs[i]) (void)hipFree(ptrs[i]);
  if (r->evpool) {
    for (int i = 0; i < 6 * STATS_TURNS; i++) (void)hipEventDestroy(r->evpool[i]);
    free(r->evpool);
  }
  for (int i = 0; i < NGRAPH; i++)
    if (r->graphs[i].exec) (void)hipGraphExecDestroy(r->graphs[i].exec);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  if (r->rh)
    for (int k = 0; k < r->R; k++) {
      RingHost *h = &r->rh[k];
      free(h->nfFill);
      free(h->Ibunch);
      free(h->bucket);
      free(h->bunch_Ib);
      free(h->Ib_frac);
      free(h->bunch_out);
    }
  free(r->rh);
  free(r->Ph);
  free(r);
}

static int check_cfg(const mbt_config *c)
{
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
  int lr_tot = c->n_lr[0] + c->n_lr[1] + c->n_lr[2];
  if ((lr_tot > 0 || c->n_res > 0 || c->n_wakefile > 0) && c->Ncell <= 0) {
    mbt_set_error("wakes need Ncell > 0");
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
  return MBT_OK;
}

static long long padded_np(long long Np) { return (Np + BUNCH_ALIGN - 1) / BUNCH_ALIGN * BUNCH_ALIGN; }

static int wg_count(const mbt_ring *r, int items)
{
  const long long CH = (long long)WG * items;
  long long n = 0;
  for (int k = 0; k < r->R; k++) n += ((long long)r->rh[k].cfg.Np + CH - 1) / CH * r->rh[k].cfg.Nbunch;
  return (int)n;
}

static void drop_graphs(mbt_ring *r)
{
  for (int i = 0; i < NGRAPH; i++)
    if (r->graphs[i].exec) {
      (void)hipGraphExecDestroy(r->graphs[i].exec);
      r->graphs[i].exec = nullptr;
    }
}

// Work-group decomposition: every bunch cut into contiguous chunks of
// WG * items macroparticles, one work-group each.
static int set_decomposition(mbt_ring *r, int items)
{
  if (items < r->min_items || items > MAX_CHUNK_ITEMS) {
    mbt_set_error("chunk of %d items per lane outside [%d, %d]", items, r->min_items, MAX_CHUNK_ITEMS);
    return MBT_EINVAL;
  }
  const int nwg = wg_count(r, items), G = r->G;
  const long long CH = (long long)WG * items;
  long long *wf = (long long *)malloc(nwg * sizeof(long long)), *we = (long long *)malloc(nwg * sizeof(long long));
  int *wb = (int *)malloc(nwg * sizeof(int)), *b0 = (int *)malloc((G + 1) * sizeof(int));
  int w = 0, g = 0;
  long long pos = 0;
  for (int k = 0; k < r->R; k++) {
    const long long Np = r->rh[k].cfg.Np, per_b = (Np + CH - 1) / CH;
    for (int bb = 0; bb < r->rh[k].cfg.Nbunch; bb++, g++) {
      b0[g] = w;
      for (long long kk = 0; kk < per_b; kk++, w++) {
        wf[w] = pos + kk * CH;
        we[w] = pos + ((kk + 1) * CH < Np ? (kk + 1) * CH : Np);
        wb[w] = g;
      }
      pos += padded_np(Np);
    }
  }
  b0[G] = w;
  int rc = MBT_OK;
  if (hipMemcpyAsync(r->wg_first, wf, nwg * sizeof(long long), hipMemcpyHostToDevice, r->stream) != hipSuccess ||
      hipMemcpyAsync(r->wg_end, we, nwg * sizeof(long long), hipMemcpyHostToDevice, r->stream) != hipSuccess ||
      hipMemcpyAsync(r->wg_bunch, wb, nwg * sizeof(int), hipMemcpyHostToDevice, r->stream) != hipSuccess ||
      hipMemcpyAsync(r->bunch_wg0, b0, (G + 1) * sizeof(int), hipMemcpyHostToDevice, r->stream) != hipSuccess ||
      hipStreamSynchronize(r->stream) != hipSuccess) {
    mbt_set_error("upload of the work-group decomposition failed");
    rc = MBT_EDEVICE;
  }
  free(wf);
  free(we);
  free(wb);
  free(b0);
  if (rc == MBT_OK) {
    r->nwg = nwg;
    r->chunk_items = items;
    drop_graphs(r); // cached graphs hold the old grid
  }
  return rc;
}

// The GEMM form of the RW sum: used when at least RWG_MIN_RINGS rings have the
// resistive wall on and all of them share h, Nmlt and T0 (one tap table).
static int setup_rw_gemm(mbt_ring *r)
{
  int nrw = 0, h = 0, Nmlt = 0;
  double T0 = 0.0;
  for (int k = 0; k < r->R; k++) {
    const mbt_config *c = &r->rh[k].cfg;
    if (!c->EnableRW_long) continue;
    if (nrw == 0) {
      h = (int)c->Nharm;
      Nmlt = c->Nmlt;
      T0 = c->T0;
    } else if ((int)c->Nharm != h || c->Nmlt != Nmlt || c->T0 != T0)
      return MBT_OK; // rings with different taps: the per-bunch sum
    nrw++;
  }
  if (nrw < (r->force_gemm ? 1 : RWG_MIN_RINGS)) return MBT_OK;
  RWGemm &g = r->rwg;
  g.h = h;
  g.M = (Nmlt + 1) * h;
  g.n = (Nmlt + 2) * h;
  g.K = ((h + g.M) + 15) / 16 * 16;
  g.HP = (h + 15) / 16 * 16;
  // K slices of up to RWG_PER steps, fitting the block's Y slice and taps in LDS
  g.per = RWG_PER;
  while (g.per > 4 && rwg_lds(g.per, g.HP) > RWG_LDS_MAX) g.per /= 2;
  if (rwg_lds(g.per, g.HP) > RWG_LDS_MAX) return MBT_OK; // h too large: the per-bunch sum
  g.KS = (g.K / 4 + g.per - 1) / g.per;
  g.N = 2 * nrw;
  int *cr = (int *)malloc(sizeof(int) * nrw), *bb = (int *)malloc(sizeof(int) * r->H);
  double *c = (double *)malloc(sizeof(double) * (g.M + 1));
  for (int k = 0, i = 0; k < r->R; k++)
    if (r->rh[k].cfg.EnableRW_long) cr[i++] = k;
  for (int q = 0; q < r->H; q++) bb[q] = -1;
  for (int k = 0, gg = 0; k < r->R; k++)
    for (int b = 0; b < r->rh[k].cfg.Nbunch; b++, gg++) bb[r->Ph[k].bucket_off + r->rh[k].cfg.bucket[b]] = gg;
  const double deltab = T0 / r->rh[cr[0]].cfg.h;
  c[0] = 0.0;
  for (long long m = 1; m <= g.M; m++) c[m] = 1.0 / sqrt(m * deltab);
  int rc = dalloc(&r->rwg_col_ring, nrw);
  if (!rc) rc = dalloc(&r->rwg_bucket_bunch, r->H);
  if (!rc) rc = dalloc(&r->rwg_c, (size_t)g.M + 1);
  if (!rc) rc = dalloc(&r->rwg_part, (size_t)g.KS * g.HP * g.N);
  if (!rc && (hipMemcpy(r->rwg_col_ring, cr, sizeof(int) * nrw, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(r->rwg_bucket_bunch, bb, sizeof(int) * r->H, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(r->rwg_c, c, sizeof(double) * (g.M + 1), hipMemcpyHostToDevice) != hipSuccess ||
              hipDeviceSynchronize() != hipSuccess))
    rc = MBT_EDEVICE;
  free(cr);
  free(bb);
  free(c);
  if (rc) return rc;
  g.col_ring = r->rwg_col_ring;
  g.bucket_bunch = r->rwg_bucket_bunch;
  g.c = r->rwg_c;
  g.part = r->rwg_part;
  r->rwg_on = 1;
  return MBT_OK;
}

// RW tiles of k_turn's leading blocks: per RW ring, groups of up to WG consecutive bunches times
// RW_TILE_TAPS consecutive taps.  Leaves rw_split off (the one-pass sum) when a
// tile's slot span does not fit the LDS budget.
static int setup_rw_split(mbt_ring *r)
{
  int ntile = 0, cmax = 0;
  size_t lds = 0;
  for (int pass = 0; pass < 2; pass++) {
    int4 *tl = pass ? (int4 *)malloc(sizeof(int4) * (ntile > 0 ? ntile : 1)) : NULL;
    int t = 0;
    for (int k = 0; k < r->R; k++) {
      const mbt_config *c = &r->rh[k].cfg;
      if (!r->Ph[k].rw) continue;
      const int h = (int)c->Nharm, M = (c->Nmlt + 1) * h;
      const int C = (M + RW_TILE_TAPS - 1) / RW_TILE_TAPS;
      if (C > cmax) cmax = C;
      for (int g0 = 0; g0 < c->Nbunch; g0 += WG) {
        const int nb = c->Nbunch - g0 < WG ? c->Nbunch - g0 : WG;
        const int span_b = c->bucket[g0 + nb - 1] - c->bucket[g0];
        for (int ct = 0; ct < C; ct++) {
          const int m0 = 1 + ct * RW_TILE_TAPS, m1 = m0 + RW_TILE_TAPS - 1 < M ? m0 + RW_TILE_TAPS - 1 : M;
          const size_t l = rw_old_lds(span_b + m1 - m0 + 1);
          if (l > lds) lds = l;
          if (pass) tl[t] = make_int4(k, r->Ph[k].bunch_off + g0, nb, ct);
          t++;
        }
      }
    }
    if (!pass) {
      ntile = t;
      if (ntile == 0 || lds > RW_OLD_LDS_MAX) {
        r->rw_split = 0;
        return MBT_OK;
      }
      continue;
    }
    int rc = dalloc(&r->rw_tiles, (size_t)ntile);
    if (!rc) rc = dalloc(&r->rw_part, (size_t)cmax * r->G * 2);
    if (!rc && hipMemcpy(r->rw_tiles, tl, sizeof(int4) * ntile, hipMemcpyHostToDevice) != hipSuccess) rc = MBT_EDEVICE;
    free(tl);
    if (rc) return rc;
  }
  r->rw_ntile = ntile;
  r->rw_lds = lds;
  return MBT_OK;
}

// Plain in-place update of the coordinate tiles, every bit written back (xor
// with a zero the compiler cannot see): the rate the slab's placement allows
// for k_turn's loads and stores, without its work.  One tile per work-group.
__global__ __launch_bounds__(WG) void k_stream(Coords C, unsigned long long zero)
{
  double *base = C.s + ((long long)blockIdx.x * 6 << TILE_SHIFT);
  for (int i = threadIdx.x; i < TILE; i += WG) {
    double v[6];
#pragma unroll
    for (int c = 0; c < 6; c++) v[c] = LDC(base + c * TILE + i);
#pragma unroll
    for (int c = 0; c < 6; c++)
      STC(__longlong_as_double(__double_as_longlong(v[c]) ^ (long long)zero), base + c * TILE + i);
  }
}

// Placement of the coordinates.  For working sets of tens of GB the six-array
// in-place stream of the turn pass runs at either 5.3-5.5 or 6.0-6.4 TB/s,
// depending on the allocation and not on the arrays' relative offsets; which
// one a fresh allocation gets depends on the allocator's state (on fresh
// devices, reserving and releasing the whole footprint first gave the fast
// placement 4 times in 4; later allocations in a process land either way).
// Working sets of a few GB stream at 5.4-5.6 TB/s wherever they lie (every
// 2 GB window of three 24 GB allocations measured 5.44-5.62), so only large
// footprints search: candidates are allocated while the device has room for
// one more (at most PLACE_CANDIDATES), each timed with k_stream, the first at
// PLACE_GOOD_GBS or above kept at once, else the fastest; the others are freed.
// The bar sits near the top of the fast placements' range, so that a fast but
// not the fastest first allocation (6.0 TB/s) is not kept without a look at
// the next ones.
#define PLACE_SEARCH_MIN (16ULL << 30)
#define PLACE_CANDIDATES 3
#define PLACE_GOOD_GBS 6300.0

// n slots (a multiple of TILE): the tiled slab of 6 n doubles
static int alloc_six(Coords &c, long long n)
{
  c.s = nullptr;
  if (hipMalloc((void **)&c.s, (size_t)n * 6 * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    c.s = nullptr;
    return MBT_ENOMEM;
  }
  return MBT_OK;
}

static void free_six(Coords &c)
{
  if (c.s) (void)hipFree(c.s);
  c = Coords{};
}

// GB/s of k_stream over one candidate's tiles (one untimed pass, one timed)
static double time_six(hipStream_t st, const Coords &c, long long n)
{
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return 0.0;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return 0.0;
  }
  const unsigned nb = (unsigned)(n >> TILE_SHIFT);
  hipLaunchKernelGGL(k_stream, dim3(nb), dim3(WG), 0, st, c, 0ULL);
  (void)hipEventRecord(e0, st);
  hipLaunchKernelGGL(k_stream, dim3(nb), dim3(WG), 0, st, c, 0ULL);
  (void)hipEventRecord(e1, st);
  float ms = 0.0f;
  double gbps = 0.0;
  if (hipGetLastError() == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
      hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.0f)
    gbps = 96.0 * (double)n / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return gbps;
}

static int place_coords(mbt_ring *r, long long n)
{
  const size_t foot = (size_t)n * 48;
  if (foot >= COORD_RESERVE_MIN) {
    void *reserve = NULL;
    if (hipMalloc(&reserve, foot) == hipSuccess)
      (void)hipFree(reserve);
    else
      (void)hipGetLastError(); // not enough memory to reserve: allocate the arrays as they come
  }
  Coords cand[PLACE_CANDIDATES];
  double rate[PLACE_CANDIDATES];
  int nc = 0, best = 0;
  for (int k = 0; k < PLACE_CANDIDATES; k++) {
    if (k > 0) {
      size_t fr = 0, tot = 0;
      if (foot < PLACE_SEARCH_MIN || hipMemGetInfo(&fr, &tot) != hipSuccess || fr < foot + (4ULL << 30)) break;
    }
    if (alloc_six(cand[k], n) != MBT_OK) {
      if (k == 0) {
        mbt_set_error("hipMalloc of the coordinates (%zu bytes) failed", foot);
        return MBT_ENOMEM;
      }
      break;
    }
    nc = k + 1;
    if (foot < PLACE_SEARCH_MIN) break;
    rate[k] = time_six(r->stream, cand[k], n);
    if (rate[k] > rate[best]) best = k;
    if (rate[k] >= PLACE_GOOD_GBS) break;
  }
  for (int k = 0; k < nc; k++)
    if (k != best) free_six(cand[k]);
  r->C = cand[best];
  r->place_candidates = nc;
  r->place_gbps = nc > 1 || foot >= PLACE_SEARCH_MIN ? rate[best] : 0.0;
  CK(hipMemsetAsync(r->C.s, 0, (size_t)n * 6 * sizeof(double), r->stream));
  CK(hipStreamSynchronize(r->stream));
  return MBT_OK;
}

// ---- initial distributions on the device -------------------------------------
// The reference's Gaussian bunch generator (weak_generate_bunch_distribution,
// src/bunch_weak.c:85-105, with c_fnorm, src/bunch.c:379-405) restated for the
// device, bit for bit: per plane, glibc's TYPE_3 rand() stream seeded with
// iseed[plane] + bucket, each normal the sum of 12 uniforms rand() / RAND_MAX
// minus 1/2, the position ps * n1 + offset and the slope so + ss * n2 + cr * ps * n1.
// The reference runs one MPI rank per bunch whose generator carries over from
// plane to plane; it reseeds at each plane whenever the planes' seeds differ,
// so with pairwise different seeds every (bunch, plane) stream starts fresh and
// one lane generates it.  The 31-word additive-feedback table of each lane lives
// in LDS (an odd stride of words per lane, so lanes reading the same index hit
// different banks).
struct GenRing {
  int on;       // the device generator covers this ring
  int track[3]; // planes generated (LON always)
  int iseed[3];
  double po[3], ps[3], so[3], ss[3], cr[3];
};

#define GEN_WG 64
#define GEN_TAB 33

__device__ __forceinline__ int gen_next(int *t, int &f, int &b)
{
  const unsigned v = (unsigned)t[f] + (unsigned)t[b];
  t[f] = (int)v;
  f = f == 30 ? 0 : f + 1;
  b = b == 30 ? 0 : b + 1;
  return (int)(v >> 1);
}

// srandom_r for TYPE_3: Park-Miller minimal standard words (Schrage's
// decomposition, signed 32-bit word), front index 3 words ahead of the rear,
// the first 310 outputs discarded
__device__ void gen_srand(int *t, int &f, int &b, unsigned seed)
{
  if (seed == 0) seed = 1;
  int word = (int)seed;
  t[0] = word;
  for (int i = 1; i < 31; i++) {
    const long long hi = word / 127773, lo = word % 127773;
    long long w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    word = (int)w;
    t[i] = word;
  }
  f = 3;
  b = 0;
  for (int k = 0; k < 310; k++) (void)gen_next(t, f, b);
}

__device__ __forceinline__ double gen_fnorm(int *t, int &f, int &b)
{
  double sum = 0.0;
  for (int i = 0; i < 12; i++) {
    const double u = ((double)gen_next(t, f, b)) / ((double)2147483647);
    const double v = u - 0.5;
    sum = sum + v;
  }
  return sum;
}

// one lane per (bunch, plane)
__global__ __launch_bounds__(GEN_WG) void k_gen(const GenRing *gr, const int *bunch_ring, const int *bucket,
                                                const long long *bunch_first, const int *np, Coords C, int G)
{
  __shared__ int tabs[GEN_WG * GEN_TAB];
  const int i = blockIdx.x * GEN_WG + threadIdx.x;
  if (i >= 3 * G) return;
  const int g = i / 3, p = i % 3;
  const GenRing &R = gr[bunch_ring[g]];
  if (!R.on || !R.track[p]) return;
  int *t = tabs + threadIdx.x * GEN_TAB;
  int f, b;
  gen_srand(t, f, b, (unsigned)(R.iseed[p] + bucket[g]));
  const long long first = bunch_first[g];
  const int Np = np[g];
  const double po = R.po[p], ps = R.ps[p], so = R.so[p], ss = R.ss[p], cr = R.cr[p];
  for (int j = 0; j < Np; j++) {
    const double tp = ps * gen_fnorm(t, f, b);
    C.s[coord_idx(first + j, 2 * p)] = po + tp;
    const double n2 = gen_fnorm(t, f, b);
    C.s[coord_idx(first + j, 2 * p + 1)] = so + ss * n2 + cr * tp;
  }
}

// Whether the device generator reproduces the ring's initial distribution:
// Gaussian planes only (no distribution files, no head-tail mode excitation),
// more than one macroparticle, pairwise different seeds of the planes it makes.
static int gen_on_device(const mbt_config *c)
{
  if (c->Np <= 1 || !c->nGen[0]) return 0;
  for (int p = 1; p < 3; p++)
    if (c->TrackPlane[p] && (!c->nGen[p] || (p == 2 && c->mode_excitation))) return 0;
  if (c->TrackPlane[1] && c->iseed[1] == c->iseed[0]) return 0;
  if (c->TrackPlane[2] && (c->iseed[2] == c->iseed[0] || (c->TrackPlane[1] && c->iseed[2] == c->iseed[1]))) return 0;
  return 1;
}

static int create_rings(const mbt_config *const *cfgs, int R, int device, int flags, mbt_ring **out)
{
  *out = NULL;
  if (R <= 0) {
    mbt_set_error("no ring");
    return MBT_EINVAL;
  }
  for (int k = 0; k < R; k++) {
    int rc = check_cfg(cfgs[k]);
    if (rc) return rc;
    if (cfgs[k]->Ncell != cfgs[0]->Ncell || cfgs[k]->NrevMon != cfgs[0]->NrevMon) {
      mbt_set_error("rings of one ensemble share Ncell and NrevMon");
      return MBT_EINVAL;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device) {
    mbt_set_error("no HIP device %d", device);
    return MBT_EDEVICE;
  }
  mbt_ring *r = (mbt_ring *)calloc(1, sizeof(mbt_ring));
  r->R = R;
  r->device = device;
  r->rh = (RingHost *)calloc(R, sizeof(RingHost));
  r->Ph = (DevParams *)calloc(R, sizeof(DevParams));
  r->mixed = (flags & MBT_RING_MIXED) ? 1 : 0;
  r->force_gemm = (flags & MBT_RING_RW_GEMM) ? 1 : 0;
  for (int k = 0; k < R; k++) copy_cfg(&r->rh[k], cfgs[k]);
  CK(hipSetDevice(device));
  // a blocking stream: ordered after work on the null stream (allocation-time
  // memsets) like every copy the library makes
  CK(hipStreamCreate(&r->stream));
  const int Nc = cfgs[0]->Ncell, NcA = Nc > 0 ? Nc : 1;
  r->Nc = Nc;
  r->nrev_mon = (int)cfgs[0]->NrevMon;
  // ---- per-ring parameters, Green functions, offsets
  double *Gf = (double *)calloc((size_t)R * 3 * NcA, sizeof(double));
  int G = 0, H = 0;
  long long Ntot = 0, Nalloc = 0, rw_total = 0, st_total = 0, fbh_total = 0;
  for (int k = 0; k < R; k++) {
    RingHost *hh = &r->rh[k];
    const mbt_config *c = &hh->cfg;
    int rc = mbt_greens_functions(c, Gf + (size_t)k * 3 * NcA, Gf + (size_t)k * 3 * NcA + NcA,
                                  Gf + (size_t)k * 3 * NcA + 2 * NcA, hh->planes);
    if (rc) {
      free(Gf);
      mbt_ring_destroy(r);
      return rc;
    }
    DevParams &P = r->Ph[k];
    fill_params(c, hh->planes, P);
    P.mixed = (flags & MBT_RING_MIXED) ? 1 : 0;
    P.ring = k;
    P.bunch_off = G;
    P.bucket_off = H;
    P.rw_off = rw_total;
    P.st_off = st_total;
    P.fbh_off = fbh_total;
    G += c->Nbunch;
    H += (int)c->Nharm;
    Ntot += (long long)c->Np * c->Nbunch;
    Nalloc += padded_np(c->Np) * c->Nbunch;
    if (c->EnableRW_long) {
      if ((long long)(c->Nmlt + 2) * c->Nharm >= (1LL << 30)) {
        // the RW sum kernels index the history with 32-bit integers
        mbt_set_error("resistive-wall history of %lld slots too long", (long long)(c->Nmlt + 2) * c->Nharm);
        free(Gf);
        mbt_ring_destroy(r);
        return MBT_EINVAL;
      }
      rw_total += 2LL * (c->Nmlt + 2) * c->Nharm;
      st_total += (long long)(c->Nmlt + 1) * c->Nharm + 1;
    }
    fbh_total += 2LL * P.fb_len;
    hh->lr_plane[0] = c->n_lr[0] > 0;
    hh->lr_plane[1] = c->n_lr[1] > 0 && c->TrackPlane[1];
    hh->lr_plane[2] = c->n_lr[2] > 0 && c->TrackPlane[2];
    r->any_qe |= P.qe;
    r->any_dip |= P.kickV || P.kickH || hh->lr_plane[1] || hh->lr_plane[2];
    r->any_gen |= P.ideal_hc || P.n_ahc > 0 || P.amp_dq;
    r->sf_any |= P.sf_active;
  }
  r->G = G;
  r->H = H;
  r->Ntot = Ntot;
  r->Nalloc = Nalloc;
  r->rw_total = rw_total;
  r->st_total = st_total;
  r->fbh_total = fbh_total;
  // ---- workgroup decomposition (contiguous chunks of one bunch): buffers sized
  // for the smallest chunk, the decomposition itself set below
  r->min_items = Ntot < SMALL_RING_MP ? 1 : MIN_CHUNK_ITEMS;
  const int nwg = wg_count(r, r->min_items);
  r->nwg_cap = nwg;
  int *nph = (int *)malloc(G * sizeof(int)), *bring = (int *)malloc(G * sizeof(int)), *bk = (int *)malloc(G * sizeof(int));
  long long *bfirst = (long long *)malloc(G * sizeof(long long));
  int *qring = (int *)malloc(H * sizeof(int)), *fill = (int *)malloc(H * sizeof(int));
  double *Ibq = (double *)malloc(H * sizeof(double)), *bIb = (double *)malloc(G * sizeof(double));
  {
    int g = 0, q = 0;
    long long pos = 0;
    for (int k = 0; k < R; k++) {
      const mbt_config *c = &r->rh[k].cfg;
      const long long Np = c->Np;
      for (int i = 0; i < (int)c->Nharm; i++, q++) {
        qring[q] = k;
        fill[q] = c->nfFill[i];
        Ibq[q] = c->Ibunch[i];
      }
      for (int bb = 0; bb < c->Nbunch; bb++, g++) {
        nph[g] = (int)Np;
        bring[g] = k;
        bk[g] = c->bucket[bb];
        bIb[g] = c->bunch_Ib[bb];
        bfirst[g] = pos;
        pos += padded_np(Np);
      }
    }
  }
  {
    const int rcp = place_coords(r, Nalloc);
    if (rcp) {
      mbt_ring_destroy(r);
      return rcp;
    }
  }
  DA(r->wg_first, nwg);
  DA(r->wg_end, nwg);
  DA(r->wg_bunch, nwg);
  DA(r->bunch_wg0, G + 1);
  DA(r->bunch_first, G);
  DA(r->bucket, G);
  DA(r->bunch_ring, G);
  DA(r->bucket_ring, H);
  DA(r->np, G);
  DA(r->nfFill, H);
  DA(r->lr_planes, 3);
  DA(r->P, R);
  DA(r->partial, (size_t)nwg * NMOM);
  DA(r->shift, (size_t)G * 6);
  DA(r->Hs.cnt, (size_t)G * NcA);
  DA(r->Hs.sx, (size_t)G * NcA * 2);
  DA(r->Hs.sz, (size_t)G * NcA * 2);
  DA(r->Hs.tcnt, (size_t)G * 2);
  DA(r->Hs.tsx, (size_t)G * 4);
  DA(r->Hs.tsz, (size_t)G * 4);
  DA(r->Hs.bad, G);
  DA(r->sfL, (size_t)G * NcA);
  DA(r->sfV, (size_t)G * NcA);
  DA(r->sfH, (size_t)G * NcA);
  DA(r->lr_wake, (size_t)3 * H * NcA);
  DA(r->fnp_ring, (size_t)3 * H * NcA);
  DA(r->Gf, (size_t)R * 3 * NcA);
  DA(r->pending, (size_t)G * 2);
  {
    long long npmax = 1;
    for (int k = 0; k < R; k++)
      if (r->rh[k].cfg.Np > npmax) npmax = r->rh[k].cfg.Np;
    DA(r->peek, (size_t)6 * npmax);
  }
  DA(r->rwhist, rw_total > 0 ? rw_total : 2);
  DA(r->sqrt_tt, st_total > 0 ? st_total : 1);
  DA(r->rsqrt_tt, st_total > 0 ? st_total : 1);
  DA(r->Ibunch, H);
  DA(r->bunch_Ib, G);
  DA(r->stats, (size_t)STATS_TURNS * G * 12);
  DA(r->amp, (size_t)STATS_TURNS * G * 6);
  DA(r->rfbuf, (size_t)STATS_TURNS * R * 2);
  DA(r->fnp_out, (size_t)G * NcA);
  DA(r->dipV_out, (size_t)G * NcA);
  DA(r->dipH_out, (size_t)G * NcA);
  if (hipHostMalloc((void **)&r->pot_host, sizeof(double) * 7 * NcA + sizeof(TurnState)) != hipSuccess) {
    r->pot_host = NULL;
    mbt_set_error("hipHostMalloc of the potentials staging buffer failed");
    mbt_ring_destroy(r);
    return MBT_ENOMEM;
  }
  DA(r->trash_tot, (size_t)G * 2);
  DA(r->trash_pend, (size_t)G * 2);
  DA(r->ts, R);
  DA(r->ring_done, R);
  DA(r->fb_hist, fbh_total > 0 ? fbh_total : 2);
  // ---- long-range resonator constants (host libm, reference formulas)
  int nres = 0;
  for (int k = 0; k < R; k++)
    for (int p = 0; p < 3; p++)
      if (r->rh[k].lr_plane[p]) nres += r->rh[k].cfg.n_lr[p];
  LRConst *lc = (LRConst *)calloc(nres > 0 ? nres : 1, sizeof(LRConst));
  {
    int act[3] = {0, 0, 0}, na = 0, pl[3] = {0, 0, 0};
    for (int k = 0; k < R; k++)
      for (int p = 0; p < 3; p++)
        if (r->rh[k].lr_plane[p] && r->rh[k].cfg.n_lr[p] > 0) act[p] = 1;
    for (int p = 0; p < 3; p++)
      if (act[p]) pl[na++] = p;
    r->lr_planes_active = make_int4(na, pl[0], pl[1], pl[2]);
  }
  int *lrf = (int *)calloc(R + 1, sizeof(int));
  long long lrc_total = 0;
  {
    int ir = 0;
    for (int k = 0; k < R; k++) {
      lrf[k] = ir;
      const mbt_config *cf = &r->rh[k].cfg;
      for (int p = 0; p < 3; p++) {
        if (!r->rh[k].lr_plane[p]) continue;
        for (int l = 0; l < cf->n_lr[p]; l++) {
          const mbt_lr_resonator *res = &cf->lr[p][l];
          LRConst &L = lc[ir++];
          L.ring = k;
          L.plane = p;
          L.l = l;
          L.h = (int)cf->Nharm;
          if (L.h > r->lr_hmax) r->lr_hmax = L.h;
          L.bucket_off = r->Ph[k].bucket_off;
          L.lrc_off = lrc_total;
          L.init_turns = (int)(cf->Nbumax / cf->h);
          lrc_total += (long long)cf->Nharm * NcA;
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
    }
  }
  lrf[R] = nres;
  r->nres = nres;
  DA(r->lc, nres > 0 ? nres : 1);
  DA(r->lr_first, R + 1);
  UP(r->lr_first, lrf, R + 1);
  DA(r->lrc, lrc_total > 0 ? lrc_total : 1);
  DA(r->phasor_end, (size_t)2 * (nres > 0 ? nres : 1));
  // ---- uploads
  UP(r->bunch_first, bfirst, G);
  UP(r->np, nph, G);
  UP(r->bunch_ring, bring, G);
  UP(r->bucket, bk, G);
  UP(r->bucket_ring, qring, H);
  UP(r->nfFill, fill, H);
  UP(r->Ibunch, Ibq, H);
  UP(r->bunch_Ib, bIb, G);
  UP(r->Gf, Gf, (size_t)R * 3 * NcA);
  UP(r->P, r->Ph, R);
  if (nres > 0) UP(r->lc, lc, nres);
  {
    int lp[3] = {0, 0, 0};
    for (int k = 0; k < R; k++)
      for (int p = 0; p < 3; p++) lp[p] |= r->rh[k].lr_plane[p];
    UP(r->lr_planes, lp, 3);
  }
  if (st_total > 0) {
    double *st = (double *)malloc(st_total * sizeof(double));
    for (int k = 0; k < R; k++) {
      const mbt_config *cf = &r->rh[k].cfg;
      if (!cf->EnableRW_long) continue;
      long long M = (long long)(cf->Nmlt + 1) * cf->Nharm;
      double deltab = cf->T0 / cf->h;
      double *s = st + r->Ph[k].st_off;
      s[0] = 1.0;
      for (long long m = 1; m <= M; m++) s[m] = sqrt(m * deltab);
    }
    UP(r->sqrt_tt, st, st_total);
    CK(hipStreamSynchronize(r->stream));
    for (long long i = 0; i < st_total; i++) st[i] = 1.0 / st[i];
    UP(r->rsqrt_tt, st, st_total);
    CK(hipStreamSynchronize(r->stream));
    free(st);
  }
  {
    TurnState *ts0 = (TurnState *)calloc(R, sizeof(TurnState));
    for (int k = 0; k < R; k++) {
      const mbt_config *cf = &r->rh[k].cfg;
      ts0[k].phai0 = cf->phai0;
      ts0[k].Vrf0 = cf->Vrf0;
      ts0[k].phai0_prev = cf->phai0;
      ts0[k].Vrf0_prev = cf->Vrf0;
      ts0[k].Iring = cf->Iring;
      ts0[k].scan_val = cf->scan_start;
    }
    UP(r->ts, ts0, R);
    CK(hipStreamSynchronize(r->stream));
    free(ts0);
  }
  free(nph);
  free(bring);
  free(bk);
  free(bfirst);
  free(qring);
  free(fill);
  free(Ibq);
  free(bIb);
  free(Gf);
  free(lc);
  free(lrf);
  RedOut &Ro = r->Ro;
  Ro.stats = r->stats;
  Ro.amp = r->amp;
  Ro.shift = r->shift;
  Ro.rwhist = r->rwhist;
  Ro.fnp_ring = r->fnp_ring;
  Ro.plane_stride = (long long)H * NcA;
  Ro.sfL = r->sfL;
  Ro.sfV = r->sfV;
  Ro.sfH = r->sfH;
  Ro.fnp_out = r->fnp_out;
  Ro.dipV_out = r->dipV_out;
  Ro.dipH_out = r->dipH_out;
  Ro.trash_tot = r->trash_tot;
  Ro.trash_pend = r->trash_pend;
  Ro.G = r->Gf;
  Ro.bunch_Ib = r->bunch_Ib;
  Ro.np = r->np;
  Ro.Gtot = G;
  Ro.rfbuf = r->rfbuf;
  Ro.R = R;
  r->use_graphs = !(flags & MBT_RING_NO_GRAPHS);
  r->lds_turn = turn_lds(Nc);
  r->lds_red = (size_t)6 * NcA * 8 + 64;
  if (r->lds_turn > 64 * 1024 || r->lds_red > 64 * 1024) {
    mbt_set_error("Ncell %d too large for the LDS histogram", Nc);
    mbt_ring_destroy(r);
    return MBT_EINVAL;
  }
  // ---- initial distributions: on the device where its generator covers the
  // ring, otherwise host threads (one glibc stream per bunch) and an upload
  if (!(flags & MBT_RING_NO_INIT_DIST)) {
    GenRing *gh = (GenRing *)calloc(R, sizeof(GenRing));
    int ndev = 0;
    for (int k = 0; k < R; k++) {
      const mbt_config *cf = &r->rh[k].cfg;
      GenRing &q = gh[k];
      q.on = !(flags & MBT_RING_HOST_INIT_DIST) && gen_on_device(cf);
      ndev += q.on;
      for (int p = 0; p < 3; p++) {
        q.track[p] = p == 0 ? 1 : cf->TrackPlane[p] != 0;
        q.iseed[p] = cf->iseed[p];
        q.po[p] = cf->pos_offset[p];
        q.ps[p] = cf->pos_sgm[p];
        q.so[p] = cf->slope_offset[p];
        q.ss[p] = cf->slope_sgm[p];
        q.cr[p] = cf->correlate[p];
      }
    }
    if (ndev > 0) {
      GenRing *gd = NULL;
      int rcg = dalloc(&gd, (size_t)R);
      if (!rcg && hipMemcpy(gd, gh, sizeof(GenRing) * R, hipMemcpyHostToDevice) != hipSuccess) rcg = MBT_EDEVICE;
      if (!rcg) {
        hipLaunchKernelGGL(k_gen, dim3((unsigned)((3LL * G + GEN_WG - 1) / GEN_WG)), dim3(GEN_WG), 0, r->stream, gd,
                           r->bunch_ring, r->bucket, r->bunch_first, r->np, r->C, G);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(r->stream) != hipSuccess) {
          mbt_set_error("initial distribution kernel failed");
          rcg = MBT_EDEVICE;
        }
      }
      if (gd) (void)hipFree(gd);
      if (rcg) {
        free(gh);
        mbt_ring_destroy(r);
        return rcg;
      }
    }
    for (int k = 0; k < R; k++) {
      if (gh[k].on) continue;
      const mbt_config *cf = &r->rh[k].cfg;
      const size_t per = (size_t)6 * cf->Np;
      double *buf = (double *)malloc(sizeof(double) * per * cf->Nbunch);
      if (!buf) {
        mbt_ring_destroy(r);
        return MBT_ENOMEM;
      }
      int rc = mbt_generate_all(cf, buf);
      for (int b = 0; b < cf->Nbunch && rc == MBT_OK; b++)
        rc = mbt_ensemble_set_bunch(r, r->Ph[k].bunch_off + b, buf + per * b);
      free(buf);
      if (rc) {
        free(gh);
        mbt_ring_destroy(r);
        return rc;
      }
    }
    free(gh);
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    mbt_set_error("device synchronisation after set-up failed");
    mbt_ring_destroy(r);
    return MBT_EDEVICE;
  }
  // default chunk: CHUNK_ITEMS, fewer items per lane where the ring has too few
  // macroparticles to give DEFAULT_WG_TARGET work-groups
  int items0 = CHUNK_ITEMS;
  while (items0 > r->min_items && wg_count(r, items0) < DEFAULT_WG_TARGET) items0 /= 2;
  if (items0 < r->min_items) items0 = r->min_items;
  int rcd = set_decomposition(r, items0);
  if (!rcd && !(flags & MBT_RING_NO_RW_GEMM)) rcd = setup_rw_gemm(r);
  r->rw_split = !r->rwg_on && r->rw_total > 0 && (r->Ntot >= RW_SPLIT_MIN_MP || (flags & MBT_RING_RW_OVERLAP));
  if (!rcd && r->rw_split) rcd = setup_rw_split(r);
  if (rcd) {
    mbt_ring_destroy(r);
    return rcd;
  }
  *out = r;
  return MBT_OK;
}

extern "C" int mbt_ring_create(const mbt_config *c, int device, int flags, mbt_ring **out)
{
  const mbt_config *one[1] = {c};
  return create_rings(one, 1, device, flags, out);
}

extern "C" int mbt_ensemble_create(const mbt_config *const *cfgs, int nrings, int device, int flags, mbt_ring **out)
{
  return create_rings(cfgs, nrings, device, flags, out);
}

static long long bunch_first_host(const mbt_ring *r, int g, int *np, int *ring = NULL)
{
  long long pos = 0;
  for (int k = 0; k < r->R; k++) {
    const mbt_config *c = &r->rh[k].cfg;
    if (g < c->Nbunch) {
      *np = c->Np;
      if (ring) *ring = k;
      return pos + (long long)g * padded_np(c->Np);
    }
    g -= c->Nbunch;
    pos += (long long)c->Nbunch * padded_np(c->Np);
  }
  *np = 0;
  return -1;
}

// One bunch's coordinates out of the tiles in SoA order, with the resistive-wall
// kick of the last turn (src/transform_weak.c:780-833), which the next pass
// applies first, added the way that pass adds it: the bunch as it stands at the
// end of the turn.  Grid (ceil(Np / 256), 6).
__global__ __launch_bounds__(256) void k_peek(Coords C, long long first, int Np, const double *pending, int g,
                                              int rwh, int rwv, double *out)
{
  const int c = blockIdx.y;
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= Np) return;
  double v = C.s[coord_idx(first + j, c)];
  if (c == 3 && rwh) v += pending[2 * g];
  if (c == 5 && rwv) v += pending[2 * g + 1];
  out[(size_t)c * Np + j] = v;
}

extern "C" int mbt_ensemble_set_bunch(mbt_ring *r, int g, const double *s)
{
  if (g < 0 || g >= r->G) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  int Np;
  const long long first = bunch_first_host(r, g, &Np);
  // the bunch starts on a tile: whole tiles as one strided copy per coordinate, then the rest
  const long long nft = Np >> TILE_SHIFT, rem = Np & (TILE - 1);
  for (int c = 0; c < 6; c++) {
    if (nft)
      CK(hipMemcpy2DAsync(r->C.s + coord_idx(first, c), 6 * TILE * sizeof(double), s + (size_t)c * Np,
                          TILE * sizeof(double), TILE * sizeof(double), nft, hipMemcpyHostToDevice, r->stream));
    if (rem)
      CK(hipMemcpyAsync(r->C.s + coord_idx(first + (nft << TILE_SHIFT), c), s + (size_t)c * Np + (nft << TILE_SHIFT),
                        rem * sizeof(double), hipMemcpyHostToDevice, r->stream));
  }
  // the coordinates handed in are the bunch at the end of a turn: no kick of the
  // last turn is left to apply to them
  CK(hipMemsetAsync(r->pending + 2 * (size_t)g, 0, 2 * sizeof(double), r->stream));
  CK(hipStreamSynchronize(r->stream));
  return MBT_OK;
}

extern "C" int mbt_ensemble_get_bunch(const mbt_ring *r, int g, double *s)
{
  if (g < 0 || g >= r->G) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  int Np, ring;
  const long long first = bunch_first_host(r, g, &Np, &ring);
  if (Np <= 0) return MBT_OK;
  const DevParams &P = r->Ph[ring];
  (void)hipGetLastError(); // the launch check below must not report an earlier call's error
  hipLaunchKernelGGL(k_peek, dim3((unsigned)((Np + 255) / 256), 6), dim3(256), 0, r->stream, r->C, first, Np,
                     (const double *)r->pending, g, P.rw && P.track_h, P.rw && P.track_v, r->peek);
  CK(hipGetLastError());
  CK(hipMemcpyAsync(s, r->peek, (size_t)6 * Np * sizeof(double), hipMemcpyDeviceToHost, r->stream));
  CK(hipStreamSynchronize(r->stream));
  return MBT_OK;
}

extern "C" int mbt_ring_set_bunch(mbt_ring *r, int ib, const double *s) { return mbt_ensemble_set_bunch(r, ib, s); }
extern "C" int mbt_ring_get_bunch(const mbt_ring *r, int ib, double *s) { return mbt_ensemble_get_bunch(r, ib, s); }

typedef void (*turn_kernel_t)(const DevParams *, Coords, const long long *, const long long *, const int *,
                              const int *, const int *, const long long *, const double *, const double *,
                              const double *, const double *, const double *, long long, const double *, double *,
                              HistPtr, const TurnState *, RwOld);

#define TK(Q, D, G, M) k_turn<MODE_TURN, Q, D, G, M>
static const turn_kernel_t g_turn_kernels[16] = {
    TK(false, false, false, false), TK(false, false, false, true), TK(false, false, true, false),
    TK(false, false, true, true),   TK(false, true, false, false), TK(false, true, false, true),
    TK(false, true, true, false),   TK(false, true, true, true),   TK(true, false, false, false),
    TK(true, false, false, true),   TK(true, false, true, false),  TK(true, false, true, true),
    TK(true, true, false, false),   TK(true, true, false, true),   TK(true, true, true, false),
    TK(true, true, true, true)};
#undef TK
#define TT(Q, D, G) k_turn<MODE_TUNE, Q, D, G, false>
static const turn_kernel_t g_tune_kernels[8] = {TT(false, false, false), TT(false, false, true), TT(false, true, false),
                                                TT(false, true, true),   TT(true, false, false), TT(true, false, true),
                                                TT(true, true, false),   TT(true, true, true)};
#undef TT

static int launch_turn_pass(mbt_ring *r, int mode)
{
  turn_kernel_t k;
  if (mode == MODE_STATS) k = k_turn<MODE_STATS, false, true, true, false>;
  else if (mode == MODE_FINAL) k = k_turn<MODE_FINAL, false, true, true, false>;
  else {
    const bool mon = ((r->rev_host + 1) % r->nrev_mon) == 0;
    k = g_turn_kernels[(r->any_qe ? 8 : 0) + (r->any_dip ? 4 : 0) + (r->any_gen ? 2 : 0) + (mon ? 1 : 0)];
  }
  const long long stride = (long long)r->H * (r->Nc > 0 ? r->Nc : 1);
  // the turn pass also carries the resistive-wall tiles of earlier turns' rows
  RwOld ro = {};
  size_t lds = r->lds_turn;
  if (mode == MODE_TURN && r->rw_split) {
    ro = {r->rw_tiles, r->rw_ntile, (r->rw_ntile + 7) / 8 * 8, r->rwhist, r->rsqrt_tt, r->Ibunch, r->nfFill,
          r->rw_part, r->G};
    if (r->rw_lds > lds) lds = r->rw_lds;
  }
  hipLaunchKernelGGL(k, dim3(r->nwg + ro.nlead), dim3(WG), lds, r->stream, r->P, r->C, r->wg_first, r->wg_end,
                     r->wg_bunch, r->bunch_ring, r->bucket, r->bunch_first, r->shift, r->sfL, r->sfV, r->sfH,
                     r->lr_wake, stride, r->pending, r->partial, r->Hs, r->ts, ro);
  CK(hipGetLastError());
  return MBT_OK;
}

static int launch_reduce(mbt_ring *r, int phase)
{
  // the statistics passes (phases 0, 1, 3) always make the dipole sums; a turn
  // pass only when a ring's kicks or LR resonators need them
  const int dip = phase != 2 || r->any_dip;
  hipLaunchKernelGGL(k_reduce, dim3(r->G), dim3(WG), r->lds_red, r->stream, r->P, r->partial, r->bunch_wg0,
                     r->bunch_ring, r->bucket, r->Hs, r->Ro, r->ts, phase, r->lr_planes, dip);
  CK(hipGetLastError());
  return MBT_OK;
}

static int launch_lr(mbt_ring *r, int init)
{
  if (r->nres == 0) return MBT_OK;
  const long long stride = (long long)r->H * (r->Nc > 0 ? r->Nc : 1);
  const size_t lds = sizeof(double) * 6 * (size_t)r->lr_hmax;
  if (lds <= LR_PAR_LDS_MAX)
    hipLaunchKernelGGL(k_lr_par, dim3(r->nres), dim3(256), lds, r->stream, r->lc, r->P, r->fnp_ring, stride,
                       r->Ibunch, r->nfFill, r->Nc, r->phasor_end, r->lrc, r->fb_hist, r->ts, init);
  else
    hipLaunchKernelGGL(k_lr, dim3((r->nres + 63) / 64), dim3(64), 0, r->stream, r->nres, r->lc, r->P, r->fnp_ring,
                       stride, r->Ibunch, r->nfFill, r->Nc, r->phasor_end, r->lrc, r->fb_hist, r->ts, init);
  CK(hipGetLastError());
  if (!init && r->lr_planes_active.x > 0) {
    const long long tot = (long long)r->lr_planes_active.x * r->G * r->Nc;
    hipLaunchKernelGGL(k_lr_sum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, r->stream, r->lr_first, r->lc,
                       r->lrc, r->P, r->bunch_ring, r->bucket, r->G, r->Nc, r->lr_planes_active,
                       (long long)r->H * r->Nc, r->lr_wake);
    CK(hipGetLastError());
  }
  return MBT_OK;
}

static int launch_fb_solve(mbt_ring *r, int init)
{
  hipLaunchKernelGGL(k_fb_solve, dim3(r->R), dim3(64), 0, r->stream, r->R, r->P, r->ts, r->fb_hist, init);
  CK(hipGetLastError());
  return MBT_OK;
}

// Long-range recurrence and RF-feedback solve that precede a turn.
static int pre_turn(mbt_ring *r)
{
  int rc;
  if (r->sf_any && (rc = launch_lr(r, 0))) return rc;
  return launch_fb_solve(r, 0);
}

static int download_slots(mbt_ring *r, long long rev0, long n, double *stats, double *ampinv, double *rf)
{
  CK(hipStreamSynchronize(r->stream));
  const int G = r->G;
  // statistics and RF rows: one copy per contiguous run of ring-buffer slots (at
  // most two per block), not one per turn
  for (long t = 0; t < n;) {
    const int slot = (int)((rev0 + t) % STATS_TURNS);
    const long run = STATS_TURNS - slot < n - t ? STATS_TURNS - slot : n - t;
    if (stats)
      CK(hipMemcpy(stats + (size_t)t * G * 12, r->stats + (size_t)slot * G * 12,
                   (size_t)run * G * 12 * sizeof(double), hipMemcpyDeviceToHost));
    if (rf)
      CK(hipMemcpy(rf + (size_t)2 * r->R * t, r->rfbuf + (size_t)2 * r->R * slot,
                   (size_t)run * 2 * r->R * sizeof(double), hipMemcpyDeviceToHost));
    t += run;
  }
  // ampinv rows exist for the monitor turns only
  if (ampinv)
    for (long t = 0; t < n; t++) {
      if ((rev0 + t + 1) % r->nrev_mon != 0) continue;
      const int slot = (int)((rev0 + t) % STATS_TURNS);
      CK(hipMemcpy(ampinv + (size_t)t * G * 6, r->amp + (size_t)slot * G * 6, (size_t)G * 6 * sizeof(double),
                   hipMemcpyDeviceToHost));
    }
  return MBT_OK;
}

extern "C" int mbt_ring_prepare(mbt_ring *r, double *init_stats)
{
  {
    const int prc = no_pending(r);
    if (prc) return prc;
  }
  CK(hipSetDevice(r->device));
  (void)hipGetLastError(); // launch checks below must not report an earlier call's error
  int rc;
  // two stats passes: means (shift 0), then moments about the means + histograms
  CK(hipMemsetAsync(r->shift, 0, (size_t)r->G * 6 * sizeof(double), r->stream));
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  if ((rc = launch_reduce(r, 0))) return rc;
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  // the initial statistics land in slot (rev=0) and the RW history rows 0..Nmlt-1
  if ((rc = launch_reduce(r, 1))) return rc;
  if (init_stats) {
    CK(hipStreamSynchronize(r->stream));
    CK(hipMemcpy(init_stats, r->stats, (size_t)r->G * 12 * sizeof(double), hipMemcpyDeviceToHost));
  }
  // LR resonators: wake_phasor_init replays, then the RF-feedback design solve
  if (r->nres > 0) {
    TurnState *ts = (TurnState *)malloc(r->R * sizeof(TurnState));
    CK(hipStreamSynchronize(r->stream));
    CK(hipMemcpy(ts, r->ts, r->R * sizeof(TurnState), hipMemcpyDeviceToHost));
    for (int k = 0; k < r->R; k++) {
      ts[k].has_fb = r->rh[k].cfg.has_rf_feedback ? 1 : 0;
      ts[k].fb_count = 0;
    }
    CK(hipMemcpyAsync(r->ts, ts, r->R * sizeof(TurnState), hipMemcpyHostToDevice, r->stream));
    CK(hipStreamSynchronize(r->stream));
    free(ts);
    // each resonator replays its ring's Nbumax / h turns (none when that is 0)
    if ((rc = launch_lr(r, 1))) return rc;
    if ((rc = launch_fb_solve(r, 1))) return rc;
  }
  if ((rc = pre_turn(r))) return rc;
  CK(hipStreamSynchronize(r->stream));
  r->rev_host = 0;
  return MBT_OK;
}

// One turn's launches.  The reference runs the long-range wakes and the
// RF-feedback solve at the start of a turn; they run here at the end of the
// previous one (nothing changes their inputs in between).  With E, events
// E[0..4] bracket the kernel classes for per-kernel timing.
static int enqueue_turn(mbt_ring *r, hipEvent_t *E)
{
  int rc;
  const int lr_launch = r->sf_any && r->nres > 0;
  const int rw_split = r->rw_split;
  if (E) CK(hipEventRecord(E[0], r->stream));
  if ((rc = launch_turn_pass(r, MODE_TURN))) return rc;
  if (E) CK(hipEventRecord(E[1], r->stream));
  if ((rc = launch_reduce(r, 2))) return rc;
  if (E) CK(hipEventRecord(E[2], r->stream));
  if (r->rwg_on) {
    const dim3 tg((r->rwg.N + 15) / 16, r->rwg.KS);
    const size_t lds = rwg_lds(r->rwg.per, r->rwg.HP);
    if (r->mixed)
      hipLaunchKernelGGL(k_rwg_tile<true>, tg, dim3(256), lds, r->stream, r->rwg, r->P, r->ts, r->rwhist, r->Ibunch,
                         r->nfFill);
    else
      hipLaunchKernelGGL(k_rwg_tile<false>, tg, dim3(256), lds, r->stream, r->rwg, r->P, r->ts, r->rwhist, r->Ibunch,
                         r->nfFill);
    const long long tot = (long long)r->rwg.h * r->rwg.N;
    hipLaunchKernelGGL(k_rwg_sum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, r->stream, r->rwg, r->P, r->ts,
                       r->pending);
    CK(hipGetLastError());
    hipLaunchKernelGGL(k_end, dim3(r->R), dim3(64), 0, r->stream, r->P, r->ts, r->bunch_Ib, r->Ibunch, r->nfFill,
                       !lr_launch);
  } else if (rw_split) {
    hipLaunchKernelGGL(k_rw_fresh, dim3(r->G), dim3(WG), 0, r->stream, r->P, r->bunch_ring, r->rwhist, r->sqrt_tt,
                       r->Ibunch, r->nfFill, r->bucket, r->ts, r->rw_part, r->G, r->pending);
    CK(hipGetLastError());
    hipLaunchKernelGGL(k_end, dim3(r->R), dim3(64), 0, r->stream, r->P, r->ts, r->bunch_Ib, r->Ibunch, r->nfFill,
                       !lr_launch);
  } else if (r->rw_total > 0)
    hipLaunchKernelGGL(k_rw, dim3(r->G), dim3(RW_WG), 0, r->stream, r->P, r->bunch_ring, r->rwhist, r->sqrt_tt,
                       r->Ibunch, r->nfFill, r->bucket, r->pending, r->ts, r->ring_done, r->bunch_Ib, !lr_launch);
  else
    hipLaunchKernelGGL(k_end, dim3(r->R), dim3(64), 0, r->stream, r->P, r->ts, r->bunch_Ib, r->Ibunch, r->nfFill,
                       !lr_launch);
  CK(hipGetLastError());
  if (E) CK(hipEventRecord(E[3], r->stream));
  if (lr_launch) {
    if ((rc = launch_lr(r, 0))) return rc;
    if ((rc = launch_fb_solve(r, 0))) return rc;
  }
  if (E) CK(hipEventRecord(E[4], r->stream));
  r->rev_host++;
  return MBT_OK;
}

// n turns from the current turn as one HIP graph.  The launches of a turn
// depend on the turn number only through the monitor-turn variant of k_turn,
// and the monitor turns of a block follow from its first one, so a graph is
// reusable for any block of n turns whose first monitor turn sits at the same
// offset (or that has none); a few are cached.  (Timed runs launch eagerly:
// the runtime does not time events recorded inside a graph.)
static int turn_graph(mbt_ring *r, long n, hipGraphExec_t *out)
{
  long mon = (long)(r->nrev_mon - 1 - r->rev_host % r->nrev_mon);
  if (mon >= n) mon = -1;
  for (int i = 0; i < NGRAPH; i++)
    if (r->graphs[i].exec && r->graphs[i].mon == mon && r->graphs[i].n == n) {
      *out = r->graphs[i].exec;
      return MBT_OK;
    }
  GraphEntry &g = r->graphs[r->graph_next];
  r->graph_next = (r->graph_next + 1) % NGRAPH;
  if (g.exec) {
    CK(hipGraphExecDestroy(g.exec));
    g.exec = nullptr;
  }
  const long long rev_save = r->rev_host;
  CK(hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal));
  int rc = MBT_OK;
  for (long t = 0; t < n && rc == MBT_OK; t++) rc = enqueue_turn(r, NULL);
  r->rev_host = rev_save;
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(r->stream, &graph);
  if (rc != MBT_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    mbt_set_error("turn graph capture failed: %s", hipGetErrorString(e));
    return MBT_EDEVICE;
  }
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g.exec = nullptr;
    mbt_set_error("turn graph instantiation failed: %s", hipGetErrorString(e));
    return MBT_EDEVICE;
  }
  g.mon = mon;
  g.n = n;
  r->graph_builds++;
  *out = g.exec;
  return MBT_OK;
}

extern "C" int mbt_ring_placement(const mbt_ring *r, int *candidates, double *gbps)
{
  if (!r || !candidates || !gbps) {
    mbt_set_error("mbt_ring_placement: null argument");
    return MBT_EINVAL;
  }
  *candidates = r->place_candidates;
  *gbps = r->place_gbps;
  return MBT_OK;
}

extern "C" int mbt_ring_graph_builds(const mbt_ring *r, long *count)
{
  if (!r || !count) {
    mbt_set_error("mbt_ring_graph_builds: null argument");
    return MBT_EINVAL;
  }
  *count = r->graph_builds;
  return MBT_OK;
}

// Turns per block of a tracking call: the statistics of a block fit the device
// ring buffer; a multiple of NrevMon where possible, so that every full block
// replays the same graph.
static long track_block(const mbt_ring *r)
{
  long block = STATS_TURNS;
  if (r->nrev_mon <= STATS_TURNS) block -= STATS_TURNS % r->nrev_mon;
  return block;
}

// Enqueues one block of turns (graph replay, or eager launches without
// graphs) and returns without waiting.
static int enqueue_block(mbt_ring *r, long n)
{
  int rc;
  if (!r->use_graphs) {
    for (long t = 0; t < n; t++)
      if ((rc = enqueue_turn(r, NULL))) return rc;
  } else {
    hipGraphExec_t exec;
    if ((rc = turn_graph(r, n, &exec))) return rc;
    CK(hipGraphLaunch(exec, r->stream));
    r->rev_host += n;
  }
  return MBT_OK;
}

extern "C" long mbt_ring_async_max(const mbt_ring *r) { return track_block(r); }

extern "C" int mbt_ring_track_async(mbt_ring *r, long nturns)
{
  int rc;
  if ((rc = no_pending(r))) return rc;
  if (nturns < 1 || nturns > track_block(r)) {
    mbt_set_error("mbt_ring_track_async: 1 .. %ld turns per call", track_block(r));
    return MBT_EINVAL;
  }
  if (r->timing) {
    mbt_set_error("mbt_ring_track_async: not with kernel timing on");
    return MBT_EINVAL;
  }
  CK(hipSetDevice(r->device));
  (void)hipGetLastError(); // launch checks below must not report an earlier call's error
  const long long rev0 = r->rev_host;
  if ((rc = enqueue_block(r, nturns))) return rc;
  r->async_rev0 = rev0;
  r->async_n = nturns;
  return MBT_OK;
}

extern "C" int mbt_ring_track_wait(mbt_ring *r, double *stats, double *ampinv, double *rf)
{
  if (!r->async_n) {
    mbt_set_error("mbt_ring_track_wait: no turns pending");
    return MBT_EINVAL;
  }
  CK(hipSetDevice(r->device));
  const long n = r->async_n;
  r->async_n = 0;
  return download_slots(r, r->async_rev0, n, stats, ampinv, rf);
}

extern "C" int mbt_ring_track(mbt_ring *r, long nturns, double *stats, double *ampinv, double *rf)
{
  int rc;
  if ((rc = no_pending(r))) return rc;
  CK(hipSetDevice(r->device));
  (void)hipGetLastError(); // launch checks below must not report an earlier call's error
  for (int i = 0; i < 4; i++) {
    r->kt_ms[i] = 0;
    r->kt_launches[i] = 0;
  }
  const long block = track_block(r);
  long done = 0;
  while (done < nturns) {
    long n = nturns - done;
    if (n > block) n = block;
    long long rev0 = r->rev_host;
    const bool timed = r->timing != 0;
    if (timed) {
      for (long t = 0; t < n; t++)
        if ((rc = enqueue_turn(r, r->evpool + 6 * t))) return rc;
    } else if ((rc = enqueue_block(r, n))) {
      return rc;
    }
    if (timed) {
      CK(hipStreamSynchronize(r->stream));
      for (long t = 0; t < n; t++) {
        hipEvent_t *E = r->evpool + 6 * t;
        float a, b, c, d;
        CK(hipEventElapsedTime(&a, E[0], E[1]));
        CK(hipEventElapsedTime(&b, E[1], E[2]));
        CK(hipEventElapsedTime(&c, E[2], E[3]));
        CK(hipEventElapsedTime(&d, E[3], E[4]));
        r->kt_ms[0] += a;
        r->kt_ms[1] += b;
        r->kt_ms[3] += c;
        r->kt_ms[2] += d;
        for (int i = 0; i < 4; i++) r->kt_launches[i]++;
      }
    }
    rc = download_slots(r, rev0, n, stats ? stats + (size_t)done * r->G * 12 : NULL,
                        ampinv ? ampinv + (size_t)done * r->G * 6 : NULL, rf ? rf + (size_t)2 * r->R * done : NULL);
    if (rc) return rc;
    done += n;
  }
  CK(hipStreamSynchronize(r->stream));
  return MBT_OK;
}

extern "C" int mbt_ring_final_stats(mbt_ring *r, double *stats)
{
  {
    const int prc = no_pending(r);
    if (prc) return prc;
  }
  CK(hipSetDevice(r->device));
  (void)hipGetLastError(); // launch checks below must not report an earlier call's error
  int rc;
  // apply the pending RW kick, then two exact passes for the statistics
  if ((rc = launch_turn_pass(r, MODE_FINAL))) return rc;
  CK(hipMemsetAsync(r->pending, 0, (size_t)r->G * 2 * sizeof(double), r->stream));
  CK(hipMemsetAsync(r->shift, 0, (size_t)r->G * 6 * sizeof(double), r->stream));
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  if ((rc = launch_reduce(r, 0))) return rc;
  if ((rc = launch_turn_pass(r, MODE_STATS))) return rc;
  if ((rc = launch_reduce(r, 3))) return rc;
  CK(hipStreamSynchronize(r->stream));
  TurnState ts;
  CK(hipMemcpy(&ts, r->ts, sizeof ts, hipMemcpyDeviceToHost));
  int slot = (int)(ts.rev % STATS_TURNS);
  CK(hipMemcpy(stats, r->stats + (size_t)slot * r->G * 12, (size_t)r->G * 12 * sizeof(double),
               hipMemcpyDeviceToHost));
  return MBT_OK;
}

extern "C" int mbt_ring_histograms(const mbt_ring *r, double *counts, int *trash)
{
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  if (r->Nc > 0) CK(hipMemcpy(counts, r->fnp_out, (size_t)r->G * r->Nc * sizeof(double), hipMemcpyDeviceToHost));
  CK(hipMemcpy(trash, r->trash_tot, (size_t)r->G * 2 * sizeof(int), hipMemcpyDeviceToHost));
  return MBT_OK;
}

extern "C" int mbt_ring_potentials(const mbt_ring *r, int ib, double *out, double *rfprev)
{
  if (ib < 0 || ib >= r->G || r->Nc <= 0) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  int ring = 0, local = ib;
  while (local >= r->rh[ring].cfg.Nbunch) local -= r->rh[ring++].cfg.Nbunch;
  const size_t Nc = r->Nc, off = (size_t)ib * Nc;
  const size_t loff = (size_t)(r->Ph[ring].bucket_off + r->rh[ring].cfg.bucket[local]) * Nc;
  const double *src[7] = {r->sfL + off, r->lr_wake + loff, r->fnp_out + off, r->sfV + off, r->dipV_out + off,
                          r->sfH + off, r->dipH_out + off};
  // eight small copies into pinned memory, one wait
  double *stage = r->pot_host;
  for (int k = 0; k < 7; k++)
    CK(hipMemcpyAsync(stage + k * Nc, src[k], Nc * sizeof(double), hipMemcpyDeviceToHost, r->stream));
  CK(hipMemcpyAsync(stage + 7 * Nc, r->ts + ring, sizeof(TurnState), hipMemcpyDeviceToHost, r->stream));
  CK(hipStreamSynchronize(r->stream));
  memcpy(out, stage, 7 * Nc * sizeof(double));
  TurnState ts;
  memcpy(&ts, stage + 7 * Nc, sizeof ts);
  if (rfprev) {
    rfprev[0] = ts.phai0_prev;
    rfprev[1] = ts.Vrf0_prev;
    rfprev[2] = ts.scan_val;
  }
  return MBT_OK;
}

extern "C" int mbt_ring_rf_state(const mbt_ring *r, int ring, double *out)
{
  if (ring < 0 || ring >= r->R || !out) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  TurnState ts;
  CK(hipMemcpy(&ts, r->ts + ring, sizeof ts, hipMemcpyDeviceToHost));
  // every turn ends with the feedback step, which first saves the phase and
  // voltage the turn's map used
  out[0] = ts.phai0_prev;
  out[1] = ts.Vrf0_prev;
  out[2] = ts.has_fb;
  return MBT_OK;
}

extern "C" int mbt_ring_rw_history(const mbt_ring *r, double *out)
{
  if (!r->rh[0].cfg.EnableRW_long) return MBT_EINVAL;
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  const mbt_config *c = &r->rh[0].cfg;
  CK(hipMemcpy(out, r->rwhist, (size_t)(c->Nmlt + 2) * c->Nharm * 2 * sizeof(double), hipMemcpyDeviceToHost));
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
extern "C" int mbt_ensemble_size(const mbt_ring *r, int *nrings, int *nbunch_total, long long *nparticles)
{
  if (nrings) *nrings = r->R;
  if (nbunch_total) *nbunch_total = r->G;
  if (nparticles) *nparticles = r->Ntot;
  return MBT_OK;
}

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

// The placement ceiling of the ring's turn pass: k_stream over the ring's own
// arrays (all Nalloc slots, 96 bytes each), reps timed passes after one untimed.
extern "C" int mbt_ring_stream_ceiling(mbt_ring *r, int reps, double *gbps)
{
  if (reps < 1 || !gbps) {
    mbt_set_error("stream ceiling: reps >= 1 and an output are required");
    return MBT_EINVAL;
  }
  CK(hipSetDevice(r->device));
  (void)hipGetLastError();
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const unsigned nb = (unsigned)(r->Nalloc >> TILE_SHIFT);
  int rc = MBT_OK;
  for (int rep = 0; rep <= reps; rep++) {
    if (rep == 1) CK(hipEventRecord(e0, r->stream));
    hipLaunchKernelGGL(k_stream, dim3(nb), dim3(WG), 0, r->stream, r->C, 0ULL);
    if (hipGetLastError() != hipSuccess) {
      mbt_set_error("stream ceiling launch failed");
      rc = MBT_EDEVICE;
      break;
    }
  }
  float t = 0.0f;
  if (!rc) {
    CK(hipEventRecord(e1, r->stream));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&t, e0, e1));
    *gbps = 96.0 * (double)r->Nalloc / ((double)t / reps * 1e-3) / 1e9;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" int mbt_ring_set_chunk(mbt_ring *r, int items)
{
  CK(hipSetDevice(r->device));
  CK(hipStreamSynchronize(r->stream));
  return set_decomposition(r, items);
}

extern "C" int mbt_ring_get_chunk(const mbt_ring *r) { return r->chunk_items; }

// Chooses the chunk (macroparticles per lane of a work-group) by timing the
// ring's k_turn variant in MODE_TUNE (the turn's work with the coordinates
// written back unchanged; it changes no state but the histogram accumulators,
// cleared afterwards) for each candidate; keeps the fastest.  The choice changes only
// how the moment sums are grouped, i.e. results at the rounding level.
extern "C" int mbt_ring_autotune(mbt_ring *r, int *best_items, double *ms)
{
  {
    const int prc = no_pending(r);
    if (prc) return prc;
  }
  static const int cand[AUTOTUNE_CANDIDATES] = {1, 2, 4, 8, 16, 32, 64};
  CK(hipSetDevice(r->device));
  (void)hipGetLastError();
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  // the turn's own variant, run as a timing pass (coordinates unchanged)
  const turn_kernel_t k = g_tune_kernels[(r->any_qe ? 4 : 0) + (r->any_dip ? 2 : 0) + (r->any_gen ? 1 : 0)];
  const long long stride = (long long)r->H * (r->Nc > 0 ? r->Nc : 1);
  // candidates interleaved over AUTOTUNE_ROUNDS rounds, each timed over
  // AUTOTUNE_REPS passes after one untimed pass; a candidate's time is its best
  // round (the least disturbed by clocks and other work)
  int best = r->chunk_items, rc = MBT_OK;
  double tbest[AUTOTUNE_CANDIDATES];
  for (int c = 0; c < AUTOTUNE_CANDIDATES; c++) tbest[c] = 1e300;
  for (int round = 0; round < AUTOTUNE_ROUNDS && rc == MBT_OK; round++)
    for (int c = 0; c < AUTOTUNE_CANDIDATES && rc == MBT_OK; c++) {
      if (cand[c] < r->min_items) continue; // below the buffers of a large ring
      if ((rc = set_decomposition(r, cand[c]))) break;
      float t = 0.0f;
      for (int rep = 0; rep <= AUTOTUNE_REPS; rep++) {
        if (rep == 1) CK(hipEventRecord(e0, r->stream));
        hipLaunchKernelGGL(k, dim3(r->nwg), dim3(WG), r->lds_turn, r->stream, r->P, r->C, r->wg_first, r->wg_end,
                           r->wg_bunch, r->bunch_ring, r->bucket, r->bunch_first, r->shift, r->sfL, r->sfV, r->sfH,
                           r->lr_wake, stride, r->pending, r->partial, r->Hs, r->ts, RwOld{});
        CK(hipGetLastError());
      }
      CK(hipEventRecord(e1, r->stream));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&t, e0, e1));
      const double m = t / AUTOTUNE_REPS;
      if (m < tbest[c]) tbest[c] = m;
    }
  double best_ms = 1e300;
  for (int c = 0; c < AUTOTUNE_CANDIDATES; c++) {
    if (ms) ms[c] = tbest[c] < 1e300 ? tbest[c] : 0.0; // 0: not tried
    if (tbest[c] < best_ms) {
      best_ms = tbest[c];
      best = cand[c];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  // the statistics passes binned into the accumulators: clear them
  const size_t GN = (size_t)r->G * (r->Nc > 0 ? r->Nc : 1);
  CK(hipMemsetAsync(r->Hs.cnt, 0, GN * sizeof(unsigned), r->stream));
  CK(hipMemsetAsync(r->Hs.sx, 0, GN * 2 * sizeof(unsigned long long), r->stream));
  CK(hipMemsetAsync(r->Hs.sz, 0, GN * 2 * sizeof(unsigned long long), r->stream));
  CK(hipMemsetAsync(r->Hs.tcnt, 0, (size_t)r->G * 2 * sizeof(unsigned), r->stream));
  CK(hipMemsetAsync(r->Hs.tsx, 0, (size_t)r->G * 4 * sizeof(unsigned long long), r->stream));
  CK(hipMemsetAsync(r->Hs.tsz, 0, (size_t)r->G * 4 * sizeof(unsigned long long), r->stream));
  CK(hipMemsetAsync(r->Hs.bad, 0, (size_t)r->G * sizeof(int), r->stream));
  if ((rc = set_decomposition(r, best))) return rc;
  if (best_items) *best_items = best;
  return MBT_OK;
}
