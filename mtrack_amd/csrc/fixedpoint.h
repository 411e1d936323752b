// Exact 128-bit fixed point of the slice dipole sums: a double x becomes the
// two's-complement integer q = round(x 2^100) (ties away from zero) held in two
// 64-bit words (lo, hi), for |x| < 2^20; sums of such integers are exact and
// independent of their order while the sum's magnitude stays below 2^27.
// Host and device code (the host side serves the unit test of these helpers).
#ifndef MBT_FIXEDPOINT_H
#define MBT_FIXEDPOINT_H
#include <hip/hip_runtime.h>
#include <string.h>

__host__ __device__ inline void fx_split(double x, unsigned long long &lo, unsigned long long &hi)
{
  unsigned long long bits;
  memcpy(&bits, &x, sizeof bits);
  const int e = (int)((bits >> 52) & 0x7ff);
  const unsigned long long m = (bits & 0xfffffffffffffULL) | (1ULL << 52);
  lo = 0;
  hi = 0;
  if (e == 0) return;
  const int s = e - 975; // x 2^100 = m 2^s
  if (s >= 64) {
    hi = m << (s - 64);
  } else if (s > 0) {
    lo = m << s;
    hi = m >> (64 - s);
  } else if (s == 0) {
    lo = m;
  } else if (s > -64) {
    const int r = -s;
    lo = (m >> r) + ((m >> (r - 1)) & 1ULL);
  } else if (s == -64) {
    lo = (m >> 63) & 1ULL;
  }
  if (bits >> 63) { // negate
    lo = ~lo + 1ULL;
    hi = ~hi + (lo == 0ULL ? 1ULL : 0ULL);
  }
}

// (lo, hi) += (a, b), modulo 2^128
__host__ __device__ inline void fx_add(unsigned long long &lo, unsigned long long &hi, unsigned long long a,
                                       unsigned long long b)
{
  const unsigned long long old = lo;
  lo += a;
  hi += b + (lo < old ? 1ULL : 0ULL);
}

__host__ __device__ inline double fx_to_double(unsigned long long lo, unsigned long long hi)
{
  const long long h = (long long)hi;
  const double v = (double)h * 18446744073709551616.0 + (double)lo;
  return v * 7.888609052210118e-31; // 2^-100
}
#endif
