/* Rounding-sensitivity switches of the weak-strong oracle -- TEST INFRASTRUCTURE.
 *
 * The oracle proper is built with none of them and follows the reference's
 * operations exactly.  Each switch builds a variant that changes the arithmetic
 * at the rounding level only, the way a parallel implementation may:
 *   ORC_SENS_RECIP  grid quotients by the reciprocal spacing (x * (1/a) in place
 *                   of x / a) in the field interpolation;
 *   ORC_SENS_TREE   sums over the ion cloud as balanced pairwise trees in place of
 *                   sequential loops;
 *   ORC_SENS_ULP    sin, cos and exp results moved by one ulp up, down or not
 *                   at all, chosen by a hash of the argument.
 * tests/test_oracle_ws_sensitivity.py measures how far each variant moves the
 * weak-strong outputs: the floor under any comparison of two implementations
 * that round differently. */
#ifndef ORACLE_SENS_H
#define ORACLE_SENS_H
#include <math.h>
#include <stdint.h>
#include <string.h>

static inline double sens_ulp(double r, double arg)
{
#ifdef ORC_SENS_ULP
  uint64_t b;
  memcpy(&b, &arg, 8);
  b ^= b >> 33;
  b *= 0xff51afd7ed558ccdULL;
  b ^= b >> 29;
  const int k = (int)(b % 3);
  if (k == 1) return nextafter(r, INFINITY);
  if (k == 2) return nextafter(r, -INFINITY);
#endif
  (void)arg;
  return r;
}

static inline double sens_ulp_exp(double x) { return sens_ulp(exp(x), x); }

#ifdef ORC_SENS_TREE
/* balanced pairwise sum of n values spaced `stride` apart */
static double sens_pairwise(const double *v, long n, long stride)
{
  if (n <= 0) return 0.0;
  if (n == 1) return v[0];
  const long h = n / 2;
  return sens_pairwise(v, h, stride) + sens_pairwise(v + h * stride, n - h, stride);
}

static double sens_pairwise_prod(const double *a, const double *b, long n)
{
  if (n <= 0) return 0.0;
  if (n == 1) return b ? a[0] * b[0] : a[0];
  const long h = n / 2;
  return sens_pairwise_prod(a, b, h) + sens_pairwise_prod(a + h, b ? b + h : NULL, n - h);
}

static void sens_cloud_sums(const double *nri, const double *x, const double *y, long n, double *sx, double *sy,
                            double *sn)
{
  *sx = sens_pairwise_prod(nri, x, n);
  *sy = sens_pairwise_prod(nri, y, n);
  *sn = sens_pairwise_prod(nri, NULL, n);
}
#endif
#endif
