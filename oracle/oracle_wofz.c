/*
 * Faddeeva function w(z) = exp(-z^2) erfc(-iz) -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates ACM TOMS Algorithm 680 (G.P.M. Poppe, C.M.J. Wijers, "More
 * efficient computation of the complex error function", ACM TOMS 16(1), 1990),
 * the algorithm behind the reference's complex_errorfunc_wofz_algo680
 * (src/physic.c:75-301): in the first quadrant, a power series for small |z|
 * (Abramowitz & Stegun 7.1.5), the Laplace continued fraction for large |z|,
 * a Taylor expansion with continued-fraction derivatives in between; other
 * quadrants by w(-z) = 2 exp(-z^2) - w(z) and w(conj z) = conj w(-z).
 * Returns 0, or 1 where the reflection would overflow.
 *
 * libm calls as the reference's -O2 build makes them (glibc's sincos and sin
 * can differ in the last bit): the series branch's sin/cos pair is one sincos
 * call there, the reflection's pair two separate calls.  Pinned bitwise against
 * the compiled reference over 2e6 points (tests/test_wofz.py).
 */
#define _GNU_SOURCE /* sincos */
#include <math.h>
#include "oracle_sens.h"
#include "oracle.h"

/* separate calls that the compiler cannot pair into sincos */
static __attribute__((noinline)) double lib_cos(double x) { return cos(x); }
static __attribute__((noinline)) double lib_sin(double x) { return sin(x); }

int orc_wofz(double xi, double yi, double *u, double *v)
{
  const double two_over_sqrtpi = 1.12837916709551257388;
  const double big = 0.5e154, big_exp = 708.503061461606, big_trig = 3.53711887601422e08;
  double ax = fabs(xi), ay = fabs(yi);
  if (ax > big || ay > big) return 1;
  double xs = ax / 6.3, ys = ay / 4.4;
  double rho2 = xs * xs + ys * ys;
  double re2 = ax * ax - ay * ay; /* Re z^2 */
  double im2 = 2.0 * ax * ay;     /* Im z^2 */
  double wr, wi, er = 0.0, ei = 0.0;
  int series = rho2 < 0.085264;
  if (series) {
    double q = (1.0 - 0.85 * ys) * sqrt(rho2);
    int n = (int)(6.0 + 72.0 * q);
    int j = 2 * n + 1;
    double sr = 1.0 / ((double)j), si = 0.0;
    for (int i = n; i >= 1; i--) {
      j -= 2;
      double t = (sr * re2 - si * im2) / ((double)i);
      si = (sr * im2 + si * re2) / ((double)i);
      sr = t + 1.0 / ((double)j);
    }
    double a = -two_over_sqrtpi * (sr * ay + si * ax) + 1.0;
    double b = two_over_sqrtpi * (sr * ax - si * ay);
    double e = sens_ulp_exp(-re2), sn, cs;
    sincos(im2, &sn, &cs);
    sn = sens_ulp(sn, im2);
    cs = sens_ulp(cs, -im2);
    er = e * cs;
    ei = -e * sn;
    wr = a * er - b * ei;
    wi = a * ei + b * er;
  } else {
    double hstep, h2 = 0.0, lam = 0.0;
    int nt, nf;
    if (rho2 > 1.0) {
      hstep = 0.0;
      nt = 0;
      double q = sqrt(rho2);
      nf = (int)(3.0 + 1442.0 / (26.0 * q + 77.0));
    } else {
      double q = (1.0 - ys) * sqrt(1.0 - rho2);
      hstep = 1.88 * q;
      h2 = 2.0 * hstep;
      nt = (int)(7.0 + 34.0 * q);
      nf = (int)(16.0 + 26.0 * q);
    }
    int taylor = hstep > 0.0;
    if (taylor) lam = pow(h2, nt);
    double rr = 0.0, ri = 0.0, sr = 0.0, si = 0.0;
    for (int n = nf; n >= 0; n--) {
      int n1 = n + 1;
      double tr = ay + hstep + n1 * rr;
      double ti = ax - n1 * ri;
      double c = 0.5 / (tr * tr + ti * ti);
      rr = c * tr;
      ri = c * ti;
      if (taylor && n <= nt) {
        double t = lam + sr;
        sr = rr * t - ri * si;
        si = ri * t + rr * si;
        lam = lam / h2;
      }
    }
    if (hstep == 0.0) {
      wr = two_over_sqrtpi * rr;
      wi = two_over_sqrtpi * ri;
    } else {
      wr = two_over_sqrtpi * sr;
      wi = two_over_sqrtpi * si;
    }
    if (ay == 0.0) wr = sens_ulp_exp(-ax * ax);
  }
  if (yi < 0.0) {
    if (series) {
      er = 2.0 * er;
      ei = 2.0 * ei;
    } else {
      double mre2 = -re2;
      if (im2 > big_trig || mre2 > big_exp) return 1;
      double e = 2.0 * sens_ulp_exp(mre2);
      er = e * sens_ulp(lib_cos(im2), im2);
      ei = -e * sens_ulp(lib_sin(im2), -im2);
    }
    wr = er - wr;
    wi = ei - wi;
    if (xi > 0.0) wi = -wi;
  } else if (xi < 0.0) {
    wi = -wi;
  }
  *u = wr;
  *v = wi;
  return 0;
}
