// Device Faddeeva function (ACM TOMS 680), shared by the evaluation kernel and
// the beam-ion field kernels.  Returns 0, or 1 where the reflection to the lower
// half plane would overflow (the reference's iflag = -1, src/physic.c:293-299).
#ifndef MBT_WOFZ_DEV_H
#define MBT_WOFZ_DEV_H
#include <hip/hip_runtime.h>

template <typename T> struct wofz_lim;
template <> struct wofz_lim<double> {
  static constexpr double big = 0.5e154, big_exp = 708.503061461606, big_trig = 3.53711887601422e08;
};
template <> struct wofz_lim<float> {
  // float range: |z|^2 must stay finite, 2 exp(-z^2) must not overflow
  static constexpr float big = 1.0e18f, big_exp = 88.0f, big_trig = 1.0e5f;
};

template <typename T> __device__ __forceinline__ T wexp(T a);
template <> __device__ __forceinline__ double wexp<double>(double a) { return exp(a); }
template <> __device__ __forceinline__ float wexp<float>(float a) { return expf(a); }
template <typename T> __device__ __forceinline__ void wsincos(T a, T *s, T *c);
template <> __device__ __forceinline__ void wsincos<double>(double a, double *s, double *c) { sincos(a, s, c); }
template <> __device__ __forceinline__ void wsincos<float>(float a, float *s, float *c) { sincosf(a, s, c); }

template <typename T> __device__ __forceinline__ int mbt_wofz_dev(T xi, T yi, T &u, T &v)
{
  const T two_over_sqrtpi = (T)1.12837916709551257388;
  T ax = fabs(xi), ay = fabs(yi);
  if (ax > wofz_lim<T>::big || ay > wofz_lim<T>::big) {
    u = v = (T)0;
    return 1;
  }
  T xs = ax / (T)6.3, ys = ay / (T)4.4;
  T rho2 = xs * xs + ys * ys;
  T re2 = ax * ax - ay * ay;
  T im2 = (T)2.0 * ax * ay;
  T wr, wi, er = 0, ei = 0;
  const bool series = rho2 < (T)0.085264;
  if (series) {
    T q = ((T)1.0 - (T)0.85 * ys) * sqrt(rho2);
    int n = (int)((T)6.0 + (T)72.0 * q);
    int j = 2 * n + 1;
    T sr = (T)1.0 / (T)j, si = 0;
    for (int i = n; i >= 1; i--) {
      j -= 2;
      T t = (sr * re2 - si * im2) / (T)i;
      si = (sr * im2 + si * re2) / (T)i;
      sr = t + (T)1.0 / (T)j;
    }
    T a = -two_over_sqrtpi * (sr * ay + si * ax) + (T)1.0;
    T b = two_over_sqrtpi * (sr * ax - si * ay);
    T e = wexp<T>(-re2), sn, cs;
    wsincos<T>(im2, &sn, &cs);
    er = e * cs;
    ei = -e * sn;
    wr = a * er - b * ei;
    wi = a * ei + b * er;
  } else {
    T hstep, h2 = 0, lam = 0;
    int nt, nf;
    if (rho2 > (T)1.0) {
      hstep = 0;
      nt = 0;
      T q = sqrt(rho2);
      nf = (int)((T)3.0 + (T)1442.0 / ((T)26.0 * q + (T)77.0));
    } else {
      T q = ((T)1.0 - ys) * sqrt((T)1.0 - rho2);
      hstep = (T)1.88 * q;
      h2 = (T)2.0 * hstep;
      nt = (int)((T)7.0 + (T)34.0 * q);
      nf = (int)((T)16.0 + (T)26.0 * q);
    }
    const bool taylor = hstep > (T)0;
    if (taylor) lam = pow(h2, (T)nt);
    T rr = 0, ri = 0, sr = 0, si = 0;
    for (int n = nf; n >= 0; n--) {
      int n1 = n + 1;
      T tr = ay + hstep + (T)n1 * rr;
      T ti = ax - (T)n1 * ri;
      T c = (T)0.5 / (tr * tr + ti * ti);
      rr = c * tr;
      ri = c * ti;
      if (taylor && n <= nt) {
        T t = lam + sr;
        sr = rr * t - ri * si;
        si = ri * t + rr * si;
        lam = lam / h2;
      }
    }
    if (hstep == (T)0) {
      wr = two_over_sqrtpi * rr;
      wi = two_over_sqrtpi * ri;
    } else {
      wr = two_over_sqrtpi * sr;
      wi = two_over_sqrtpi * si;
    }
    if (ay == (T)0) wr = wexp<T>(-ax * ax);
  }
  if (yi < (T)0) {
    if (series) {
      er = (T)2.0 * er;
      ei = (T)2.0 * ei;
    } else {
      T mre2 = -re2;
      if (im2 > wofz_lim<T>::big_trig || mre2 > wofz_lim<T>::big_exp) {
        u = v = (T)0;
        return 1;
      }
      T e = (T)2.0 * wexp<T>(mre2), sn, cs;
      wsincos<T>(im2, &sn, &cs);
      er = e * cs;
      ei = -e * sn;
    }
    wr = er - wr;
    wi = ei - wi;
    if (xi > (T)0) wi = -wi;
  } else if (xi < (T)0) {
    wi = -wi;
  }
  u = wr;
  v = wi;
  return 0;
}
#endif
