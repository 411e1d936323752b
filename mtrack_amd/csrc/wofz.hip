// Faddeeva function w(z) = exp(-z^2) erfc(-iz) on gfx950, one evaluation per
// lane: ACM TOMS Algorithm 680 (Poppe & Wijers 1990), the algorithm of the
// reference's complex_errorfunc_wofz_algo680 (src/physic.c:75-301).  First
// quadrant: power series for small |z|, Laplace continued fraction for large
// |z|, Taylor expansion with continued-fraction derivatives in between; the
// other quadrants by reflection.  Lanes of a wavefront take different branch
// and term counts; the per-lane loops are short (<= 78 terms), so divergence
// costs little against the FP64 divides.  Templated on the arithmetic type:
// FP64 for the reference path, FP32 for the mixed-precision beam-ion field.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include "mbt_host.h"
#include "wofz_dev.h"

#pragma clang fp contract(off)

template <typename T>
__global__ void k_wofz(long n, const T *x, const T *y, T *wr, T *wi, int *flag)
{
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T u, v;
  int f = mbt_wofz_dev<T>(x[i], y[i], u, v);
  wr[i] = u;
  wi[i] = v;
  flag[i] = f;
}

template <typename T>
static int run_wofz(int device, long n, const T *x, const T *y, T *wr, T *wi, int *flag)
{
  if (n <= 0) return MBT_OK;
  if (hipSetDevice(device) != hipSuccess) {
    mbt_set_error("no HIP device %d", device);
    return MBT_EDEVICE;
  }
  T *d = NULL;
  int *df = NULL;
  size_t bytes = (size_t)n * sizeof(T);
  if (hipMalloc((void **)&d, 4 * bytes) != hipSuccess || hipMalloc((void **)&df, n * sizeof(int)) != hipSuccess) {
    if (d) (void)hipFree(d);
    mbt_set_error("hipMalloc failed");
    return MBT_ENOMEM;
  }
  int rc = MBT_OK;
  if (hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n, y, bytes, hipMemcpyHostToDevice) != hipSuccess)
    rc = MBT_EDEVICE;
  if (rc == MBT_OK) {
    hipLaunchKernelGGL(k_wofz<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, d, d + n, d + 2 * n, d + 3 * n,
                       df);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = MBT_EDEVICE;
  }
  if (rc == MBT_OK && (hipMemcpy(wr, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(wi, d + 3 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
                       hipMemcpy(flag, df, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess))
    rc = MBT_EDEVICE;
  (void)hipFree(d);
  (void)hipFree(df);
  if (rc) mbt_set_error("mbt_wofz: HIP failure");
  return rc;
}

extern "C" int mbt_wofz(int device, long n, const double *x, const double *y, double *wr, double *wi, int *flag)
{
  return run_wofz<double>(device, n, x, y, wr, wi, flag);
}

extern "C" int mbt_wofz_f32(int device, long n, const float *x, const float *y, float *wr, float *wi, int *flag)
{
  return run_wofz<float>(device, n, x, y, wr, wi, flag);
}
