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
#include <mutex>
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

// ---- allocation-free variant ------------------------------------------------
// For callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc): the staging of points
// and results lives in a file-scope device array instead of hipMalloc'ed
// buffers.  Each device holds its own copy of the array for the life of the
// process; the host reaches it only through the address hipGetSymbolAddress
// returns on that device (the array's name on the host is not a device
// address).  Every call on a device shares that device's array, so calls that
// overlap on one device race; nothing in it needs clearing, since a call
// writes every slot it reads ([0, n) of the points by the copy in, of the
// results by the kernel).
#define WOFZ_NOALLOC_MAX (1L << 18) // points per call: 8 MiB of doubles + 1 MiB of flags
#define WOFZ_SCRATCH_BYTES (WOFZ_NOALLOC_MAX * (4 * sizeof(double) + sizeof(int)))
#define WOFZ_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[WOFZ_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[WOFZ_MAX_DEVICES];

// The device address of `scratch` on `device`, looked up once per device with
// that device current.
static int scratch_on(int device, unsigned char **out)
{
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = NULL;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mbt_set_error("mbt_wofz_noalloc: no address for the scratch array on device %d: %s", device,
                    hipGetErrorString(e));
      return MBT_EDEVICE;
    }
    g_scratch_ptr[device] = (unsigned char *)p;
  }
  *out = g_scratch_ptr[device];
  return MBT_OK;
}

extern "C" int mbt_wofz_noalloc(int device, long n, const double *x, const double *y, double *wr, double *wi,
                                int *flag)
{
  if (n < 0 || n > WOFZ_NOALLOC_MAX) {
    mbt_set_error("mbt_wofz_noalloc: %ld points outside [0, %ld], the size of its fixed scratch array", n,
                  WOFZ_NOALLOC_MAX);
    return MBT_EINVAL;
  }
  if (n == 0) return MBT_OK;
  if (device < 0 || device >= WOFZ_MAX_DEVICES || hipSetDevice(device) != hipSuccess) {
    mbt_set_error("no HIP device %d", device);
    return MBT_EDEVICE;
  }
  unsigned char *base = NULL;
  int rc = scratch_on(device, &base);
  if (rc) return rc;
  double *d = (double *)base; // x, y, wr, wi: WOFZ_NOALLOC_MAX doubles each
  int *df = (int *)(base + 4 * WOFZ_NOALLOC_MAX * sizeof(double));
  const size_t bytes = (size_t)n * sizeof(double);
  const long M = WOFZ_NOALLOC_MAX;
  if (hipMemcpy(d, x, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + M, y, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    mbt_set_error("mbt_wofz_noalloc: copy of the points failed");
    return MBT_EDEVICE;
  }
  hipLaunchKernelGGL(k_wofz<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, n, d, d + M, d + 2 * M,
                     d + 3 * M, df);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
    mbt_set_error("mbt_wofz_noalloc: kernel failed");
    return MBT_EDEVICE;
  }
  if (hipMemcpy(wr, d + 2 * M, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(wi, d + 3 * M, bytes, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(flag, df, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    mbt_set_error("mbt_wofz_noalloc: copy of the results failed");
    return MBT_EDEVICE;
  }
  return MBT_OK;
}
