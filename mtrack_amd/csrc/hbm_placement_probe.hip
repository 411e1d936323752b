// HBM placement probe (a measurement aid, not part of the library): the rate of
// the turn pass's traffic pattern -- six FP64 arrays updated in place with
// nontemporal loads and stores, 96 bytes per slot -- in each of eight 24 GB
// allocations, allocated twice over (all eight freed in between), with the six
// arrays at 2 MiB-aligned offsets and staggered by 512 KiB, and the same memory
// as one stream.  Build and run:
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         mtrack_amd/csrc/hbm_placement_probe.hip -o hbm_placement_probe && ./hbm_placement_probe
#include <hip/hip_runtime.h>
#include <stdio.h>

#define NTL(p) __builtin_nontemporal_load(p)
#define NTS(v, p) __builtin_nontemporal_store(v, p)

// six arrays of n doubles at base + j * stride, 2,048 slots per block
__global__ __launch_bounds__(256) void k_six(double *base, long n, long stride, double s)
{
  double *a = base, *b = a + stride, *c = b + stride, *d = c + stride, *e = d + stride, *f = e + stride;
  const long i0 = (long)blockIdx.x * 2048;
  const long i1 = i0 + 2048 < n ? i0 + 2048 : n;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) {
    const double x0 = NTL(a + i), x1 = NTL(b + i), x2 = NTL(c + i), x3 = NTL(d + i), x4 = NTL(e + i),
                 x5 = NTL(f + i);
    NTS(x0 * s + x5, a + i);
    NTS(x1 * s + x0, b + i);
    NTS(x2 * s + x1, c + i);
    NTS(x3 * s + x2, d + i);
    NTS(x4 * s + x3, e + i);
    NTS(x5 * s + x4, f + i);
  }
}

// one stream: in-place update of n6 consecutive doubles, 12,288 per block
__global__ __launch_bounds__(256) void k_one(double *a, long n6, double s)
{
  const long i0 = (long)blockIdx.x * 12288;
  const long i1 = i0 + 12288 < n6 ? i0 + 12288 : n6;
  for (long i = i0 + threadIdx.x; i < i1; i += 256) NTS(NTL(a + i) * s + 1.0, a + i);
}

// GB/s (96 bytes per slot of n) over 5 launches after 2 untimed
static double rate(int one, double *p, long n, long stride)
{
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (int w = 0; w < 7; w++) {
    if (w == 2) (void)hipEventRecord(e0);
    if (one)
      k_one<<<(unsigned)((6 * n + 12287) / 12288), 256>>>(p, 6 * n, 1.0);
    else
      k_six<<<(unsigned)((n + 2047) / 2048), 256>>>(p, n, stride, 1.0);
  }
  (void)hipEventRecord(e1);
  (void)hipEventSynchronize(e1);
  float ms = 0.0f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 96.0 * n / (ms / 5 * 1e-3) / 1e9;
}

#define NCH 8
int main()
{
  const long n = 500000000L;       // doubles per array (4 GB)
  const long stride = 500170752L;  // n rounded up to a 2 MiB multiple
  const long stagger = 65536L;     // 512 KiB
  const size_t bytes = (size_t)(stride + stagger) * 48;
  double *ch[NCH];
  for (int round = 0; round < 2; round++) {
    for (int k = 0; k < NCH; k++) {
      if (hipMalloc((void **)&ch[k], bytes) != hipSuccess || hipMemset(ch[k], 0, bytes) != hipSuccess) {
        printf("allocation %d of round %d failed\n", k, round);
        return 1;
      }
    }
    if (hipDeviceSynchronize() != hipSuccess) return 1;
    for (int k = 0; k < NCH; k++) {
      printf("round %d allocation %d: six arrays %.0f GB/s, staggered 512 KiB %.0f GB/s, one stream %.0f GB/s\n",
             round, k, rate(0, ch[k], n, stride), rate(0, ch[k], n, stride + stagger), rate(1, ch[k], n, 0));
      fflush(stdout);
    }
    for (int k = 0; k < NCH; k++) (void)hipFree(ch[k]);
  }
  return hipDeviceSynchronize() == hipSuccess ? 0 : 1;
}
