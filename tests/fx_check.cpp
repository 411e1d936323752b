// Host driver of the fixed-point helpers (mtrack_amd/csrc/fixedpoint.h) for
// tests/test_fixedpoint.py: reads "n" then n doubles (as 16 hex digits of
// their bits) and prints each value's (lo, hi) split, then the 128-bit sum of
// all of them and its conversion back to double.
#include <stdio.h>
#include <string.h>
#include "fixedpoint.h"

int main()
{
  long n;
  if (scanf("%ld", &n) != 1) return 1;
  unsigned long long slo = 0, shi = 0;
  for (long i = 0; i < n; i++) {
    unsigned long long b;
    if (scanf("%llx", &b) != 1) return 1;
    double x;
    memcpy(&x, &b, sizeof x);
    unsigned long long lo, hi;
    fx_split(x, lo, hi);
    printf("%016llx %016llx\n", hi, lo);
    fx_add(slo, shi, lo, hi);
  }
  const double v = fx_to_double(slo, shi);
  unsigned long long vb;
  memcpy(&vb, &v, sizeof vb);
  printf("sum %016llx %016llx %016llx\n", shi, slo, vb);
  return 0;
}
