/* Internal declarations shared by the host-side C units of libmbtrack_mi355x. */
#ifndef MBT_HOST_H
#define MBT_HOST_H

#include <stdio.h>
#include "mbtrack_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SI prefixes and constants with the reference's values (src/def.h:10-28). */
#define MBT_FPICO 1.0e-12
#define MBT_FNANO 1.0e-09
#define MBT_FMICRO 1.0e-06
#define MBT_FMILLI 1.0e-03
#define MBT_FKILO 1.0e+03
#define MBT_FMEGA 1.0e+06
#define MBT_FGIGA 1.0e+09
#define MBT_FTERA 1.0e+12
#define MBT_C_LIGHT 2.997925e+08
#define MBT_Z0 376.73
#define MBT_PI 3.14159265358979323846

void mbt_set_error(const char *fmt, ...);

/* glibc random(3) TYPE_3 additive generator, restated: the stream the
 * reference draws through srand/rand (src/bunch.c:379-436). */
typedef struct mbt_glibc_rand {
  int r[31];
  int f, b; /* front and rear indices into the 31-word table */
} mbt_glibc_rand;
void mbt_glibc_srand(mbt_glibc_rand *g, unsigned seed);
int mbt_glibc_rand_next(mbt_glibc_rand *g);

/* Output writers with the reference's formats. */
typedef struct mbt_stat_accum {
  double sum[4][3], sum_sqr[4][3]; /* 0 pos, 1 slope, 2 pos_sigma, 3 slope_sigma */
} mbt_stat_accum;
void mbt_accum_reset(mbt_stat_accum *a);
void mbt_accum_add(mbt_stat_accum *a, const double *st12);
void mbt_accum_update(mbt_stat_accum *a, long N);
void mbt_stats_fprintf(FILE *fp, const mbt_config *cfg, long rev, const mbt_stat_accum *a, double scan);
void mbt_worker_stat_header(FILE *fp, const mbt_config *cfg, const double *st12, double Ib);

#ifdef __cplusplus
}
#endif
#endif
