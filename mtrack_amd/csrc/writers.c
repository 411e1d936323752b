/*
 * Output files of the weak-weak run, with the reference's formats:
 * statistics accumulators (src/statistics.c:36-45,203-260) and the writers of
 * src/bunch.c:552-804 and src/bunch_weak.c:295-447.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "mbt_host.h"

enum { A_POS = 0, A_SLOPE = 1, A_PSIG = 2, A_SSIG = 3 };

void mbt_accum_reset(mbt_stat_accum *a) { memset(a, 0, sizeof *a); }

/* weak_bunch_add_statistics, src/statistics.c:203-220; st12 = pos[3] pos_sigma[3] slope[3] slope_sigma[3] */
void mbt_accum_add(mbt_stat_accum *a, const double *st12)
{
  for (int p = 0; p < 3; p++) {
    double v[4] = {st12[p], st12[6 + p], st12[3 + p], st12[9 + p]};
    for (int k = 0; k < 4; k++) {
      a->sum[k][p] = a->sum[k][p] + v[k];
      a->sum_sqr[k][p] = a->sum_sqr[k][p] + v[k] * v[k];
    }
  }
}

/* weak_bunch_update_statistics, src/statistics.c:222-245 */
void mbt_accum_update(mbt_stat_accum *a, long N)
{
  for (int k = 0; k < 4; k++)
    for (int p = 0; p < 3; p++) {
      if (N > 1) {
        double s2 = a->sum[k][p] * a->sum[k][p];
        s2 = (-1.0 / ((double)N)) * s2;
        double v = s2 + a->sum_sqr[k][p];
        a->sum_sqr[k][p] = (1.0 / ((double)(N - 1))) * v;
      } else {
        a->sum_sqr[k][p] = 0.0;
      }
      a->sum[k][p] = (1.0 / ((double)N)) * a->sum[k][p];
    }
}

/* bunch_stats_fprintf, src/bunch.c:791-804 */
void mbt_stats_fprintf(FILE *fp, const mbt_config *c, long rev, const mbt_stat_accum *a, double scan)
{
  fprintf(fp, "%4ld   %.4e   LON    %.4e   %.4e   %.4e   %.4e   %.4e   %.4e   %.4e   %.4e", rev + 1, scan,
          a->sum[A_POS][0], sqrt(a->sum_sqr[A_POS][0]), a->sum[A_SLOPE][0], sqrt(a->sum_sqr[A_SLOPE][0]),
          a->sum[A_PSIG][0], sqrt(a->sum_sqr[A_PSIG][0]), a->sum[A_SSIG][0], sqrt(a->sum_sqr[A_SSIG][0]));
  for (int p = 1; p < 3; p++)
    if (c->TrackPlane[p])
      fprintf(fp, "  %s    %.4e   %.4e   %.4e   %.4e   %.4e   %.4e   %.4e   %.4e", p == 1 ? "HOR" : "VER",
              a->sum[A_POS][p], sqrt(a->sum_sqr[A_POS][p]), a->sum[A_SLOPE][p], sqrt(a->sum_sqr[A_SLOPE][p]),
              a->sum[A_PSIG][p], sqrt(a->sum_sqr[A_PSIG][p]), a->sum[A_SSIG][p], sqrt(a->sum_sqr[A_SSIG][p]));
  fprintf(fp, "\n");
}

/* worker_weak_stat_output, src/bunch.c:772-786 */
void mbt_worker_stat_header(FILE *fp, const mbt_config *c, const double *s, double Ib)
{
  fprintf(fp, " # bunch statistics, averaged over turns, Ib = %g A ;\n", Ib);
  fprintf(fp, " # turn - scan_var - AVE_CM_LON - SIG_CM_LON - AVE_delE_LON - SIG_delE_LON -  AVE_sigmaL_LON - "
              "SIG_sigmaL_LON - AVE_sigmaE_LON - SIG_sigmaE_LON... (_HOR) ... (_VER)\n # ");
  fprintf(fp, "beg      -/-    LON    %.4e       -/-        %.4e       -/-        %.4e      -/-        %.4e        -/- ",
          s[0], s[6], s[3], s[9]);
  if (c->TrackPlane[1])
    fprintf(fp, "    HOR     %.4e       -/-        %.4e       -/-        %.4e      -/-        %.4e        -/- ", s[1],
            s[7], s[4], s[10]);
  if (c->TrackPlane[2])
    fprintf(fp, "    VER     %.4e       -/-       %.4e       -/-        %.4e      -/-        %.4e       -/- ", s[2],
            s[8], s[5], s[11]);
  fprintf(fp, "\n");
}
