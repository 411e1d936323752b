/*
 * Weak-strong drop-in driver: the reference's manager_weakstrong and its
 * worker ranks (src/manager_weakstrong.c:7-263, src/worker_weakstrong.c:14-247)
 * as one host process driving one GPU.  Files and formats are the reference's:
 *
 *   the generated macro-electrons (gendst_datafile[plane], relative to the
 *   working directory, src/bunch.c:318-337), ampinv_{LON,HOR,VER}_<turn>.dat on
 *   monitor turns (src/bunch.c:552-696), CMImotionsH.out / CMImotionsV.out with
 *   FBII (src/fbi_weakstrong.c:583-685), CMmotionsV.out every turn (:688-737),
 *   ampinv_{HOR,VER}_mean.dat and cmpos_{HOR,VER}_mean.dat at the end
 *   (src/bunch.c:703-770), and the manager's progress lines.
 *
 * Reference behaviour kept on purpose: the monitor-turn statistics are stored
 * at [turn/NrevMon * Nbunch + bucket] and the per-turn ampinv files read entry
 * [turn/NrevMon + bucket] of that array (src/manager_weakstrong.c:130-140,
 * src/bunch.c:604); entries the reference never wrote read as 0 here (the
 * reference's array is smaller than its writes when buckets outnumber bunches,
 * and it writes and reads past it).  Bunch currents are the reference's zeros
 * unless MBT_WS_BUNCH_CURRENTS is passed.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "mbt_host.h"

#define NAN_ NAN

static FILE *wsopen(const mbt_config *c, const char *name, const char *mode)
{
  char path[2 * MBT_PATH];
  snprintf(path, sizeof path, "%s/%s", c->work_path, name);
  FILE *fp = fopen(path, mode);
  if (!fp) {
    fprintf(stderr, "Error: cannot open %s !\n", path);
    mbt_set_error("cannot open %s for writing", path);
  }
  return fp;
}

/* generate_initial_sbunch_distribution2's writeout, src/bunch.c:318-337 */
static int ws_write_dist(const mbt_config *c, const double *d)
{
  const int Np = c->Np;
  if (Np == 1) return MBT_OK;
  for (int pl = 0; pl < 3; pl++) {
    if (!c->nGen[pl]) continue;
    FILE *fp = fopen(c->gendst_datafile[pl], "w");
    if (!fp) {
      fprintf(stderr, "Error ! Cannot open %s\n", c->gendst_datafile[pl]);
      return MBT_EIO;
    }
    const double *u = d + (size_t)(2 * pl) * Np, *p = d + (size_t)(2 * pl + 1) * Np;
    for (int j = 0; j < Np; j++) {
      if (pl == 0) fprintf(fp, "%5d %15.12g %15.12g\n", j, u[j] / MBT_FNANO, p[j]);
      else fprintf(fp, "%5d %15.12g %15.12g\n", j, u[j] / MBT_FMILLI, p[j] / MBT_FMILLI);
    }
    fclose(fp);
  }
  return MBT_OK;
}

int mbt_manager_weakstrong(const mbt_config *c, int device, int flags)
{
  const int H = (int)c->Nharm, G = c->Nbunch;
  const long NrevTot = c->NrevTot, NrevMon = c->NrevMon;
  mkdir(c->work_path, 0755);
  printf("Starting job \"%s\"\n\n", c->jobtitle);
  printf("Generating initial distribution...\n");
  fflush(stdout);
  mbt_ws *w = NULL;
  int rc = mbt_ws_create(c, device, flags, &w);
  if (rc) return rc;
  {
    double *d = malloc(sizeof(double) * 6 * (size_t)c->Np);
    mbt_ws_distribution(w, d);
    rc = ws_write_dist(c, d);
    free(d);
    if (rc) {
      mbt_ws_destroy(w);
      return rc;
    }
  }
  /* the manager's monitor arrays (one vector per plane), indexed as it does */
  const long nmon = NrevTot / NrevMon + 1;
  const size_t hn = (size_t)nmon * G + H + 2;
  double(*h_cm)[3] = calloc(hn, sizeof *h_cm), (*h_amp)[3] = calloc(hn, sizeof *h_amp);
  double(*h_ampcm)[3] = calloc(hn, sizeof *h_ampcm);
  const long B = mbt_ws_async_max(w);
  double *mon = malloc(sizeof(double) * 18 * G * B), *ion = malloc(sizeof(double) * 3 * G * B);
  double *zcm = malloc(sizeof(double) * H * B);
  printf("Starting multi bunch tracking...\n");
  fflush(stdout);
  const char *pn[3] = {"LON", "HOR", "VER"};
  /* blocks of B turns: the next block is enqueued before the files of the
     last one are written, so that the writing runs while the GPU tracks */
  long rev = 0, pend = NrevTot < B ? NrevTot : B;
  if (pend > 0) rc = mbt_ws_track_async(w, pend);
  while (rc == MBT_OK && rev < NrevTot) {
    const long n = pend;
    if ((rc = mbt_ws_track_wait(w, NULL, mon, ion, zcm))) break;
    pend = NrevTot - rev - n < B ? NrevTot - rev - n : B;
    if (pend > 0 && (rc = mbt_ws_track_async(w, pend))) break;
    for (long t = 0; t < n; t++, rev++) {
      if (rev % 100 == 0) {
        if (rev + 99 >= NrevTot) printf("Transformation for turn %3ld to %3ld", rev, NrevTot - 1);
        else printf("Transformation for turn %3ld to %3ld", rev, rev + 99);
        fflush(stdout);
      }
      if (rev % NrevMon == 0) {
        const long irm = rev / NrevMon;
        for (int ib = 0; ib < G; ib++) {
          const double *m = mon + ((size_t)t * G + ib) * 18;
          const size_t k = (size_t)irm * G + c->bucket[ib];
          for (int p = 0; p < 3; p++) {
            h_cm[k][p] = m[p];
            h_amp[k][p] = m[12 + p];
            h_ampcm[k][p] = m[15 + p];
          }
        }
        for (int p = 0; p < 3; p++) {
          if (!c->TrackPlane[p]) continue;
          char name[64];
          snprintf(name, sizeof name, "ampinv_%s_%ld.dat", pn[p], rev + 1);
          FILE *fa = wsopen(c, name, "w");
          if (!fa) continue;
          fprintf(fa, "# turn: %-5ld\n", rev + 1);
          fprintf(fa, "# bunch   ampinv_cm   ampinv\n");
          for (int jb = 0; jb < H; jb++)
            if (c->nfFill[jb]) {
              const size_t k = (size_t)irm + jb;
              fprintf(fa, " %3d %15.8e %15.8e\n", jb, k < hn ? h_ampcm[k][p] : 0.0, k < hn ? h_amp[k][p] : 0.0);
            }
          fclose(fa);
        }
      }
      if (c->EnableFBII) {
        FILE *fh = wsopen(c, "CMImotionsH.out", rev == 0 ? "w" : "a");
        FILE *fv = wsopen(c, "CMImotionsV.out", rev == 0 ? "w" : "a");
        if (fh && fv) {
          fprintf(fh, "%3d: ", (int)rev);
          fprintf(fv, "%3d: ", (int)rev);
          for (int ib = 0; ib < G; ib++) {
            const double *s = ion + ((size_t)t * G + ib) * 3;
            if (s[2] == 0.0) {
              fprintf(fh, " %12.8lf", NAN_);
              fprintf(fv, " %12.8lf", NAN_);
            } else {
              fprintf(fh, " %12.8lf", MBT_FKILO * (s[0] / s[2]) / c->sgm_xx);
              fprintf(fv, " %12.8lf", MBT_FKILO * (s[1] / s[2]) / c->sgm_zz);
            }
          }
          putc(0x0a, fh);
          putc(0x0a, fv);
        }
        if (fh) fclose(fh);
        if (fv) fclose(fv);
      }
      FILE *fz = wsopen(c, "CMmotionsV.out", rev == 0 ? "w" : "a");
      if (fz) {
        for (int kb = 0; kb < H; kb++) fprintf(fz, "%-15.8lf ", 10000.0 * MBT_FKILO * zcm[(size_t)t * H + kb]);
        fprintf(fz, "\n");
        fclose(fz);
      }
      if (rev % 10 == 9) printf(".");
      fflush(stdout);
      if (rev % 100 == 99 || rev == NrevTot - 1) printf(" done\n");
    }
  }
  if (rc == MBT_OK)
    for (int p = 1; p < 3; p++) {
      if (!c->TrackPlane[p]) continue;
      char name[64], name2[64];
      snprintf(name, sizeof name, "ampinv_%s_mean.dat", pn[p]);
      snprintf(name2, sizeof name2, "cmpos_%s_mean.dat", pn[p]);
      FILE *fa = wsopen(c, name, "w"), *fpos = wsopen(c, name2, "w");
      if (fa && fpos)
        for (long r0 = 0; r0 < NrevTot; r0 += NrevMon) {
          const long irm = r0 / NrevMon;
          double sp = 0.0, sa = 0.0, sc = 0.0;
          int nb = 0;
          for (int kb = 0; kb < H; kb++)
            if (c->nfFill[kb]) {
              const size_t k = (size_t)irm * G + kb;
              sp += k < hn ? h_cm[k][p] : 0.0;
              sa += k < hn ? h_amp[k][p] : 0.0;
              sc += k < hn ? h_ampcm[k][p] : 0.0;
              nb++;
            }
          double fnb = (double)nb;
          fprintf(fa, " %3ld %15.8e %15.8e\n", r0, sa / fnb, sc / fnb);
          fprintf(fpos, " %3ld %15.8e\n", r0, sp / fnb);
        }
      if (fa) fclose(fa);
      if (fpos) fclose(fpos);
    }
  free(mon);
  free(ion);
  free(zcm);
  free(h_cm);
  free(h_amp);
  free(h_ampcm);
  mbt_ws_destroy(w);
  return rc;
}
