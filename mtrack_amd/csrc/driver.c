/*
 * Weak-weak drop-in driver: the reference's manager rank and all worker ranks
 * (src/manager_weakweak.c:14-350, src/worker_weakweak.c:19-471) as one host
 * process driving one GPU.  The GPU ring tracks all bunches; per-turn
 * statistics come back in blocks of turns and are accumulated and written here
 * with the reference's formats and file names:
 *
 *   mean_bunch_all.dat, mean_bunch_end.dat, bunchbybunch_{LON,HOR,VER}.dat,
 *   ampinv_{HOR,VER}_mean.dat, ampinv_{LON,HOR,VER}_<turn>.dat,
 *   mean_bunch_<kb>.dat, potentials_bunch_<kb>.dat, wakes.dat,
 *   dipoles_history.dat and the final distribution files.
 *
 * Reference quirks reproduced on purpose (they define the file contents):
 *  - bunchbybunch_*.dat prints scan_val_hist[0] on every line (m never advances,
 *    src/bunch_weak.c:368-373);
 *  - ampinv_*_mean.dat and ampinv_*_<turn>.dat index the per-bunch history by
 *    bucket number (src/bunch_weak.c:421-434, src/bunch.c:561-581, 604); entries
 *    beyond the history the reference allocates read as 0 here (the reference
 *    reads past its allocation there);
 *  - "%i" of the turn counter in mean_bunch_end.dat / dipoles_history.dat.
 * Where every kb_out rank of the reference writes the same file (wakes.dat,
 * the final distribution), it is written once, from the last output bunch.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include "mbt_host.h"

#define FKILO MBT_FKILO

typedef struct vec3 {
  double v[3];
} vec3;

static FILE *wopen(const mbt_config *c, const char *name, const char *mode)
{
  char path[2 * MBT_PATH];
  snprintf(path, sizeof path, "%s/%s", c->work_path, name);
  FILE *fp = fopen(path, mode);
  if (!fp) mbt_set_error("cannot open %s for writing", path);
  return fp;
}

/* turns per call of the tracking loop (the library's statistics block) */
#define DRV_BLOCK 256

/* whether the potentials files get a block at the end of turn t (the tables of
   the output turn before its kick, src/transform_weak.c:735-768) */
static int pot_due(const mbt_config *c, int nout, long t)
{
  return nout > 0 && c->Ncell > 0 && (t + 1) % c->NrevMon == 0 &&
         t + 1 >= c->NrevPotentialsOut + (t / c->NrevScan) * c->NrevScan;
}

/* weak_writeout_bunch_distribution, src/bunch_weak.c:295-332 */
static int write_distribution(const mbt_config *c, int plane, const double *soa6, int Np)
{
  char name[3 * MBT_PATH];
  if (c->nGen[plane] == 0) {
    const char *g = c->gendst_datafile[plane];
    size_t st = 0;
    for (size_t i = 0; g[i]; i++)
      if (g[i] == '/') st = i;
    snprintf(name, sizeof name, "%s%s", c->work_path, g + st);
  } else {
    snprintf(name, sizeof name, "%s", c->gendst_datafile[plane]);
  }
  FILE *fp = fopen(name, "w");
  if (!fp) {
    fprintf(stderr, "Error ! Cannot open %s\n", name);
    return MBT_EIO;
  }
  const double pp[3] = {MBT_FGIGA, FKILO, FKILO}, sp[3] = {1.0, FKILO, FKILO};
  const double *pos = soa6 + (size_t)(2 * plane) * Np, *slope = soa6 + (size_t)(2 * plane + 1) * Np;
  for (int j = 0; j < Np; j++) fprintf(fp, "%5d %15.12g %15.12g\n", j, pos[j] * pp[plane], slope[j] * sp[plane]);
  fclose(fp);
  return MBT_OK;
}

static int is_out(const mbt_config *c, int kb)
{
  for (int i = 0; i < c->Nbunch_out; i++)
    if (c->bunch_out[i] == kb) return 1;
  return 0;
}

static void put_vec(vec3 *dst, const double *st12, int off)
{
  for (int p = 0; p < 3; p++) dst->v[p] = st12[off + p];
}

int mbt_manager_weakweak(const mbt_config *c, int device)
{
  const int Nb = c->Nbunch, h = (int)c->Nharm, Nc = c->Ncell;
  const long NrevTot = c->NrevTot, NrevMon = c->NrevMon, NrevScan = c->NrevScan;
  mkdir(c->work_path, 0755);
  printf("Starting job \"%s\"\n\n", c->jobtitle);
  printf("Generating initial distribution...\n");
  fflush(stdout);
  mbt_ring *r = NULL;
  int rc = mbt_ring_create(c, device, 0, &r);
  if (rc) return rc;
  /* turns per library call: whole monitor blocks, up to the library's block of
     DRV_BLOCK turns of statistics (one replayed graph per call) */
  const long blk = NrevMon > 0 && NrevMon < DRV_BLOCK ? (DRV_BLOCK / NrevMon) * NrevMon : (NrevMon > 0 ? NrevMon : 1);
  /* a block can hold one more turn: the previous block's potentials turn */
  double *st = malloc(sizeof(double) * 12 * Nb * (size_t)(blk + 1));
  double *st2 = malloc(sizeof(double) * 12 * Nb * (size_t)(blk + 1));
  double *am = calloc(6 * (size_t)Nb * (blk + 1), sizeof(double));
  double *am2 = calloc(6 * (size_t)Nb * (blk + 1), sizeof(double));
  double *init = malloc(sizeof(double) * 12 * Nb);
  rc = mbt_ring_prepare(r, init);
  if (rc) goto done;

  /* ---- manager: initial statistics over all bunches (src/manager_weakweak.c:105-129) */
  FILE *fall = wopen(c, "mean_bunch_all.dat", "w+");
  if (!fall) {
    rc = MBT_EIO;
    goto done;
  }
  mbt_stat_accum all;
  mbt_accum_reset(&all);
  for (int k = 0; k < Nb; k++) mbt_accum_add(&all, init + 12 * k);
  mbt_accum_update(&all, Nb);
  fprintf(fall, " # bunch statistics, AVEraged over all bunches & turns, Ib = %g A ;\n", c->bunch_Ib[0]);
  fprintf(fall, " # turn - scan_var - AVE_CM_LON - SIG_CM_LON - AVE_delE_LON - SIG_delE_LON -  AVE_sigmaL_LON - "
                "SIG_sigmaL_LON - AVE_sigmaE_LON - SIG_sigmaE_LON... (_HOR) ... (_VER)\n # ");
  mbt_stats_fprintf(fall, c, -100, &all, 0.0);
  fprintf(fall, "\n");
  mbt_accum_reset(&all);

  /* ---- workers with kb_out: files opened at set-up (src/worker_weakweak.c:96-107,225-250) */
  int nout = 0;
  int *outb = calloc(Nb, sizeof(int));
  for (int k = 0; k < Nb; k++)
    if (is_out(c, c->bucket[k])) outb[nout++] = k;
  FILE **fmean = calloc(nout > 0 ? nout : 1, sizeof(FILE *));
  FILE **fpot = calloc(nout > 0 ? nout : 1, sizeof(FILE *));
  mbt_stat_accum *wacc = calloc(nout > 0 ? nout : 1, sizeof(mbt_stat_accum));
  long *wN = calloc(nout > 0 ? nout : 1, sizeof(long));
  double *G = NULL;
  int planes[3] = {0, 0, 0};
  if (Nc > 0) {
    G = calloc(3 * (size_t)Nc, sizeof(double));
    mbt_greens_functions(c, G, G + Nc, G + 2 * Nc, planes);
  }
  if (nout > 0) {
    FILE *fw = wopen(c, "wakes.dat", "w+");
    if (fw) {
      fprintf(fw, "# bin (%.2e s)   total wake L     - V -      - H -\n", c->dT * c->sigma_tau);
      for (int n = 0; n < Nc; n++) fprintf(fw, "%d  %e   %e   %e\n", n, G[n], G[Nc + n], G[2 * Nc + n]);
      fclose(fw);
    }
  }
  for (int i = 0; i < nout; i++) {
    int k = outb[i], kb = c->bucket[k];
    char name[64];
    snprintf(name, sizeof name, "mean_bunch_%d.dat", kb);
    fmean[i] = wopen(c, name, "w+");
    if (!fmean[i]) {
      rc = MBT_EIO;
      goto done;
    }
    mbt_worker_stat_header(fmean[i], c, init + 12 * k, c->bunch_Ib[k]);
    mbt_accum_reset(&wacc[i]);
    if (c->NrevPotentialsOut <= NrevTot) {
      snprintf(name, sizeof name, "potentials_bunch_%d.dat", kb);
      fpot[i] = wopen(c, name, "w+");
      if (fpot[i]) {
        fprintf(fpot[i], " # potentials and distributions, Ib = %g A ;\n", c->bunch_Ib[k]);
        fprintf(fpot[i], " # rev  scan_val         tau         LON_wake_pot       HC_pot        rf_pot      active_cav "
                         "     ideal_HC        bunch_shape");
        if (planes[2] > 0) fprintf(fpot[i], "     VER_wake_pot     VER_dipole_mom");
        if (planes[1] > 0) fprintf(fpot[i], "     HOR_wake_pot     HOR_dipole_mom");
        fprintf(fpot[i], "\n");
      }
    }
  }

  /* ---- history for ampinv / bunch-by-bunch files (src/manager_weakweak.c:25-31) */
  const long Nmonitors = NrevTot / NrevMon;
  const size_t hist_n = (size_t)Nmonitors * Nb + 2;
  vec3 *H_amp = calloc(hist_n, sizeof(vec3)), *H_ampcm = calloc(hist_n, sizeof(vec3)), *H_cm = calloc(hist_n, sizeof(vec3));
  double *scan_hist = calloc((size_t)(NrevTot / NrevMon + c->Nscan + 1), sizeof(double));
  int mscan = 0;
  double scan_val = c->scan_start;
  double Iring = c->Iring;
  double *bunchIb = malloc(sizeof(double) * Nb);
  memcpy(bunchIb, c->bunch_Ib, sizeof(double) * Nb);
  double *pot = Nc > 0 ? malloc(sizeof(double) * 7 * Nc) : NULL;

  printf("Starting multi bunch tracking...\n");
  fflush(stdout);
  long Nstat = 0;
  long rev = 0;  /* turns the per-turn bookkeeping has processed */
  long trev = 0; /* turns planned: the start of the next block of whole monitor blocks */
  const long amax = mbt_ring_async_max(r);
  double *stb[2] = {st, st2}, *amb[2] = {am, am2};
  long pn = 0;    /* turns of the previous block, waiting for the bookkeeping */
  long carry = 0; /* the previous block's potentials turn, tracked at the head of the next block */
  int pb = 0, cb = 0;
  while (trev < NrevTot || carry > 0 || pn > 0) {
    long n = 0, n1 = carry;
    int want_pot = 0, pending = 0;
    if (trev < NrevTot) {
      /* the next block: up to the next monitor boundary, then whole monitor
         blocks more while no potentials output falls due at a block's end;
         the turn of a due output is left for the head of the block after,
         so that the output's tables are read between the two */
      n = NrevMon - (trev % NrevMon);
      if (trev + n > NrevTot) n = NrevTot - trev;
      while (carry + n + NrevMon <= amax && trev + n + NrevMon <= NrevTot && !pot_due(c, nout, trev + n - 1))
        n += NrevMon;
      want_pot = pot_due(c, nout, trev + n - 1);
      n1 = carry + (want_pot ? n - 1 : n);
    }
    if (n1 > 0 && n1 <= amax) {
      /* enqueued now, collected after the previous block's bookkeeping */
      if ((rc = mbt_ring_track_async(r, n1))) goto done;
      pending = 1;
    } else if (n1 > 0 && (rc = mbt_ring_track(r, n1, stb[cb], amb[cb], NULL))) {
      goto done;
    }
    /* per-turn bookkeeping of the manager and the kb_out workers for the
       previous block, while the next one runs */
    const double *st_p = stb[pb], *am_p = amb[pb];
    for (long t = 0; t < pn; t++, rev++) {
      if (rev % 5000 == 0) {
        if (rev + 4999 >= NrevTot) printf("Transformation for turn %3ld to %3ld", rev, NrevTot - 1);
        else printf("Transformation for turn %3ld to %3ld", rev, rev + 4999);
        fflush(stdout);
      }
      const double *s = st_p + (size_t)12 * Nb * t;
      for (int i = 0; i < nout; i++) {
        mbt_accum_add(&wacc[i], s + 12 * outb[i]);
        wN[i]++;
        if (wN[i] % NrevMon == 0 || (rev + 1) % NrevScan == 0) {
          mbt_accum_update(&wacc[i], wN[i]);
          mbt_stats_fprintf(fmean[i], c, rev, &wacc[i], scan_val);
          mbt_accum_reset(&wacc[i]);
          wN[i] = 0;
        }
      }
      for (int k = 0; k < Nb; k++) {
        if ((rev + 1) % NrevMon == 0) {
          long irevmon = rev / NrevMon;
          size_t idx = (size_t)irevmon * Nb + k;
          if (idx < hist_n) {
            for (int p = 0; p < 3; p++) {
              H_amp[idx].v[p] = am_p[(size_t)6 * (Nb * t + k) + p];
              H_ampcm[idx].v[p] = am_p[(size_t)6 * (Nb * t + k) + 3 + p];
            }
            put_vec(&H_cm[idx], s + 12 * k, 0);
          }
        }
        mbt_accum_add(&all, s + 12 * k);
      }
      Nstat++;
      if (Nstat % NrevMon == 0 || (rev + 1) % NrevScan == 0) {
        Nstat *= Nb;
        mbt_accum_update(&all, Nstat);
        scan_hist[mscan++] = scan_val;
        if (c->EnableAmpinv_out) {
          /* bunch_weak_writeout_ampinv, src/bunch.c:588-640 */
          long irevmon = rev / NrevMon;
          const char *pn[3] = {"LON", "HOR", "VER"};
          for (int p = 0; p < 3; p++) {
            if (!c->TrackPlane[p]) continue;
            char name[64];
            snprintf(name, sizeof name, "ampinv_%s_%ld.dat", pn[p], rev + 1);
            FILE *fa = wopen(c, name, "w");
            if (!fa) continue;
            fprintf(fa, "# turn: %-5ld\n", rev + 1);
            fprintf(fa, "# bunch   ampinv_cm   ampinv\n");
            for (int jb = 0; jb < h; jb++)
              if (c->nfFill[jb]) {
                size_t idx = (size_t)irevmon + jb;
                double acm = idx < hist_n ? H_ampcm[idx].v[p] : 0.0, a = idx < hist_n ? H_amp[idx].v[p] : 0.0;
                fprintf(fa, " %3d %15.8e %15.8e\n", jb, acm, a);
              }
            fclose(fa);
          }
        }
        mbt_stats_fprintf(fall, c, rev, &all, scan_val);
        mbt_accum_reset(&all);
        Nstat = 0;
      }
      if ((rev + 1) % NrevScan == 0 && rev > 0 && c->scan == 1) {
        /* tracking_scan_step_manager, src/tracking.c:10-37 */
        Iring += c->scan_step;
        scan_val = Iring;
        for (int k = 0; k < Nb; k++) bunchIb[k] += c->scan_step / (((double)Nb) * FKILO);
      }
      if (rev % 500 == 499) printf(".");
      fflush(stdout);
      if (rev % 5000 == 4999 || rev == NrevTot - 1) {
        time_t now = time(NULL);
        struct tm *tn = localtime(&now);
        printf(" done (at time: %d:%d:%d)\n", tn->tm_hour, tn->tm_min, tn->tm_sec);
      }
    }
    pn = 0;
    if (n1 > 0) {
      if (pending && (rc = mbt_ring_track_wait(r, stb[cb], amb[cb], NULL))) goto done;
      if (want_pot) {
        const long pot_turn = trev + n - 1;
        /* src/transform_weak.c:735-768: tables of the output turn, before its kick */
        for (int i = 0; i < nout; i++) {
          if (!fpot[i]) continue;
          int k = outb[i], kb = c->bucket[k];
          double extra[3];
          if ((rc = mbt_ring_potentials(r, k, pot, extra))) goto done;
          double eps_const = extra[1] / (c->E0 * FKILO);
          double FacI = c->EnableIdealHC == 0 ? 0. : eps_const * c->HC_k;
          for (int ic = 0; ic < Nc; ic++) {
            double taucell = -(-c->Nsigma + c->dT * (ic - 0.5)) * c->sigma_tau;
            double rftest = eps_const * sin(c->wrf * taucell + extra[0]) - c->Urad;
            double active = 0.0;
            for (int a = 0; a < c->n_ahc; a++)
              active += c->ahc[a].Vpeak / (c->E0 * MBT_FGIGA) *
                        sin(c->wrf * c->ahc[a].nHC * taucell + c->ahc[a].phi_aHC + c->ahc[a].nHC * kb * 2 * MBT_PI);
            double ideal = FacI * sin(c->m_aHC * c->wrf * taucell + c->m_aHC * c->phi_n);
            fprintf(fpot[i], "\n %ld   %e   %e   %e    %e   %e   %e   %e   %e", pot_turn + 1, extra[2], taucell,
                    pot[ic], pot[Nc + ic], rftest, active, ideal, pot[2 * Nc + ic]);
            if (planes[2] > 0) fprintf(fpot[i], "   %e   %e", pot[3 * Nc + ic], pot[4 * Nc + ic]);
            if (planes[1] > 0) fprintf(fpot[i], "   %e   %e", pot[5 * Nc + ic], pot[6 * Nc + ic]);
          }
          fprintf(fpot[i], "\n");
        }
      }
      pn = n1;
      pb = cb;
      cb ^= 1;
    }
    carry = want_pot ? 1 : 0;
    trev += n;
  }

  /* ---- end of run */
  double *fin = malloc(sizeof(double) * 12 * Nb);
  if (c->scan == 0) {
    if ((rc = mbt_ring_final_stats(r, fin))) goto done;
    FILE *fe = wopen(c, "mean_bunch_end.dat", "w+");
    if (fe) {
      fprintf(fe, " # bunch statistics end of run, turns = %i; Ib = %g A ;\n", (int)rev, bunchIb[0]);
      fprintf(fe, " # bunch_num    CM    bunch_length    energy_dev    energy_spread;\n");
      for (int k = 0; k < Nb; k++)
        fprintf(fe, " %d     %e    %e    %e    %e\n", c->bucket[k], fin[12 * k], fin[12 * k + 3], fin[12 * k + 6],
                fin[12 * k + 9]);
      fclose(fe);
    }
  }
  {
    /* weak_bunch_writeout_tbtbbb + weak_bunch_writeout_mean_ampinv, src/bunch_weak.c:342-447 */
    const char *pn[3] = {"LON", "HOR", "VER"};
    for (int p = 0; p < 3; p++) {
      if (p > 0 && !c->TrackPlane[p]) continue;
      if (p > 0) {
        char name[64];
        snprintf(name, sizeof name, "ampinv_%s_mean.dat", pn[p]);
        FILE *fm = wopen(c, name, "w");
        if (fm) {
          fprintf(fm, "# turn   AVE_ampinv    AVE_ampinv_cm   current \n");
          int m = 0;
          for (long rv = 0; rv < NrevTot; rv += NrevMon) {
            long irevmon = rv / NrevMon;
            double sa = 0.0, sc = 0.0;
            int nb = 0;
            for (int ib = 0; ib < h; ib++)
              if (c->nfFill[ib]) {
                size_t idx = (size_t)irevmon * Nb + ib;
                sa += idx < hist_n ? H_amp[idx].v[p] : 0.0;
                sc += idx < hist_n ? H_ampcm[idx].v[p] : 0.0;
                nb++;
              }
            double fnb = (double)nb;
            fprintf(fm, " %3ld %15.8e %15.8e  %15.8e\n", rv, sa / fnb, sc / fnb, scan_hist[m]);
            m++;
          }
          fclose(fm);
        }
      }
      char name[64];
      snprintf(name, sizeof name, "bunchbybunch_%s.dat", pn[p]);
      FILE *fb = wopen(c, name, "w");
      if (fb) {
        fprintf(fb, "# turn    position of bunch 1,2,3...\n");
        for (long rv = NrevMon; rv < NrevTot + 1; rv += NrevMon) {
          long irevmon = rv / NrevMon - 1;
          fprintf(fb, " %3ld %15.8e", rv, scan_hist[0]);
          for (int ib = 0; ib < Nb; ib++) fprintf(fb, " %15.8e", H_cm[(size_t)irevmon * Nb + ib].v[p]);
          fprintf(fb, "\n");
        }
        fclose(fb);
      }
    }
  }
  /* kb_out workers: final distribution, dipole history (src/worker_weakweak.c:407-439) */
  if (nout > 0) {
    int k = outb[nout - 1];
    double *soa = malloc(sizeof(double) * 6 * (size_t)c->Np);
    if (mbt_ring_get_bunch(r, k, soa) == MBT_OK) {
      write_distribution(c, 0, soa, c->Np);
      if (c->TrackPlane[1]) write_distribution(c, 1, soa, c->Np);
      if (c->TrackPlane[2]) write_distribution(c, 2, soa, c->Np);
    }
    free(soa);
    if (c->EnableRW_long) {
      double *rh = malloc(sizeof(double) * 2 * (size_t)(c->Nmlt + 2) * h);
      FILE *fd = wopen(c, "dipoles_history.dat", "w+");
      if (fd && mbt_ring_rw_history(r, rh) == MBT_OK) {
        fprintf(fd, " # history of dipole moments at rev = %i; Ib = %g A ;\n", (int)rev, bunchIb[k]);
        fprintf(fd, " # Nhist  kb   rev   dipole_HOR    dipole_VER\n");
        const long nn = (long)(c->Nmlt + 2) * h;
        for (int j = 0; j < c->Nmlt; j++)
          for (int i = 0; i < h; i++) {
            long idx = ((long)j * h + i) % nn;
            fprintf(fd, " %d    %d    %d           %e    %e \n", j * h + i, i, j, rh[2 * idx], rh[2 * idx + 1]);
          }
      }
      if (fd) fclose(fd);
      free(rh);
    }
  }
  {
    int *trash = malloc(sizeof(int) * 2 * Nb);
    double *cnt = malloc(sizeof(double) * (size_t)Nb * (Nc > 0 ? Nc : 1));
    const int have_hist = Nc > 0 && mbt_ring_histograms(r, cnt, trash) == MBT_OK;
    double rf[3] = {0.0, 0.0, 0.0};
    const int have_rf = mbt_ring_rf_state(r, 0, rf) == MBT_OK && rf[2] != 0.0;
    /* each worker's last lines, in bunch order (src/worker_weakweak.c:441-448) */
    for (int k = 0; k < Nb; k++) {
      if (have_hist && (trash[2 * k] != 0 || trash[2 * k + 1] != 0))
        printf("\n WARNING - Binning for selffields bunch #%d:\n %d particles below first bin,    %d particles above "
               "last bin\n",
               c->bucket[k], trash[2 * k], trash[2 * k + 1]);
      if (have_rf && is_out(c, c->bucket[k])) printf("\nFinal RF voltage and phase: %.7f MV, %.5f\n", rf[1], rf[0]);
    }
    free(trash);
    free(cnt);
  }
  free(fin);
  for (int i = 0; i < nout; i++) {
    if (fmean[i]) fclose(fmean[i]);
    if (fpot[i]) fclose(fpot[i]);
  }
  fclose(fall);
  free(H_amp);
  free(H_ampcm);
  free(H_cm);
  free(scan_hist);
  free(bunchIb);
  free(pot);
  free(G);
  free(fmean);
  free(fpot);
  free(wacc);
  free(wN);
  free(outb);
done:
  free(st);
  free(st2);
  free(am);
  free(am2);
  free(init);
  mbt_ring_destroy(r);
  return rc;
}
