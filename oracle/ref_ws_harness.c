/*
 * Weak-strong reference harness: test infrastructure only.
 *
 * Links the reference's own translation units (built in place from
 * /root/reference/src by oracle/Makefile, never copied) and replays
 * manager_weakstrong / worker_weakstrong (src/manager_weakstrong.c:7-263,
 * src/worker_weakstrong.c:14-247) in one process, in the order the MPI
 * messages impose:
 *
 *   set-up   strong_generate_eCMs (written out, as the manager does),
 *            generate_initial_general_distribution2, fbii_model_init
 *   phase 1  per filled bunch: transformSbunch_shortrange, fbi_weakstrong_bunch
 *            (when EnableFBII), statistics on monitor turns; the manager
 *            appends every bunch's CM to the history
 *   phase 2  sortCMdata, the 5000-double history message, resortCMdata;
 *            per bunch: transformSbunch_optic, transformSbunch_longrange,
 *            bunch_strong_updateCM
 *   phase 3  bunch by bunch in train order: fbi_weakstrong_train
 *
 * Per-worker state (bunch copy, local ion beam with its fields and map) is kept
 * per bunch; the ion cloud the manager relays is one shared ion_beam_t.
 *
 * Output (%.17g), in outdir:
 *   cm.txt    turn ib  x x' z z'    (CM each bunch sends after phase 1)
 *   mon.txt   turn ib  pos[3] pos_sigma[3] slope[3] slope_sigma[3] ampinv[3] ampinv_cm[3]
 *   ion.txt   turn ib  sum Nri x, sum Nri y, sum Nri   (after the bunch's train step)
 *   zcm.txt   turn  z CM of every bucket (writeoutVCMmotionsPM2's value before scaling)
 *   slices.txt  ib cell xtau xeps xx xp zz zp npop   (after the last turn)
 *
 * With write_files = 1 the manager's own writers also run, at the points where
 * manager_weakstrong calls them, and write its files into outdir:
 * bunch_strong_writeout_ampinv (monitor turns), writeoutCMImotions (after every
 * bunch of the FBII relay), writeoutVCMmotionsPM2 (every turn),
 * bunch_writeout_mean_ampinv_pos (end), and its progress lines go to stdout.
 * (The MPI program itself cannot run this mode below h = 500: its workers
 * receive 500 doubles into ring.Ibunch, allocated with h entries.)
 *
 * Usage: refwsharness input.conf outdir [nturns] [ibunch_fill] [write_files]
 * ibunch_fill = 1 sets ring.Ibunch from the filling as the weak-weak manager
 * does (src/manager_weakweak.c:58-80); 0 keeps the reference's zeros.
 */
#include <stdio.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <unistd.h>
#include "types.h"
#include "bunch.h"
#include "input.h"
#include "fbi.h"
#include "rw.h"

ring_t ring;
tracking_t track;
grid_t fbii_grid;

int transformSbunch_shortrange(int in, unsigned int kb, const e_beam_t ebeam, const bunch_macroparticle_model_t MPmodel,
                               bunch_strong_t *bunch0, bunch_strong_t *bunch);
int transformSbunch_optic(int in, int kb, const e_beam_t ebeam, bunch_strong_t *bCM);
int transformSbunch_longrange(int in, unsigned int kb, const plane_t plane, const e_beam_t ebeam,
                              const bunch_macroparticle_model_t MPmodel, const bunch_CM_history_strong_t *bCMhist,
                              bunch_strong_t *bCM);
int strong_generate_eCMs(const bunch_macroparticle_model_t MPmodel, e_beam_t *ebeam, bunch_strong_t *bunch0,
                         int writeout);
int generate_initial_general_distribution2(bunch_strong_t *bunches, const bunch_strong_t *bunch0,
                                           const bunch_macroparticle_model_t MPmodel, const ring_t ring,
                                           const e_beam_t ebeam);
int sortCMdata(int in, bunch_CM_history_strong_t *bCM);
int resortCMdata(int in, bunch_CM_history_strong_t *CMhist);
int bunch_strong_distribution_statistics(bunch_strong_t *bunch, int slices);

static void put3(FILE *fp, const vector_t *v)
{
  for (int p = 0; p < 3; p++) fprintf(fp, " %.17g", v->v[p]);
}

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: refwsharness input.conf outdir [nturns] [ibunch_fill]\n");
    return 2;
  }
  long nturns_override = argc > 3 ? atol(argv[3]) : -1;
  int ib_fill = argc > 4 ? atoi(argv[4]) : 0;
  int wfiles = argc > 5 ? atoi(argv[5]) : 0;
  char conf[FILENAME_MAX];
  if (argv[1][0] == '/') snprintf(conf, sizeof conf, "%s", argv[1]);
  else {
    char cwd[FILENAME_MAX];
    if (!getcwd(cwd, sizeof cwd)) return 1;
    snprintf(conf, sizeof conf, "%s/%s", cwd, argv[1]);
  }
  if (chdir(argv[2])) return 1; /* the manager writes its distribution files here */
  snprintf(track.work_path, FILENAME_MAX, ".");

  bunch_macroparticle_model_t mp;
  selffield_model_t sf0;
  e_beam_t ebeam;
  memset(&mp, 0, sizeof mp);
  memset(&sf0, 0, sizeof sf0);
  memset(&ebeam, 0, sizeof ebeam);
  if (!read_input(conf, &ring, &track, &ebeam, &mp, &sf0)) return 1;
  /* the project's reader reads beffH3/beffV3 whenever RW long-range is on
     (src/input.c:465-478 reads them only with EnableRW_short) */
  if (track.EnableRW_long && !track.EnableRW_short) {
    FILE *cf = fopen(conf, "r");
    config_get_fprint = false;
    if (track.TrackPlane[HOR]) config_get_double(cf, "ring", "beffH3", &ring.beffH[0]);
    if (track.TrackPlane[VER]) config_get_double(cf, "ring", "beffV3", &ring.beffV[0]);
    fclose(cf);
  }
  setup_ring_parameters(&ring);
  macrop_model_setup_parameters(ring, track, &mp);
  fbii_grid_setup_parameters(&fbii_grid, &ebeam, mp);
  if (!e_beam_setup(&track, &ring, &ebeam)) return 1;
  if (nturns_override > 0) track.NrevTot = nturns_override;
  if (ib_fill) {
    int k = 0;
    for (unsigned kb = 0; kb < ring.Nharm; kb++)
      if (ebeam.nfFill[kb]) {
        double bunch_Ib = ebeam.Ib_frac[k] * ring.Iring / (((double)ebeam.Nbunch) * FKILO);
        ring.Ibunch[kb] = bunch_Ib * FKILO;
        k++;
      }
  }

  const int slices = 2 * ebeam.distrib.Nslc + 1;
  const int h = (int)ring.Nharm;
  /* manager: history, eCMs, per-bucket distributions */
  bunch_CM_history_strong_t *mh = calloc(1, sizeof *mh), *wh = calloc(1, sizeof *wh);
  bunch_CM_history_strong_init(mh, (int)(track.NrevTot / track.NrevMon * ebeam.Nbunch));
  bunch_CM_history_strong_init(wh, (int)(track.NrevTot / track.NrevMon * ebeam.Nbunch));
  bunch_strong_t bunch0;
  memset(&bunch0, 0, sizeof bunch0);
  if (wfiles) {
    printf("Starting job \"%s\"\n\n", track.jobtitle);
    printf("Generating initial distribution...\n");
  }
  if (strong_generate_eCMs(mp, &ebeam, &bunch0, 1) < 0) return 1;
  bunch_strong_t *bunches = calloc(500, sizeof(bunch_strong_t));
  if (generate_initial_general_distribution2(bunches, &bunch0, mp, ring, ebeam) < 0) return 1;
  fbii_model_t fm;
  fbii_model_init(&fm);
  fm.factBE = sqrt(2 * (mp.sgm_xx * mp.sgm_xx - mp.sgm_zz * mp.sgm_zz)) / FKILO;

  /* workers */
  bunch_strong_t *wb = calloc(h, sizeof(bunch_strong_t));
  ion_beam_local_t **loc = calloc(h, sizeof(ion_beam_local_t *));
  ion_model_t *imodel = calloc(h, sizeof(ion_model_t));
  for (int kb = 0; kb < h; kb++) {
    if (!ebeam.nfFill[kb]) continue;
    loc[kb] = calloc(1, sizeof(ion_beam_local_t));
    int ic = 0;
    const int offst = fbii_grid.offst;
    for (int cell = 0; cell < slices; cell++)
      for (int msx = -offst; msx < offst + 1; msx++)
        for (int msz = -offst; msz < offst + 1; msz++) loc[kb]->mapslice2i[cell][msx + offst][msz + offst] = ic++;
    memcpy(wb[kb].npop, bunch0.npop, sizeof bunch0.npop);
  }
  ion_beam_t *ionbeam = calloc(1, sizeof(ion_beam_t));

  char path[FILENAME_MAX];
  FILE *fcm = fopen("cm.txt", "w"), *fmon = fopen("mon.txt", "w"), *fion = fopen("ion.txt", "w");
  FILE *fz = fopen("zcm.txt", "w");
  (void)path;
  if (wfiles) printf("Starting multi bunch tracking...\n");
  struct timespec t_a, t_b;
  clock_gettime(CLOCK_MONOTONIC, &t_a);
  for (long rev = 0; rev < track.NrevTot; rev++) {
    if (wfiles && rev % 100 == 0) {
      if (rev + 99 >= track.NrevTot) printf("Transformation for turn %3ld to %3ld", rev, track.NrevTot - 1);
      else printf("Transformation for turn %3ld to %3ld", rev, rev + 99);
      fflush(stdout);
    }
    int ib = 0;
    for (int kb = 0; kb < h; kb++) {
      if (!ebeam.nfFill[kb]) continue;
      /* worker receives the manager's arrays into bunch0, then bunch = bunch0 */
      bunch_strong_t b0 = wb[kb];
      memcpy(b0.xtau, bunches[kb].xtau, sizeof b0.xtau);
      memcpy(b0.xeps, bunches[kb].xeps, sizeof b0.xeps);
      memcpy(b0.xx, bunches[kb].xx, sizeof b0.xx);
      memcpy(b0.xp, bunches[kb].xp, sizeof b0.xp);
      memcpy(b0.zz, bunches[kb].zz, sizeof b0.zz);
      memcpy(b0.zp, bunches[kb].zp, sizeof b0.zp);
      bunch_strong_t b = b0;
      transformSbunch_shortrange((int)rev, kb, ebeam, mp, &b0, &b);
      if (track.EnableFBII && fbi_weakstrong_bunch(mp, &imodel[kb], fm, ebeam, loc[kb], &b, kb) < 0)
        fprintf(stderr, "fbi_weakstrong_bunch failed\n");
      fprintf(fcm, "%ld %d %.17g %.17g %.17g %.17g\n", rev, ib, b.stats.pos.x, b.stats.slope.x, b.stats.pos.z,
              b.stats.slope.z);
      bunch_CM_history_append(mh, kb, b.stats.pos.x, b.stats.slope.x, b.stats.pos.z, b.stats.slope.z);
      if (rev % track.NrevMon == 0) {
        bunch_strong_distribution_statistics(&b, slices);
        fprintf(fmon, "%ld %d", rev, ib);
        put3(fmon, &b.stats.pos);
        put3(fmon, &b.stats.pos_sigma);
        put3(fmon, &b.stats.slope);
        put3(fmon, &b.stats.slope_sigma);
        put3(fmon, &b.stats.ampinv);
        put3(fmon, &b.stats.ampinv_cm);
        fprintf(fmon, "\n");
        /* the manager's store (src/manager_weakstrong.c:130-140); its arrays hold
           NrevTot/NrevMon*Nbunch + 2 entries, which bucket-indexed writes can
           pass when buckets outnumber bunches: those are skipped here */
        const int irevmon = rev / track.NrevMon;
        const long nout = (long)(track.NrevTot / track.NrevMon * ebeam.Nbunch) + 2;
        if (wfiles && irevmon * ebeam.Nbunch + kb < nout) {
          mh->cm[irevmon * ebeam.Nbunch + kb] = b.stats.pos;
          mh->rms[irevmon * ebeam.Nbunch + kb] = b.stats.pos_sigma;
          mh->cm_slope[irevmon * ebeam.Nbunch + kb] = b.stats.slope;
          mh->rms_slope[irevmon * ebeam.Nbunch + kb] = b.stats.slope_sigma;
          mh->ampinv[irevmon * ebeam.Nbunch + kb] = b.stats.ampinv;
          mh->ampinv_cm[irevmon * ebeam.Nbunch + kb] = b.stats.ampinv_cm;
        }
      }
      wb[kb] = b;
      bunches[kb] = b;
      ib++;
    }
    if (wfiles && rev % track.NrevMon == 0) bunch_strong_writeout_ampinv(rev, rev / track.NrevMon, mh, ebeam);
    sortCMdata((int)rev, mh);
    memcpy(wh->dataXX, mh->dataXX, 5000 * sizeof(double));
    memcpy(wh->dataXP, mh->dataXP, 5000 * sizeof(double));
    memcpy(wh->dataVV, mh->dataVV, 5000 * sizeof(double));
    memcpy(wh->dataVP, mh->dataVP, 5000 * sizeof(double));
    resortCMdata((int)rev, wh);
    for (int kb = 0; kb < h; kb++) {
      if (!ebeam.nfFill[kb]) continue;
      bunch_strong_t *b = &wb[kb];
      transformSbunch_optic((int)rev, kb, ebeam, b);
      if (track.EnableRW_long) {
        if (track.TrackPlane[HOR]) transformSbunch_longrange((int)rev, kb, HOR, ebeam, mp, wh, b);
        if (track.TrackPlane[VER]) transformSbunch_longrange((int)rev, kb, VER, ebeam, mp, wh, b);
      }
      bunch_strong_updateCM(b, slices, HOR);
      bunch_strong_updateCM(b, slices, VER);
      bunches[kb] = *b;
    }
    ionbeam->model.nFio = 0;
    ionbeam->model.kbFio = 0;
    ib = 0;
    for (int kb = 0; kb < h; kb++) {
      if (!ebeam.nfFill[kb]) continue;
      if (track.EnableFBII) {
        ionbeam->model.Nions = imodel[kb].Nions;
        if (fbi_weakstrong_train(mp, fm, ebeam, loc[kb], ionbeam, &wb[kb], ib, kb) < 0)
          fprintf(stderr, "fbi_weakstrong_train failed\n");
      }
      double sx = 0.0, sy = 0.0, sn = 0.0;
      for (int ip = 0; ip < ionbeam->model.nFio; ip++) {
        sx = sx + ionbeam->NriFio[ip] * ionbeam->xxFio[ip];
        sy = sy + ionbeam->NriFio[ip] * ionbeam->yyFio[ip];
        sn = sn + ionbeam->NriFio[ip];
      }
      fprintf(fion, "%ld %d %.17g %.17g %.17g\n", rev, ib, sx, sy, sn);
      bunches[kb] = wb[kb];
      if (wfiles && track.EnableFBII) writeoutCMImotions(track.work_path, (int)rev, ib, ionbeam, &ebeam, mp);
      ib++;
    }
    if (wfiles) {
      writeoutVCMmotionsPM2(track.work_path, (int)rev, bunches, &ebeam);
      if (rev % 10 == 9) printf(".");
      fflush(stdout);
      if (rev % 100 == 99 || rev == track.NrevTot - 1) printf(" done\n");
    }
    fprintf(fz, "%ld", rev);
    for (int kb = 0; kb < h; kb++) {
      double z = 0.0;
      if (ebeam.nfFill[kb]) {
        for (int cell = 0; cell < slices; cell++) z = z + bunches[kb].zz[cell];
        z = z / ((double)slices);
      }
      fprintf(fz, " %.17g", z);
    }
    fprintf(fz, "\n");
  }
  clock_gettime(CLOCK_MONOTONIC, &t_b);
  fprintf(stderr, "turn_loop_seconds %.6f\n", (t_b.tv_sec - t_a.tv_sec) + 1e-9 * (t_b.tv_nsec - t_a.tv_nsec));
  if (wfiles) {
    if (track.TrackPlane[HOR]) bunch_writeout_mean_ampinv_pos(track.NrevTot, track.NrevMon, mh, ebeam, HOR);
    if (track.TrackPlane[VER]) bunch_writeout_mean_ampinv_pos(track.NrevTot, track.NrevMon, mh, ebeam, VER);
  }
  fclose(fcm);
  fclose(fmon);
  fclose(fion);
  fclose(fz);
  FILE *fs = fopen("slices.txt", "w");
  int ib = 0;
  for (int kb = 0; kb < h; kb++) {
    if (!ebeam.nfFill[kb]) continue;
    for (int cell = 0; cell < slices; cell++)
      fprintf(fs, "%d %d %.17g %.17g %.17g %.17g %.17g %.17g %d\n", ib, cell, bunches[kb].xtau[cell],
              bunches[kb].xeps[cell], bunches[kb].xx[cell], bunches[kb].xp[cell], bunches[kb].zz[cell],
              bunches[kb].zp[cell], bunch0.npop[cell]);
    ib++;
  }
  fclose(fs);
  return 0;
}
