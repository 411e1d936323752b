/*
 * .conf reader and derived machine parameters (host C).
 *
 * Grammar, keys, defaults and required/optional rules follow the reference
 * reader src/input.c:278-1100; derived quantities follow src/input.c:1553-1831
 * (setup_ring_parameters, setup_tracking_parameters,
 * macrop_model_setup_parameters) and the fillings of src/input.c:1289-1512.
 * The derived formulas keep the reference's operation order so that every
 * derived double is bit-identical to the reference's.
 *
 * Grammar (pinned by the reference's src/test_input.c / src/test_input.input):
 *  - a section header is "[name]" as the first word of a line (leading blanks
 *    allowed, the word ends at a blank);
 *  - a key line is "identifier = value" or "identifier : value"; blanks around
 *    the operator are skipped; lines whose first non-blank is '#' or ';' are
 *    comments;
 *  - numeric values are read with scanf semantics (leading number, rest
 *    ignored); string values are the whole rest of the line, verbatim;
 *  - the first parsable occurrence of a key inside any copy of its section wins;
 *    a section counts as present only if it appears exactly once.
 * Deviation (documented): integers of long type are parsed as long; the
 * reference scans "%d" into a long (src/input.c:1075-1081), which only works for
 * zero-initialised non-negative values.
 */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "mbt_host.h"

static __thread char g_err[512];

void mbt_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}

const char *mbt_last_error(void) { return g_err; }

/* ---------------------------------------------------------------- parser */

typedef struct conf_text {
  char **line;
  int n;
  int print_errors;
} conf_text;

static int conf_load(const char *path, conf_text *t)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return -1;
  t->line = NULL;
  t->n = 0;
  int cap = 0;
  char *buf = NULL;
  size_t bcap = 0;
  ssize_t len;
  while ((len = getline(&buf, &bcap, fp)) >= 0) {
    if (len > 0 && buf[len - 1] == '\n') buf[--len] = '\0';
    if (t->n == cap) {
      cap = cap ? 2 * cap : 256;
      t->line = realloc(t->line, cap * sizeof(char *));
    }
    t->line[t->n++] = strdup(buf);
  }
  free(buf);
  fclose(fp);
  return 0;
}

static void conf_free(conf_text *t)
{
  for (int i = 0; i < t->n; i++) free(t->line[i]);
  free(t->line);
}

/* -1: not a section line, 0: another section, 1: this section */
static int scan_section(const char *line, const char *name)
{
  if (line[0] == '#' || line[0] == ';') return -1;
  size_t i = 0;
  while (line[i] == ' ' || line[i] == '\t') i++;
  if (line[i] != '[') return -1;
  i++;
  size_t begin = i;
  while (line[i] != '\0' && line[i] != ' ' && line[i] != '\t') i++;
  if (i == 0 || line[i - 1] != ']') return -1;
  size_t n = (i - 1) - begin; /* name length */
  if (strlen(name) != n) return 0;
  return strncmp(line + begin, name, n) == 0 ? 1 : 0;
}

/* Returns pointer to the value text if the line assigns identifier, else NULL. */
static const char *scan_key(const char *line, const char *key)
{
  size_t i = 0;
  while (line[i] == ' ' || line[i] == '\t') i++;
  if (line[i] == '#' || line[i] == ';') return NULL;
  size_t begin = i;
  while (line[i] != '\0' && line[i] != ' ' && line[i] != '\t' && line[i] != ':' && line[i] != '=') i++;
  size_t n = i - begin;
  while (line[i] == ' ' || line[i] == '\t') i++;
  if (line[i] != ':' && line[i] != '=') return NULL;
  i++;
  while (line[i] == ' ' || line[i] == '\t') i++;
  if (strlen(key) != n || strncmp(line + begin, key, n) != 0) return NULL;
  return line + i;
}

static int have_section(const conf_text *t, const char *sec)
{
  int count = 0;
  for (int i = 0; i < t->n; i++)
    if (scan_section(t->line[i], sec) > 0) count++;
  if (count > 1 && t->print_errors) fprintf(stderr, "ERROR: Section [%s] is not unique !\n", sec);
  return count == 1;
}

enum { K_INT, K_LONG, K_DOUBLE, K_STR };

static int conf_get(const conf_text *t, const char *sec, const char *key, int kind, void *out, size_t outlen)
{
  int in_sec = 0;
  for (int i = 0; i < t->n; i++) {
    int s = scan_section(t->line[i], sec);
    if (s == 0) {
      in_sec = 0;
      continue;
    }
    if (s > 0) {
      in_sec = 1;
      continue;
    }
    if (!in_sec) continue;
    const char *v = scan_key(t->line[i], key);
    if (!v) continue;
    int ok = 0;
    switch (kind) {
    case K_INT: ok = sscanf(v, "%d", (int *)out) == 1; break;
    case K_LONG: ok = sscanf(v, "%ld", (long *)out) == 1; break;
    case K_DOUBLE: ok = sscanf(v, "%lf", (double *)out) == 1; break;
    case K_STR:
      snprintf((char *)out, outlen, "%s", v);
      ok = 1;
      break;
    }
    if (ok) return 1;
  }
  if (t->print_errors) fprintf(stderr, "ERROR: Could not get identifier %s in section [%s]\n", key, sec);
  return 0;
}

#define GET_I(sec, key, p) conf_get(&T, sec, key, K_INT, (p), 0)
#define GET_L(sec, key, p) conf_get(&T, sec, key, K_LONG, (p), 0)
#define GET_D(sec, key, p) conf_get(&T, sec, key, K_DOUBLE, (p), 0)
#define GET_S(sec, key, p) conf_get(&T, sec, key, K_STR, (p), sizeof(p))

static const char *const fill_names[] = {"UNIFORM", "ONETHIRD", "TWOTHIRD", "SINGLE", "SINGLE2",
                                         "SINGLE3", "THREEFOURTH", "ONEFOURTH", "SINGLE10",
                                         "SINGLE20", "SINGLE30", "EVENHALF", "EVENFOURTH"};

static int plane_of(const char *s, int *pl)
{
  if (strcmp(s, "LON") == 0) *pl = MBT_LON;
  else if (strcmp(s, "HOR") == 0) *pl = MBT_HOR;
  else if (strcmp(s, "VER") == 0) *pl = MBT_VER;
  else return 0;
  return 1;
}

/* defaults of read_input for either input format, src/input.c:166-182 */
static void input_defaults(mbt_config *c)
{
  c->Nslc = 0;
  c->fYokoyaV = 1.0;
  c->fYokoyaH = 1.0;
  c->QH0 = 18.20;
  c->QV0 = 10.30;
  c->EnableFBII = 0;
  c->EnableQuantum = 1;
  c->EnableAmpinv_out = 0;
  c->EnableRW_short = 0;
  c->EnableRW_long = 0;
  c->filling = 5; /* SINGLE3 */
  c->triggerRW = 0;
  c->bunch_model = 0;
}

/* One data line of the legacy format (c_strng, src/input.c:89-111): lines whose
 * first character is 'C' or 'c' are comments; the next line is the data.  Reads
 * at most len-1 characters of it (the reference writes past its 81-byte buffer
 * on longer lines, and never returns at end of file). */
static int inp_line(FILE *fp, char *str, size_t len)
{
  int c0 = getc(fp);
  while (c0 == 'C' || c0 == 'c') {
    int c1;
    do c1 = getc(fp);
    while (c1 != '\n' && c1 != EOF);
    if (c1 == EOF) return 0;
    c0 = getc(fp);
  }
  if (c0 == EOF) return 0;
  size_t i = 0;
  while (c0 != '\n' && c0 != EOF) {
    if (i + 1 < len) str[i++] = (char)c0;
    c0 = getc(fp);
  }
  str[i] = 0;
  return 1;
}

/* The legacy fixed-order input (read_input_file, src/input.c:198-276): 21 data
 * lines; every quantity it does not name keeps the defaults above or zero. */
static int read_inp(FILE *fp, mbt_config *c)
{
  char s[MBT_PATH];
  input_defaults(c);
#define LINE()                                                                                   \
  do {                                                                                           \
    if (!inp_line(fp, s, sizeof s)) {                                                            \
      mbt_set_error("legacy input ends early");                                                  \
      return 0;                                                                                  \
    }                                                                                            \
  } while (0)
  LINE();
  snprintf(c->jobtitle, 80, "%s", s);
  LINE();
  sscanf(s, "%lf", &c->E0);
  LINE();
  sscanf(s, "%lf%lf", &c->Lc, &c->rho0);
  LINE();
  sscanf(s, "%lf%lf%lf", &c->beta1[MBT_HOR], &c->beta1[MBT_VER], &c->dispH1);
  LINE();
  sscanf(s, "%lf%lf%lf", &c->alpha1[MBT_HOR], &c->alpha1[MBT_VER], &c->disppH1);
  LINE();
  sscanf(s, "%lf%lf", &c->emittanceH, &c->couplbeta);
  LINE();
  sscanf(s, "%lf%lf", &c->ac, &c->Je);
  LINE();
  sscanf(s, "%lf%lf", &c->Vrf0, &c->frf);
  LINE();
  sscanf(s, "%lf", &c->Iring);
  LINE();
  sscanf(s, "%lf%lf", &c->ApertureH, &c->ApertureV);
  LINE();
  sscanf(s, "%lf%lf", &c->Gzix, &c->Gziz);
  LINE();
  sscanf(s, "%d%d%d", &c->TrackPlane[0], &c->TrackPlane[1], &c->TrackPlane[2]);
  LINE();
  sscanf(s, "%ld%d%ld", &c->NrevTot, &c->Nmlt, &c->NrevMon);
  c->NmultiT = c->Nmlt + 2;
  /* no current scan; the reference leaves NrevScan 0 on this path and then
   * takes every turn number modulo it -- the .conf reader's default instead */
  c->NrevScan = c->NrevTot + 1;
  LINE();
  sscanf(s, "%d", &c->Np);
  c->fNp = (double)c->Np;
  LINE();
  sscanf(s, "%d%d%d%d%d%d", &c->nGen[0], &c->iseed[0], &c->nGen[1], &c->iseed[1], &c->nGen[2], &c->iseed[2]);
  for (int p = 0; p < 3; p++) {
    LINE();
    char w[MBT_PATH];
    if (sscanf(s, "%1023s", w) == 1) snprintf(c->gendst_datafile[p], MBT_PATH, "%s", w);
  }
  LINE();
  sscanf(s, "%lf%lf%lf%lf", &c->xtauCM_offset, &c->xepsCM_offset, &c->sgm_xtau, &c->sgm_xeps);
  LINE();
  sscanf(s, "%lf%lf%lf%lf", &c->xxCM_offset, &c->xpCM_offset, &c->sgm_xx, &c->sgm_xp);
  LINE();
  sscanf(s, "%lf%lf%lf%lf", &c->zzCM_offset, &c->zpCM_offset, &c->sgm_zz, &c->sgm_zp);
#undef LINE
  return 1;
}

/* src/input.c:278-850 */
static int read_conf(conf_text *T_, mbt_config *c)
{
  conf_text T = *T_;
  char sec[32], str[MBT_PATH];
  double tmp;

  input_defaults(c);

  snprintf(sec, sizeof sec, "tracking");
  T.print_errors = 1;
  if (!GET_S(sec, "jobtitle", c->jobtitle)) return 0;
  if (GET_S(sec, "filling", str)) {
    int found = 0;
    for (int i = 0; i < 13; i++)
      if (strcmp(fill_names[i], str) == 0) {
        c->filling = i;
        found = 1;
      }
    if (!found) {
      c->filling = MBT_FILL_FROMFILE;
      snprintf(c->fill_filename, sizeof c->fill_filename, "%s", str);
    }
  }
  if (!GET_S(sec, "BunchModel", str)) return 0;
  if (strcmp(str, "WEAK") == 0) c->bunch_model = 0;
  else if (strcmp(str, "STRONG") == 0) c->bunch_model = 1;
  else return 0;
  if (!GET_I(sec, "TrackPlane_LON", &c->TrackPlane[0])) return 0;
  if (c->TrackPlane[0] == 0) {
    fprintf(stderr, "ERROR: LON tracking switched off, has to be on ALWAYS!\n");
    return 0;
  }
  if (!GET_I(sec, "TrackPlane_HOR", &c->TrackPlane[1])) return 0;
  if (!GET_I(sec, "TrackPlane_VER", &c->TrackPlane[2])) return 0;
  if (!GET_L(sec, "NrevTot", &c->NrevTot)) return 0;
  if (!GET_L(sec, "NrevMon", &c->NrevMon)) return 0;
  if (!GET_I(sec, "EnableRW_short", &c->EnableRW_short)) return 0;
  if (!GET_I(sec, "EnableRW_long", &c->EnableRW_long)) return 0;
  if (!GET_I(sec, "EnableFBII", &c->EnableFBII)) return 0;
  if (!GET_I(sec, "EnableQuantum", &c->EnableQuantum)) return 0;
  if (!GET_I(sec, "EnableIdealHC", &c->EnableIdealHC)) return 0;
  T.print_errors = 0;
  if (!GET_L(sec, "NrevPotentialsOut", &c->NrevPotentialsOut)) c->NrevPotentialsOut = 0;
  if (!GET_I(sec, "EnableRW_short_LON", &c->EnableRW_short_LON)) c->EnableRW_short_LON = c->EnableRW_short;
  if (!GET_I(sec, "triggerRW", &c->triggerRW)) c->triggerRW = 0;
  if (!GET_I(sec, "Nmlt", &c->Nmlt) && c->EnableRW_long) c->Nmlt = 0;
  c->NmultiT = c->Nmlt + 2;
  if (!GET_I(sec, "EnableAmpinv_out", &c->EnableAmpinv_out)) c->EnableAmpinv_out = 0;
  if (!GET_I(sec, "EnableResonators", &c->EnableResonators)) c->EnableResonators = 1;
  if (GET_I(sec, "EnableDiffCurr", &c->EnableDiffCurr)) {
    T.print_errors = 1;
    if (!GET_D(sec, "current_ratio", &c->current_ratio)) return 0;
    if (c->current_ratio < 0. || c->current_ratio > 1.) {
      fprintf(stderr, "ERROR: current_ratio not between 0 and 1.\n");
      return 0;
    }
  }

  snprintf(sec, sizeof sec, "ring");
  T.print_errors = 1;
  if (!GET_D(sec, "E0", &c->E0)) return 0;
  if (!GET_D(sec, "Lc", &c->Lc)) return 0;
  if (!GET_D(sec, "rho0", &c->rho0)) return 0;
  if (!GET_D(sec, "betaH", &c->beta1[MBT_HOR])) return 0;
  if (!GET_D(sec, "betaV", &c->beta1[MBT_VER])) return 0;
  if (!GET_D(sec, "dispH", &c->dispH1)) return 0;
  if (!GET_D(sec, "alphaH", &c->alpha1[MBT_HOR])) return 0;
  if (!GET_D(sec, "alphaV", &c->alpha1[MBT_VER])) return 0;
  if (!GET_D(sec, "disppH", &c->disppH1)) return 0;
  if (!GET_D(sec, "ac", &c->ac)) return 0;
  if (!GET_D(sec, "Je", &c->Je)) return 0;
  if (!GET_D(sec, "Vrf0", &c->Vrf0)) return 0;
  if (!GET_D(sec, "frf", &c->frf)) return 0;
  if (!GET_D(sec, "Iring", &c->Iring)) return 0;
  /* both are looked up before failing (non-short-circuit | in the reference) */
  {
    int a = GET_D(sec, "ApertureH", &c->ApertureH);
    int b = GET_D(sec, "ApertureV", &c->ApertureV);
    if (!a || !b) return 0;
    a = GET_D(sec, "QH0", &c->QH0);
    b = GET_D(sec, "QV0", &c->QV0);
    if (!a || !b) return 0;
  }
  T.print_errors = 0;
  /* src/input.c:414-417: a missing CHV/CVH/CVV zeroes CHH; the missing one keeps 0 */
  if (!GET_D(sec, "CHH", &c->CHH)) c->CHH = 0;
  if (!GET_D(sec, "CHV", &c->CHV)) c->CHH = 0;
  if (!GET_D(sec, "CVH", &c->CVH)) c->CHH = 0;
  if (!GET_D(sec, "CVV", &c->CVV)) c->CHH = 0;
  if (!GET_I(sec, "mIdeal", &c->m_aHC)) c->m_aHC = 3;
  if (!GET_D(sec, "Gzix", &c->Gzix)) {
    T.print_errors = 1;
    if (GET_D(sec, "Cx", &tmp)) c->Gzix = tmp / c->QH0;
    else return 0;
  }
  T.print_errors = 0;
  if (!GET_D(sec, "Gziz", &c->Gziz)) {
    T.print_errors = 1;
    if (GET_D(sec, "Cz", &tmp)) c->Gziz = tmp / c->QV0;
    else return 0;
  }
  if (!GET_D(sec, "emittanceH", &c->emittanceH)) return 0;
  T.print_errors = 0;
  if (!GET_D(sec, "couplbeta", &c->couplbeta)) {
    T.print_errors = 1;
    if (GET_D(sec, "emittanceV", &tmp)) c->couplbeta = tmp / c->emittanceH;
    else return 0;
  }
  if (c->EnableRW_short == 1 || c->EnableRW_long == 1 || c->EnableRW_short_LON) {
    if (!GET_D(sec, "rhorw", &c->rhorw)) return 0;
    if (!GET_D(sec, "beffL1", &c->beffL[0])) return 0;
    if (!GET_D(sec, "beffL2", &c->beffL[1])) return 0;
    /* The reference reads beffH3/beffV3 only with EnableRW_short (src/input.c:465-478)
       although the long-range RW kick divides by them (src/transform_weak.c:801-811);
       with RW long-range alone they are read here as well (superset read). */
    int need_h = c->TrackPlane[MBT_HOR] && (c->EnableRW_short || c->EnableRW_long);
    int need_v = c->TrackPlane[MBT_VER] && (c->EnableRW_short || c->EnableRW_long);
    if (need_h) {
      if (!GET_D(sec, "beffH3", &c->beffH[0])) return 0;
      if (c->EnableRW_short && !GET_D(sec, "beffH4", &c->beffH[1])) return 0;
      if (!c->EnableRW_short) {
        T.print_errors = 0;
        GET_D(sec, "beffH4", &c->beffH[1]);
        T.print_errors = 1;
      }
    }
    if (need_v) {
      if (!GET_D(sec, "beffV3", &c->beffV[0])) return 0;
      if (c->EnableRW_short && !GET_D(sec, "beffV4", &c->beffV[1])) return 0;
      if (!c->EnableRW_short) {
        T.print_errors = 0;
        GET_D(sec, "beffV4", &c->beffV[1]);
        T.print_errors = 1;
      }
    }
  } else {
    c->rhorw = 0.0;
    c->beffL[0] = c->beffL[1] = 0.0;
    c->beffH[0] = c->beffH[1] = 0.0;
    c->beffV[0] = c->beffV[1] = 0.0;
  }
  if (c->EnableIdealHC != 1) c->m_aHC = 1;
  T.print_errors = 0;
  GET_D(sec, "fYokoyaV", &c->fYokoyaV);
  GET_D(sec, "fYokoyaH", &c->fYokoyaH);
  c->U0 = -1.0;
  c->taue = -1.0;
  if (GET_D(sec, "U0", &tmp)) c->U0 = tmp * MBT_FKILO;
  if (GET_D(sec, "taue", &tmp)) c->taue = tmp * MBT_FMILLI;

  snprintf(sec, sizeof sec, "distribution");
  T.print_errors = 1;
  if (!GET_I(sec, "Nslc", &c->Nslc)) return 0;
  if (!GET_I(sec, "Np", &c->Np)) return 0;
  c->fNp = (double)c->Np;
  T.print_errors = 0;
  if (!GET_I(sec, "modeHT", &c->modeHT)) c->mode_excitation = 0;
  else c->mode_excitation = 1;
  T.print_errors = 1;
  {
    int a = GET_I(sec, "nGen_LON", &c->nGen[0]);
    int b = GET_I(sec, "nGen_HOR", &c->nGen[1]);
    int d = GET_I(sec, "nGen_VER", &c->nGen[2]);
    if (!a || !b || !d) return 0;
  }
  T.print_errors = 0;
  int tmp_seed = (int)time(NULL);
  if (!GET_I(sec, "iseed_LON", &c->iseed[0])) c->iseed[0] = tmp_seed;
  if (!GET_I(sec, "iseed_HOR", &c->iseed[1])) c->iseed[1] = tmp_seed;
  if (!GET_I(sec, "iseed_VER", &c->iseed[2])) c->iseed[2] = tmp_seed;
  T.print_errors = 1;
  {
    int a = GET_S(sec, "gendst_datafile_LON", c->gendst_datafile[0]);
    int b = GET_S(sec, "gendst_datafile_HOR", c->gendst_datafile[1]);
    int d = GET_S(sec, "gendst_datafile_VER", c->gendst_datafile[2]);
    if (!a || !b || !d) return 0;
  }
  {
    int a = GET_D(sec, "xtauCM_offset", &c->xtauCM_offset);
    int b = GET_D(sec, "xepsCM_offset", &c->xepsCM_offset);
    int d = GET_D(sec, "sgm_xtau", &c->sgm_xtau);
    int e = GET_D(sec, "sgm_xeps", &c->sgm_xeps);
    if (!a || !b || !d || !e) return 0;
    a = GET_D(sec, "xxCM_offset", &c->xxCM_offset);
    b = GET_D(sec, "xpCM_offset", &c->xpCM_offset);
    d = GET_D(sec, "sgm_xx", &c->sgm_xx);
    e = GET_D(sec, "sgm_xp", &c->sgm_xp);
    if (!a || !b || !d || !e) return 0;
    a = GET_D(sec, "zzCM_offset", &c->zzCM_offset);
    b = GET_D(sec, "zpCM_offset", &c->zpCM_offset);
    d = GET_D(sec, "sgm_zz", &c->sgm_zz);
    e = GET_D(sec, "sgm_zp", &c->sgm_zp);
    if (!a || !b || !d || !e) return 0;
  }

  /* [resonator_N] */
  T.print_errors = 1;
  int i = 1;
  snprintf(sec, sizeof sec, "resonator_%d", i);
  while (have_section(&T, sec) && c->EnableResonators != 0) {
    if (i > MBT_MAX_RES) {
      mbt_set_error("more than %d resonators", MBT_MAX_RES);
      return 0;
    }
    mbt_resonator *r = &c->res[i - 1];
    if (!GET_D(sec, "Rs", &r->Rs)) return 0;
    if (!GET_D(sec, "fres", &tmp)) return 0;
    r->wr = (tmp * MBT_FGIGA) * 2.0 * MBT_PI;
    if (!GET_D(sec, "Qfactor", &r->Qfactor)) return 0;
    if (r->Qfactor == 0.5) fprintf(stderr, "WARNING: resonator %d has Qfactor = 0.5 and will be ignored! \n", i);
    if (!GET_S(sec, "plane", str) || !plane_of(str, &r->plane)) return 0;
    i++;
    snprintf(sec, sizeof sec, "resonator_%d", i);
  }
  c->n_res = i - 1;

  i = 1;
  snprintf(sec, sizeof sec, "wake_file_%d", i);
  while (have_section(&T, sec)) {
    if (i > MBT_MAX_WAKEFILE) return 0;
    if (!GET_S(sec, "filename", c->wakefile[i - 1])) return 0;
    if (!GET_S(sec, "plane", str) || !plane_of(str, &c->wakefile_plane[i - 1])) return 0;
    i++;
    snprintf(sec, sizeof sec, "wake_file_%d", i);
  }
  c->n_wakefile = i - 1;

  /* [harmonic_cavity_N] then [longrange_resonator_N] (LON ones appended) */
  i = 1;
  snprintf(sec, sizeof sec, "harmonic_cavity_%d", i);
  while (have_section(&T, sec)) {
    if (i > MBT_MAX_LR) return 0;
    mbt_lr_resonator *r = &c->lr[MBT_LON][i - 1];
    if (!GET_I(sec, "m", &r->m)) return 0;
    if (!GET_D(sec, "Rs", &r->Rs)) return 0;
    if (!GET_D(sec, "detune", &r->detune)) return 0;
    if (!GET_D(sec, "Qfactor", &r->Qfactor)) return 0;
    i++;
    snprintf(sec, sizeof sec, "harmonic_cavity_%d", i);
  }
  c->n_lr[MBT_LON] = i - 1;
  int nL = c->n_lr[MBT_LON], nH = 0, nV = 0;
  i = 1;
  snprintf(sec, sizeof sec, "longrange_resonator_%d", i);
  while (have_section(&T, sec)) {
    int pl;
    if (!GET_S(sec, "plane", str) || !plane_of(str, &pl)) return 0;
    mbt_lr_resonator *r;
    if (pl == MBT_LON) r = &c->lr[MBT_LON][nL++];
    else if (pl == MBT_HOR) r = &c->lr[MBT_HOR][nH++];
    else r = &c->lr[MBT_VER][nV++];
    if (nL > MBT_MAX_LR || nH > MBT_MAX_LR || nV > MBT_MAX_LR) return 0;
    r->m = 0;
    r->detune = 0;
    if (!GET_D(sec, "fres", &tmp)) return 0;
    r->wr = (tmp * MBT_FGIGA) * 2.0 * MBT_PI;
    if (!GET_D(sec, "Rs", &r->Rs)) return 0;
    if (!GET_D(sec, "Qfactor", &r->Qfactor)) return 0;
    i++;
    snprintf(sec, sizeof sec, "longrange_resonator_%d", i);
  }
  c->n_lr[MBT_LON] = nL;
  c->n_lr[MBT_HOR] = nH;
  c->n_lr[MBT_VER] = nV;

  snprintf(sec, sizeof sec, "rf_feedback");
  c->has_rf_feedback = 0;
  if (have_section(&T, sec)) {
    c->has_rf_feedback = 1;
    if (!GET_I(sec, "resonator", &c->rffb_resonator)) return 0;
    if (!GET_I(sec, "averaging_length", &c->rffb_len_average)) return 0;
  }

  T.print_errors = 1;
  i = 1;
  snprintf(sec, sizeof sec, "active_HC_%d", i);
  while (have_section(&T, sec)) {
    if (i > MBT_MAX_AHC) return 0;
    mbt_active_hc *a = &c->ahc[i - 1];
    if (!GET_D(sec, "Vpeak", &a->Vpeak)) return 0;
    if (!GET_D(sec, "nHC", &a->nHC)) return 0;
    if (!GET_D(sec, "phi_aHC", &a->phi_aHC)) return 0;
    i++;
    snprintf(sec, sizeof sec, "active_HC_%d", i);
  }
  c->n_ahc = i - 1;

  snprintf(sec, sizeof sec, "selffield");
  c->RsisL = c->aindL = c->RsisV = c->aindV = c->RsisH = c->aindH = 0.0;
  T.print_errors = 0;
  GET_D(sec, "RsisL", &c->RsisL);
  if (GET_D(sec, "aindL", &tmp)) c->aindL = tmp / (2.0 * MBT_PI * MBT_FGIGA);
  GET_D(sec, "RsisV", &c->RsisV);
  if (GET_D(sec, "aindV", &tmp)) c->aindV = tmp / (2.0 * MBT_PI * MBT_FGIGA);
  GET_D(sec, "RsisH", &c->RsisH);
  if (GET_D(sec, "aindH", &tmp)) c->aindH = tmp / (2.0 * MBT_PI * MBT_FGIGA);
  T.print_errors = 1;
  int lr_tot = c->n_lr[0] + c->n_lr[1] + c->n_lr[2];
  if (c->n_res > 0 || c->EnableRW_short == 1 || lr_tot > 0 || c->n_wakefile > 0 || c->EnableRW_short_LON) {
    if (!have_section(&T, sec)) return 0;
    int a = GET_D(sec, "Nsigma", &c->Nsigma);
    int b = GET_I(sec, "Ncell", &c->Ncell);
    if (!a || !b) return 0;
  } else {
    c->Nsigma = 1.0;
    c->Ncell = 0;
  }

  c->scan = 0;
  c->Nscan = 1;
  c->NrevScan = c->NrevTot + 1;
  c->scan_start = 0.0;
  c->scan_step = 0.0;
  snprintf(sec, sizeof sec, "scan");
  if (have_section(&T, sec)) {
    if (!GET_I(sec, "scan_type", &c->scan)) return 0;
    if (c->scan != 0) {
      if (!GET_L(sec, "Nscan", &c->Nscan)) return 0;
      if (!GET_L(sec, "NrevScan", &c->NrevScan)) return 0;
      if (c->scan == 1) {
        if (!GET_D(sec, "scan_start", &c->scan_start)) return 0;
        if (!GET_D(sec, "scan_step", &c->scan_step)) return 0;
      }
      if (c->scan == 4)
        if (!GET_D(sec, "scan_start", &c->scan_start)) return 0;
      c->NrevTot = c->Nscan * c->NrevScan;
    }
  }
  return 1;
}

/* ------------------------------------------------------- derived parameters */

/* src/input.c:1553-1695 */
static int setup_ring(mbt_config *c)
{
  const double Cg = 8.85e-05;
  c->R = c->Lc / (2.0 * MBT_PI);
  c->Gamma = 1957.0 * c->E0;
  c->Gamma2 = c->Gamma * c->Gamma;
  c->Gamma3 = c->Gamma * c->Gamma2;
  c->wrf = 2.0 * MBT_PI * c->frf * MBT_FMEGA;
  c->h = ((int)(c->frf * MBT_FMEGA * c->Lc / MBT_C_LIGHT + 0.5));
  c->Nharm = (unsigned)c->h;
  c->T0 = c->Nharm / c->frf / MBT_FMEGA;
  c->w0 = (2.0 * MBT_PI) / c->T0;
  c->AmpDependdQH = (c->CHH != 0.0) || (c->CHV != 0.0);
  c->AmpDependdQV = (c->CVH != 0.0) || (c->CVV != 0.0);
  if (c->U0 >= 0 && c->taue >= 0) {
    c->Urad = c->U0 / (c->E0 * MBT_FGIGA);
    if (c->Urad != 0) {
      if (c->Je < 0) c->Je = 2 * c->T0 / (c->taue * c->Urad);
      if (c->rho0 < 0) c->rho0 = Cg * c->E0 * c->E0 * c->E0 * c->E0 * MBT_FGIGA / c->U0;
    }
    c->De = 2 * c->T0 / (c->taue);
    c->tauC = c->Je * c->taue;
    c->aexpe = 1.0 / c->taue;
  } else {
    c->U0 = Cg * c->E0 * c->E0 * c->E0 * c->E0 * MBT_FGIGA / c->rho0;
    c->Urad = c->U0 / (c->E0 * MBT_FGIGA);
    c->tauC = 4.0 * MBT_PI * c->R * c->rho0 / (Cg * MBT_C_LIGHT) / pow(c->E0, 3);
    c->taue = c->tauC / c->Je;
    c->aexpe = 1.0 / c->taue;
    c->De = 2.0 * c->aexpe * c->T0;
  }
  c->q = c->U0 / c->Vrf0 / MBT_FMEGA;
  c->Fq = 2.0 * (sqrt(1.0 / c->q / c->q - 1.0) - acos(c->q));
  c->epsMax = c->U0 / (MBT_PI * c->ac * c->h * c->E0 * MBT_FGIGA) * c->Fq;
  c->epsMax = sqrt(c->epsMax);
  c->Vrfp = MBT_FMEGA * c->Vrf0 * c->wrf * sqrt(1.0 - c->q * c->q);
  c->wso = sqrt(fabs(c->ac) * c->Vrfp / (c->T0 * c->E0 * MBT_FGIGA));
  c->wso2 = c->wso * c->wso;
  c->fso = c->wso / (2.0 * MBT_PI);
  c->Qso = c->fso * c->T0;
  c->wgziH = c->QH0 * c->w0 * c->Gzix / c->ac;
  c->wgziV = c->QV0 * c->w0 * c->Gziz / c->ac;
  c->sn0 = c->q;
  c->phai0 = asin(c->q);
  c->fc1 = c->ac / c->wso;
  c->fc12 = pow(c->fc1, 2);
  c->Jz = 1.0;
  c->Jx = 4.0 - c->Jz - c->Je;
  c->taux = c->tauC / c->Jx;
  c->tauz = c->tauC / c->Jz;
  c->gamma1[MBT_HOR] = (1.0 + c->alpha1[MBT_HOR] * c->alpha1[MBT_HOR]) / c->beta1[MBT_HOR];
  c->gamma1[MBT_VER] = (1.0 + c->alpha1[MBT_VER] * c->alpha1[MBT_VER]) / c->beta1[MBT_VER];

  unsigned Ntmp = 0;
  double vb = 0, genphase = 0, genphase_harm = 0, eloss_harm = 0;
  for (int i = 1; i <= c->n_lr[MBT_LON]; i++) {
    mbt_lr_resonator *r = &c->lr[MBT_LON][i - 1];
    if (r->m > 0) {
      r->wr = r->m * c->wrf + r->detune * 2 * MBT_PI;
      if (r->m > 1) {
        genphase_harm = cos(atan(r->Qfactor * (r->wr / (c->wrf * r->m) - (c->wrf * r->m) / r->wr)));
        eloss_harm += 2 * c->Iring * r->Rs * genphase_harm * genphase_harm * MBT_FMILLI / MBT_FMEGA;
      } else {
        genphase = atan(r->Qfactor * (r->wr / c->wrf - c->wrf / r->wr));
        vb = 2 * c->Iring * r->Rs * cos(genphase) * MBT_FMILLI / MBT_FMEGA;
      }
    }
    double tmp = (r->wr * 0.5 / r->Qfactor);
    r->Nturn = (unsigned)(log(2) * 10 / tmp / c->T0) + 1;
    r->Nbu = r->Nturn * c->h;
    if (r->Nbu < 2) {
      mbt_set_error("Given longrange resonator is shortrange!");
      fprintf(stderr, "Given longrange resonator is shortrange!\n");
      return 0;
    }
    if (r->Nbu > Ntmp) Ntmp = r->Nbu;
  }
  c->phai0 = asin(c->q + eloss_harm / c->Vrf0);
  c->Nbumax = Ntmp;
  if (c->has_rf_feedback) {
    c->rffb_vrf_design = 1 * c->Vrf0;
    if (c->ac < 0) c->rffb_phi0_design = MBT_PI - c->phai0;
    else c->rffb_phi0_design = 1 * c->phai0;
    c->Vrf0 = sqrt(vb * vb + c->rffb_vrf_design * c->rffb_vrf_design +
                   2 * vb * c->rffb_vrf_design * sin(genphase + c->rffb_phi0_design));
    c->phai0 = asin(vb / c->Vrf0 * sin(MBT_PI / 2 + c->rffb_phi0_design + genphase)) + c->rffb_phi0_design;
    if (c->rffb_len_average == -1) c->has_rf_feedback = 0;
    else if (fabs((double)c->rffb_len_average) > c->Nbumax) c->Nbumax = (unsigned)fabs((double)c->rffb_len_average);
  } else {
    c->phai0 += asin(vb / c->Vrf0 * cos(c->phai0 + genphase));
  }
  if (c->ac < 0) c->phai0 = MBT_PI - c->phai0;
  return 1;
}

/* src/input.c:1775-1831 */
static void setup_macrop(mbt_config *c)
{
  const double Cq = 3.84e-13;
  if (c->TrackPlane[MBT_LON] && c->sgm_xeps < 0.0) c->sgm_xeps = sqrt(Cq * c->Gamma2 / (c->Je * c->rho0));
  if (c->TrackPlane[MBT_LON] && c->sgm_xtau < 0.0) c->sgm_xtau = MBT_FGIGA * fabs(c->ac) / c->wso * c->sgm_xeps;
  if (c->sgm_xx < 0.0) c->sgm_xx = MBT_FKILO * sqrt(c->emittanceH / MBT_FGIGA * c->beta1[MBT_HOR]);
  if (c->sgm_xp < 0.0) {
    c->sgm_xp = MBT_FKILO * sqrt(c->emittanceH / MBT_FGIGA / c->beta1[MBT_HOR]);
    c->corr_xpx = -c->alpha1[MBT_HOR] / c->beta1[MBT_HOR];
  } else c->corr_xpx = 0.0;
  if (c->sgm_zz < 0.0) c->sgm_zz = MBT_FKILO * sqrt(c->emittanceH / MBT_FGIGA * c->beta1[MBT_VER] * c->couplbeta);
  if (c->sgm_zp < 0.0) {
    c->sgm_zp = MBT_FKILO * sqrt(c->emittanceH / MBT_FGIGA / c->beta1[MBT_VER] * c->couplbeta);
    c->corr_zpz = -c->alpha1[MBT_VER] / c->beta1[MBT_VER];
  } else c->corr_zpz = 0.0;
  c->sgmatau = c->sgm_xtau / MBT_FGIGA;
  c->corr_epstau = 0.0;
  c->pos_offset[0] = c->xtauCM_offset * MBT_FNANO;
  c->pos_offset[1] = c->xxCM_offset * MBT_FMILLI;
  c->pos_offset[2] = c->zzCM_offset * MBT_FMILLI;
  c->slope_offset[0] = c->xepsCM_offset;
  c->slope_offset[1] = c->xpCM_offset * MBT_FMILLI;
  c->slope_offset[2] = c->zpCM_offset * MBT_FMILLI;
  c->pos_sgm[0] = c->sgm_xtau * MBT_FNANO;
  c->pos_sgm[1] = c->sgm_xx * MBT_FMILLI;
  c->pos_sgm[2] = c->sgm_zz * MBT_FMILLI;
  c->slope_sgm[0] = c->sgm_xeps;
  c->slope_sgm[1] = c->sgm_xp * MBT_FMILLI;
  c->slope_sgm[2] = c->sgm_zp * MBT_FMILLI;
  c->correlate[0] = c->corr_epstau;
  c->correlate[1] = c->corr_xpx;
  c->correlate[2] = c->corr_zpz;
}

/* src/input.c:1289-1512 */
static int setup_filling(mbt_config *c)
{
  unsigned H = c->Nharm;
  c->nfFill = calloc(H, sizeof(int));
  int *out = NULL, nout = 0;
  if (c->filling == MBT_FILL_FROMFILE) {
    FILE *fp = fopen(c->fill_filename, "r");
    if (!fp) {
      fprintf(stderr, "ERROR: %s is not a valid filename or filling pattern.\n", c->fill_filename);
      mbt_set_error("%s is not a valid filename or filling pattern", c->fill_filename);
      return 0;
    }
    int *bucket_out = calloc(H, sizeof(int));
    double *ratio = calloc(H, sizeof(double));
    double tot = 0.0;
    int nb = 0, nbo = 0;
    unsigned kb;
    for (kb = 0; kb < H; kb++) {
      int b, o;
      double r;
      if (fscanf(fp, "%d %lf %d\n", &b, &r, &o) != 3) break;
      if (b >= (int)H) break;
      while (b > (int)kb) {
        bucket_out[kb] = 0;
        ratio[kb] = 0;
        kb++;
      }
      if (r > 0) {
        c->nfFill[kb] = 1;
        ratio[kb] = r;
        tot += r;
        nb++;
        if (o) {
          bucket_out[kb] = 1;
          nbo++;
        } else bucket_out[kb] = 0;
      } else {
        bucket_out[kb] = 0;
        ratio[kb] = 0.0;
      }
    }
    fclose(fp);
    double mean = tot / nb;
    c->Nbunch_out = nbo;
    c->bunch_out = calloc(nbo > 0 ? nbo : 1, sizeof(int));
    c->Nbunch = nb;
    c->Ib_frac = calloc(nb > 0 ? nb : 1, sizeof(double));
    nb = nbo = 0;
    for (kb = 0; kb < H; kb++)
      if (ratio[kb] > 0) {
        c->Ib_frac[nb++] = ratio[kb] / mean;
        if (bucket_out[kb]) c->bunch_out[nbo++] = kb;
      }
    free(bucket_out);
    free(ratio);
  } else {
    unsigned n = 0, step = 1;
    int o3[3];
    switch (c->filling) {
    case 0: n = H; nout = 1; o3[0] = 0; break;
    case 7: n = H / 4; nout = 3; o3[0] = 0; o3[1] = (int)H / 8; o3[2] = (int)H / 4 - 1; break;
    case 1: n = H / 3; nout = 3; o3[0] = 0; o3[1] = (int)H / 6; o3[2] = (int)H / 3 - 1; break;
    case 2: n = 2 * H / 3; nout = 3; o3[0] = 0; o3[1] = (int)H / 3; o3[2] = (int)(2 * H / 3) - 1; break;
    case 6: n = 3 * H / 4; nout = 3; o3[0] = 0; o3[1] = (int)(3 * H / 8); o3[2] = (int)(3 * H / 4) - 1; break;
    case 3: n = 1; nout = 1; o3[0] = 0; break;
    case 4: n = 2; nout = 2; o3[0] = 0; o3[1] = 1; break;
    case 5: n = 3; nout = 3; o3[0] = 0; o3[1] = 1; o3[2] = 2; break;
    case 8: n = 10; nout = 3; o3[0] = 0; o3[1] = 4; o3[2] = 9; break;
    case 9: n = 20; nout = 3; o3[0] = 0; o3[1] = 9; o3[2] = 19; break;
    case 10: n = 30; nout = 3; o3[0] = 0; o3[1] = 14; o3[2] = 29; break;
    case 11: n = H / 2; step = 2; nout = 1; o3[0] = 0; break;
    case 12: n = H / 4; step = 4; nout = 1; o3[0] = 0; break;
    default: return 0;
    }
    if ((n - 1) * step >= H) {
      mbt_set_error("filling pattern does not fit the harmonic number %u", H);
      return 0;
    }
    for (unsigned k = 0; k < n; k++) c->nfFill[k * step] = 1;
    c->Nbunch_out = nout;
    out = calloc(nout, sizeof(int));
    for (int k = 0; k < nout; k++) out[k] = o3[k];
    c->bunch_out = out;
    c->Nbunch = 0;
    for (unsigned k = 0; k < H; k++)
      if (c->nfFill[k]) c->Nbunch++;
    c->Ib_frac = calloc(c->Nbunch, sizeof(double));
    for (int k = 0; k < c->Nbunch; k++) {
      if (c->EnableDiffCurr) c->Ib_frac[k] = (k % 2 == 0) ? c->current_ratio : 2.0 - c->current_ratio;
      else c->Ib_frac[k] = 1.0;
    }
  }
  if (c->Nbunch <= 0) {
    mbt_set_error("no filled bucket");
    return 0;
  }
  c->Ibunch = calloc(H, sizeof(double));
  c->bunch_Ib = calloc(c->Nbunch, sizeof(double));
  c->bucket = calloc(c->Nbunch, sizeof(int));
  for (unsigned k = 0, n = 0; k < H; k++)
    if (c->nfFill[k]) c->bucket[n++] = k;
  return 1;
}

/* Bunch currents as the manager sends them, src/manager_weakweak.c:58-93: after
 * setup_tracking_parameters, which may replace Iring by the scan start. */
static void setup_currents(mbt_config *c)
{
  for (int k = 0; k < c->Nbunch; k++) {
    c->bunch_Ib[k] = c->Ib_frac[k] * c->Iring / (((double)c->Nbunch) * MBT_FKILO);
    c->Ibunch[c->bucket[k]] = c->bunch_Ib[k] * MBT_FKILO;
  }
}

/* src/input.c:1698-1772 */
static int setup_tracking(mbt_config *c)
{
  if (c->EnableIdealHC == 1) {
    c->phai0 = asin(c->m_aHC * c->m_aHC / (c->m_aHC * c->m_aHC - 1.) * c->q);
    c->phi_n = atan(tan(c->phai0) / (double)c->m_aHC) / (double)c->m_aHC;
    if (c->ac < 0) {
      c->phai0 = MBT_PI - c->phai0;
      c->phi_n = MBT_PI - c->phi_n;
    }
    c->HC_k = -cos(c->phai0) / (c->m_aHC * cos(c->m_aHC * c->phi_n));
  }
  if (c->scan == 1) c->Iring = c->scan_start;
  if (c->scan == 3 || c->scan == 4) {
    if (c->n_lr[MBT_LON] < 1) {
      fprintf(stderr, "Qfactor scan but no longrange resonator given!\n");
      return 0;
    }
    if (c->scan == 3) {
      c->scan_start = 1. / c->Nscan;
      c->scan_step = 1. / c->Nscan;
      for (int i = 0; i < c->n_lr[MBT_LON]; i++) c->lr[MBT_LON][i].Qfactor = c->lr[MBT_LON][i].Qfactor / c->Nscan;
    } else {
      c->scan_step = (1. - c->scan_start) / (c->Nscan - 1);
      for (int i = 0; i < c->n_lr[MBT_LON]; i++) c->lr[MBT_LON][i].Qfactor = c->lr[MBT_LON][i].Qfactor * c->scan_start;
    }
  }
  double b = c->beffL[1];
  if (c->beffH[1] > b && c->TrackPlane[1] == 1) b = c->beffH[1];
  else if (c->beffV[1] > b && c->TrackPlane[2] == 1) b = c->beffV[1];
  if (c->EnableRW_short || c->EnableRW_short_LON) {
    c->s0 = pow((2 * b * b * c->rhorw / MBT_Z0), 1.0 / 3.0);
    double dT = 2.0 * c->Nsigma / ((double)c->Ncell - 1);
    double rel = dT * c->pos_sgm[0] * MBT_C_LIGHT;
    c->mult = rel / c->s0;
    if (c->mult < 20.) {
      fprintf(stderr, "Selffield Model has too small binwidth for RW treatment!\n");
      mbt_set_error("Selffield Model has too small binwidth for RW treatment");
      return 0;
    }
  }
  if (c->EnableRW_long && c->Nmlt == 0) {
    double tmp = c->Lc * sqrt(c->rhorw) / pow(b, 3);
    c->Nmlt = 1 + ((int)500e3 / tmp);
    c->NmultiT = c->Nmlt + 2;
    if (c->Nmlt == 0) {
      fprintf(stderr, "Given RW is to shortrange, RW_long will be switched off!\n");
      c->EnableRW_long = 0;
    }
  }
  return 1;
}

int mbt_config_read(const char *conf_path, const char *work_path, mbt_config *c)
{
  memset(c, 0, sizeof *c);
  snprintf(c->input_filename, sizeof c->input_filename, "%s", conf_path);
  snprintf(c->work_path, sizeof c->work_path, "%s", work_path ? work_path : "work/");
  const char *dot = strrchr(conf_path, '.');
  if (!dot) {
    fprintf(stderr, "ERROR: Couldn't get input file extension.\n");
    mbt_set_error("input file has no extension");
    return MBT_EINVAL;
  }
  const int inp = strcmp(dot + 1, "inp") == 0;
  if (strcmp(dot + 1, "conf") != 0 && !inp) {
    fprintf(stderr, "ERROR: Unknown input file extension.\n");
    mbt_set_error("unknown input file extension %s (.conf or .inp)", dot + 1);
    return MBT_EINVAL;
  }
  int ok;
  if (inp) {
    FILE *fp = fopen(conf_path, "r");
    if (!fp) {
      fprintf(stderr, "ERROR: Cannot open file %s for reading.\n", conf_path);
      mbt_set_error("cannot open %s", conf_path);
      return MBT_EIO;
    }
    ok = read_inp(fp, c);
    fclose(fp);
  } else {
    conf_text T;
    if (conf_load(conf_path, &T) != 0) {
      fprintf(stderr, "ERROR: Cannot open file %s for reading.\n", conf_path);
      mbt_set_error("cannot open %s", conf_path);
      return MBT_EIO;
    }
    ok = read_conf(&T, c);
    conf_free(&T);
  }
  if (!ok) {
    if (!g_err[0]) mbt_set_error("read_input failed for %s", conf_path);
    return MBT_EINVAL;
  }
  if (!setup_ring(c)) return MBT_EINVAL;
  setup_macrop(c);
  if (!setup_filling(c)) {
    if (!g_err[0]) mbt_set_error("e_beam_setup failed");
    return MBT_EINVAL;
  }
  if (!setup_tracking(c)) return MBT_EINVAL;
  setup_currents(c);
  if (c->Ncell > 0) {
    c->dT = 2.0 * c->Nsigma / ((double)c->Ncell - 1);
    c->sigma_tau = c->pos_sgm[0];
  }
  return MBT_OK;
}

void mbt_config_free(mbt_config *c)
{
  free(c->nfFill);
  free(c->Ib_frac);
  free(c->Ibunch);
  free(c->bunch_Ib);
  free(c->bucket);
  free(c->bunch_out);
  c->nfFill = NULL;
  c->Ib_frac = c->Ibunch = c->bunch_Ib = NULL;
  c->bucket = c->bunch_out = NULL;
}

int mbt_config_dump(const mbt_config *c, const char *path)
{
  FILE *fp = fopen(path, "w");
  if (!fp) return MBT_EIO;
#define DD(n) fprintf(fp, #n " %.17g\n", (double)(c->n))
  DD(Nharm); DD(h); DD(T0); DD(w0); DD(wrf); DD(R); DD(Gamma); DD(Gamma2); DD(Gamma3);
  DD(U0); DD(Urad); DD(De); DD(taue); DD(tauC); DD(aexpe); DD(Je); DD(rho0);
  DD(q); DD(Fq); DD(epsMax); DD(Vrfp); DD(wso); DD(wso2); DD(fso); DD(Qso); DD(wgziH); DD(wgziV);
  DD(sn0); DD(phai0); DD(Vrf0); DD(fc1); DD(fc12); DD(Jx); DD(Jz); DD(taux); DD(tauz);
  DD(gamma1[1]); DD(gamma1[2]); DD(beta1[1]); DD(beta1[2]); DD(alpha1[1]); DD(alpha1[2]);
  DD(dispH1); DD(disppH1); DD(Gzix); DD(Gziz); DD(QH0); DD(QV0);
  DD(CHH); DD(CHV); DD(CVH); DD(CVV); DD(AmpDependdQH); DD(AmpDependdQV);
  DD(m_aHC); DD(phi_n); DD(HC_k); DD(Nbumax);
  DD(rhorw); DD(beffL[0]); DD(beffL[1]); DD(beffH[0]); DD(beffH[1]); DD(beffV[0]); DD(beffV[1]);
  DD(NrevTot); DD(NrevMon); DD(NrevScan); DD(Nscan); DD(scan); DD(scan_start); DD(scan_step);
  DD(Nmlt); DD(NmultiT); DD(triggerRW); DD(EnableRW_long); DD(EnableQuantum); DD(EnableIdealHC);
  DD(TrackPlane[0]); DD(TrackPlane[1]); DD(TrackPlane[2]);
  DD(Np); DD(iseed[0]); DD(iseed[1]); DD(iseed[2]);
  DD(pos_offset[0]); DD(pos_offset[1]); DD(pos_offset[2]);
  DD(slope_offset[0]); DD(slope_offset[1]); DD(slope_offset[2]);
  DD(pos_sgm[0]); DD(pos_sgm[1]); DD(pos_sgm[2]);
  DD(slope_sgm[0]); DD(slope_sgm[1]); DD(slope_sgm[2]);
  DD(correlate[0]); DD(correlate[1]); DD(correlate[2]);
  DD(Ncell); DD(Nsigma); DD(dT); DD(RsisL); DD(aindL);
  DD(Nbunch); DD(Nbunch_out); DD(has_rf_feedback); DD(rffb_vrf_design); DD(rffb_phi0_design);
  DD(n_res); DD(n_lr[0]); DD(n_lr[1]); DD(n_lr[2]); DD(n_ahc);
#undef DD
  for (int p = 0; p < 3; p++)
    for (int i = 0; i < c->n_lr[p]; i++)
      fprintf(fp, "lr%d_%d %.17g %.17g %.17g %u %u\n", p, i, c->lr[p][i].wr, c->lr[p][i].Rs,
              c->lr[p][i].Qfactor, c->lr[p][i].Nturn, c->lr[p][i].Nbu);
  for (int i = 0; i < c->Nbunch_out; i++) fprintf(fp, "bunch_out %d\n", c->bunch_out[i]);
  for (unsigned k = 0; k < c->Nharm; k++)
    if (c->nfFill[k]) fprintf(fp, "Ibunch_%u %.17g\n", k, c->Ibunch[k]);
  fclose(fp);
  return MBT_OK;
}

/* ------------------------------------------------- accessors for bindings */

#include <stddef.h>

typedef struct field_desc {
  const char *name;
  size_t off;
  char kind; /* d double, i int, l long, u unsigned */
} field_desc;

#define FD(n, k) {#n, offsetof(mbt_config, n), k}
static const field_desc g_fields[] = {
    FD(Nbunch, 'i'), FD(Np, 'i'), FD(Ncell, 'i'), FD(Nharm, 'u'), FD(NrevTot, 'l'), FD(NrevMon, 'l'),
    FD(NrevScan, 'l'), FD(Nscan, 'l'), FD(scan, 'i'), FD(Nmlt, 'i'), FD(EnableQuantum, 'i'),
    FD(EnableRW_long, 'i'), FD(EnableIdealHC, 'i'), FD(EnableAmpinv_out, 'i'), FD(bunch_model, 'i'),
    FD(Nbunch_out, 'i'), FD(has_rf_feedback, 'i'), FD(n_res, 'i'), FD(NrevPotentialsOut, 'l'),
    FD(T0, 'd'), FD(h, 'd'), FD(E0, 'd'), FD(sigma_tau, 'd'), FD(Nsigma, 'd'), FD(dT, 'd'), FD(taue, 'd'),
    FD(taux, 'd'), FD(tauz, 'd'), FD(De, 'd'), FD(Urad, 'd'), FD(phai0, 'd'), FD(Vrf0, 'd'), FD(fNp, 'd'),
    FD(QH0, 'd'), FD(QV0, 'd'), FD(Qso, 'd'), FD(ac, 'd'), FD(Iring, 'd'), FD(Lc, 'd'), FD(rhorw, 'd'),
    FD(scan_start, 'd'), FD(scan_step, 'd'), FD(Gzix, 'd'), FD(Gziz, 'd'), FD(wrf, 'd'), FD(Nslc, 'i'),
};
#undef FD

static const field_desc *find_field(const char *name)
{
  for (size_t i = 0; i < sizeof g_fields / sizeof g_fields[0]; i++)
    if (strcmp(g_fields[i].name, name) == 0) return &g_fields[i];
  if (strcmp(name, "TrackPlane_HOR") == 0) {
    static field_desc f = {"TrackPlane_HOR", offsetof(mbt_config, TrackPlane) + sizeof(int), 'i'};
    return &f;
  }
  if (strcmp(name, "TrackPlane_VER") == 0) {
    static field_desc f = {"TrackPlane_VER", offsetof(mbt_config, TrackPlane) + 2 * sizeof(int), 'i'};
    return &f;
  }
  return NULL;
}

mbt_config *mbt_config_alloc(void) { return (mbt_config *)calloc(1, sizeof(mbt_config)); }

void mbt_config_release(mbt_config *c)
{
  if (!c) return;
  mbt_config_free(c);
  free(c);
}

int mbt_config_get(const mbt_config *c, const char *name, double *v)
{
  const field_desc *f = find_field(name);
  if (!f) {
    mbt_set_error("unknown configuration field %s", name);
    return MBT_EINVAL;
  }
  const char *p = (const char *)c + f->off;
  switch (f->kind) {
  case 'd': *v = *(const double *)p; break;
  case 'i': *v = *(const int *)p; break;
  case 'u': *v = *(const unsigned *)p; break;
  default: *v = (double)*(const long *)p; break;
  }
  return MBT_OK;
}

int mbt_config_set(mbt_config *c, const char *name, double v)
{
  const field_desc *f = find_field(name);
  if (!f) {
    mbt_set_error("unknown configuration field %s", name);
    return MBT_EINVAL;
  }
  char *p = (char *)c + f->off;
  switch (f->kind) {
  case 'd': *(double *)p = v; break;
  case 'i': *(int *)p = (int)v; break;
  case 'u': *(unsigned *)p = (unsigned)v; break;
  default: *(long *)p = (long)v; break;
  }
  return MBT_OK;
}

int mbt_config_bucket(const mbt_config *c, int k) { return (k >= 0 && k < c->Nbunch) ? c->bucket[k] : -1; }

/* ------------------------------------------ single-key lookups (src/input.c:972-1100) */

static int lookup(const char *path, const char *sec, const char *key, int kind, void *out, size_t outlen)
{
  conf_text T;
  if (conf_load(path, &T) != 0) {
    mbt_set_error("cannot open %s", path);
    return MBT_EIO;
  }
  T.print_errors = 0;
  int ok = conf_get(&T, sec, key, kind, out, outlen);
  conf_free(&T);
  return ok ? 1 : 0;
}

int mbt_conf_get_double(const char *path, const char *section, const char *key, double *v)
{
  return lookup(path, section, key, K_DOUBLE, v, 0);
}

int mbt_conf_get_int(const char *path, const char *section, const char *key, int *v)
{
  return lookup(path, section, key, K_INT, v, 0);
}

int mbt_conf_get_str(const char *path, const char *section, const char *key, char *v, int len)
{
  return lookup(path, section, key, K_STR, v, (size_t)len);
}

int mbt_conf_have_section(const char *path, const char *section)
{
  conf_text T;
  if (conf_load(path, &T) != 0) return MBT_EIO;
  T.print_errors = 0;
  int ok = have_section(&T, section);
  conf_free(&T);
  return ok;
}
