/* sl_rex.c -- rex v2 files for batched runs, and the datum override of the
 * env and rex trailers.
 *
 * The frames' bit streams come from the device (slg_batch_rex_frames); this
 * file lays out the file around them as the reference's rex writer does:
 *   header        RexOpen + RexSaveHeader   sorc/slosh.fd/rex.c:150-180, 264-379
 *   frame         RexSaveStep               rex.c:387-447 (track, date, offset)
 *   envelope      RexSaveEnv + Data_WriteTrk rex.c:505-545, 57-230
 *   header patch  RexSetMinMaxHt            rex.c:232-256
 * All numbers are little-endian, as FWRITE_LIT writes them.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/slosh_hip.h"

struct slg_rex {
   FILE *fp;
   int32_t off;        /* bytes written so far (rex->Offset) */
   int32_t trk_off;    /* where the last frame's next-frame offset goes */
   int32_t minft_off;  /* header position of the minimum-height field */
};

static int wr(slg_rex *r, const void *p, size_t n)
{
   return fwrite(p, 1, n, r->fp) == n ? 0 : -1;
}

static int wr_u16(slg_rex *r, unsigned v)
{
   const uint16_t s = (uint16_t) v;
   return wr(r, &s, 2);
}

/* RexOpen (version 2) and RexSaveHeader: "rex2:" + 0xAB + the comment (double
 * quotes made single), its last 12 bytes moved behind the fixed fields; a
 * comment shorter than 12 characters is padded there with its NUL and the
 * start of "ArthurTaylor", as rex.c:308-322 leaves them in its buffer. */
int slg_rex_open(const char *path, const char *comment, const char *abbrev, int imxb, int jmxb, slg_rex **out)
{
   *out = NULL;
   if (!path || imxb < 2 || jmxb < 2 || imxb > 65536 || jmxb > 65536) return SLG_EINPUT;
   char com[201];
   size_t L = comment ? strlen(comment) : 0;
   if (L > 200) L = 200;
   for (size_t k = 0; k < L; k++) com[k] = comment[k] == '"' ? '\'' : comment[k];
   com[L] = '\0';
   /* the identifier and comment as one buffer: str_len bytes in all */
   char buf[6 + 200 + 13];
   memset(buf, 0, sizeof buf);
   memcpy(buf, "rex2:", 5);
   buf[5] = (char) 171;
   memcpy(buf + 6, com, L);
   size_t str_len = 6 + L;
   if (L < 12) {
      static const char pad[] = " ArthurTaylor";
      memcpy(buf + 6 + L, pad, 12 - L);
      str_len = 18;
      buf[6 + L] = '\0';
   }
   char abv[4] = {0, 0, 0, 0};
   if (abbrev) {
      const size_t la = strlen(abbrev);
      memcpy(abv, abbrev, la < 4 ? la : 4);
   }
   slg_rex *r = (slg_rex *) calloc(1, sizeof *r);
   if (!r) return SLG_EINPUT;
   r->fp = fopen(path, "w+b");
   if (!r->fp) {
      free(r);
      return SLG_EINPUT;
   }
   const unsigned char c0 = (unsigned char) (str_len - 12);
   const unsigned si = (unsigned) (imxb - 1), sj = (unsigned) (jmxb - 1);
   const unsigned char eof = 26;
   const int32_t zero = 0;
   int bad = wr(r, &c0, 1) || wr(r, buf, str_len - 12) || wr(r, abv, 4) || wr_u16(r, si) || wr_u16(r, sj) ||
             wr_u16(r, 0) || wr_u16(r, si) || wr_u16(r, 0) || wr_u16(r, sj);
   r->minft_off = (int32_t) (1 + str_len - 12 + 4 + 4 + 8);
   /* minft, maxft, startfrm, stopfrm, minimum frame display time */
   bad = bad || wr_u16(r, 0) || wr_u16(r, 100) || wr_u16(r, 0) || wr_u16(r, 0) || wr_u16(r, 100);
   bad = bad || wr(r, buf + str_len - 12, 12) || wr(r, &eof, 1) || wr(r, &zero, 4);
   r->off = (int32_t) (1 + str_len - 12 + 4 + 4 + 8 + 10 + 12 + 1 + 4);
   if (bad) {
      slg_rex_close(r);
      return SLG_EINPUT;
   }
   *out = r;
   return 0;
}

/* the previous frame's next-frame offset becomes the current position */
static int link_prev(slg_rex *r)
{
   if (r->trk_off) {
      if (fseek(r->fp, r->trk_off, SEEK_SET) || wr(r, &r->off, 4)) return -1;
      if (fseek(r->fp, r->off, SEEK_SET)) return -1;
   }
   return 0;
}

/* myRound (myutil.c:387-393) */
static float round_to(double v, int place)
{
   const double p = place == 4 ? 10000. : 100.;
   return (float) (floor(v * p + .5) / p);
}

/* Clock_PrintDate (clock.c:636-650) for seconds since 1970-01-01 00:00 UTC;
 * the calendar date by the days-to-civil conversion of H. Hinnant's
 * chrono-compatible date algorithms (proleptic Gregorian). */
static void clock_date(double clock, int *year, int *mon, int *day, int *hr, int *mn, double *sec)
{
   const double SEC_DAY = 86400.;
   const long tot = (long) floor(clock / SEC_DAY);
   long z = tot + 719468;
   const long era = (z >= 0 ? z : z - 146096) / 146097;
   const long doe = z - era * 146097;
   const long yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
   const long doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
   const long mp = (5 * doy + 2) / 153;
   *day = (int) (doy - (153 * mp + 2) / 5 + 1);
   *mon = (int) (mp < 10 ? mp + 3 : mp - 9);
   *year = (int) (yoe + era * 400 + (*mon <= 2));
   double s = clock - (double) tot * SEC_DAY;
   const int32_t is = (int32_t) s;
   *hr = (int) ((is % 86400) / 3600);
   *mn = (int) ((is % 3600) / 60);
   *sec = (is % 60) + (s - is);
}

/* RexSaveStep (rex.c:387-447): track values rounded as myRound does, the date
 * of the model clock, a next-frame offset (patched by the next frame), then
 * the frame's bit stream. */
int slg_rex_frame(slg_rex *r, const double track[6], double clock, const unsigned char *bits, int nbytes)
{
   if (!r || !track || nbytes < 0 || (nbytes && !bits)) return SLG_EINPUT;
   if (link_prev(r)) return SLG_EINPUT;
   const float f[6] = {round_to(track[0], 4), round_to(track[1], 4), round_to(track[2], 2),
                       round_to(track[3], 2), round_to(track[4], 2), round_to(track[5], 2)};
   int year, mon, day, hr, mn;
   double sec;
   clock_date(clock, &year, &mon, &day, &hr, &mn, &sec);
   const unsigned char d[5] = {(unsigned char) hr, (unsigned char) mn, (unsigned char) sec, (unsigned char) mon,
                               (unsigned char) day};
   const int16_t y = (int16_t) year;
   const int32_t zero = 0;
   if (wr(r, f, sizeof f) || wr(r, d, 5) || wr(r, &y, 2)) return SLG_EINPUT;
   r->off += 24 + 7;
   r->trk_off = r->off;
   if (wr(r, &zero, 4)) return SLG_EINPUT;
   r->off += 4;
   if (nbytes && wr(r, bits, (size_t) nbytes)) return SLG_EINPUT;
   r->off += nbytes;
   return 0;
}

/* InitWater_Anomaly, InitWater_TideMode, InitWater_New and the surge-only
 * branch of InitWater_TideModeOverride (slosh2.c:345-482): the datums of a
 * surge-only run, in float as the reference holds them. */
void sl_datum_override(float *ht1, float *ht2)
{
   sl_datum_override_tide(ht1, ht2, 0);
}

void sl_datum_override_tide(float *ht1, float *ht2, int f_tide)
{
   const float h = *ht1;
   float anom;
   if ((double) h == 99.9 || (double) h == 999.9) anom = 99.9f;
   else anom = (float) (fmod((double) (h + 50.f), 100.) - 50.);
   int mode;
   const int tenth = (int) floor((double) (h * 10.f) + .5);
   if (tenth == 999) mode = 10;
   else if (tenth == 9999) mode = 11;
   else {
      static const int by_adjust[9] = {0, 99, 9, -1, 1, -2, 2, -3, 3};
      const int adjust = (int) ((h + 50.f) / 100.f);
      mode = (adjust >= 0 && adjust <= 8 && adjust != 1) ? by_adjust[adjust] : 0;
   }
   if (f_tide == 0) {
      if (mode == 1 || mode == 2 || mode == 3 || mode == -1 || mode == -2 || mode == -3) mode = 9;
   } else {   /* the gridded tide V2 modes (slosh2.c:444-452) */
      if (mode == 10 || mode == 11) {
         anom = 0.f;
         *ht2 = 0.f;
      }
      mode = 2;
   }
   int a;
   switch (mode) {   /* InitWater_New, slosh2.c:345-381 */
      case 10: *ht1 = 99.9f; return;
      case 11: *ht1 = 999.9f; return;
      case 9: a = 2; break;
      case 2: a = 6; break;
      default: a = 0; break;   /* mode 0 */
   }
   *ht1 = (float) (a * 100) + anom;
}

/* Data_WriteTrk (rex.c:57-230): "envl", then the track file's first 100 track
 * rows as six floats each, begin / end / landfall hour as bytes and the two
 * datums (after the surge-only override), read with the reference's fixed
 * line positions and fields.  Lines are read into one 200-byte buffer as
 * fgets leaves it; at the end of the file the last line is read again. */
static int write_trk(slg_rex *r, const char *trk)
{
   FILE *fp = fopen(trk, "rt");
   if (!fp) return -1;
   char b[200];
   memset(b, 0, sizeof b);
   float lat = 0.f, lon = 0.f, spd = 0.f, dir = 0.f, delp = 0.f, rmax = 0.f, ht1, ht2;
   int i = 0, j = 0, k = 0, bad = wr(r, "envl", 4);
#define GET() ((void) !fgets(b, 200, fp))
   GET();
   rewind(fp);
   if (strncmp(b, "FileVersion=", 12) == 0) {
      for (int n = 0; n < 16; n++) GET();
      for (int n = 0; n < 100; n++) {
         GET();
         b[0] = ' ';
         (void) sscanf(b, "%d %f %f %f %f", &j, &lat, &lon, &delp, &rmax);
         spd = 0.f;
         dir = 0.f;
         const float v[6] = {lat, lon, spd, dir, delp, rmax};
         bad = bad || wr(r, v, sizeof v);
      }
      rewind(fp);
      for (int n = 0; n < 6; n++) GET();
      GET();
      k = atoi(b + 13);
      GET();
      i = atoi(b + 10);
      GET();
      j = atoi(b + 8);
      const char c[3] = {(char) i, (char) j, (char) k};
      bad = bad || wr(r, c, 3);
      GET();
      GET();
      ht1 = (float) atof(b + 11);
      GET();
      ht2 = (float) atof(b + 10);
   } else {
      GET();
      GET();
      for (int n = 0; n < 100; n++) {
         GET();
         memset(b, ' ', 18);
         (void) sscanf(b, "%d %f %f %f %f %f %f %d", &j, &lat, &lon, &spd, &dir, &delp, &rmax, &k);
         const float v[6] = {lat, lon, spd, dir, delp, rmax};
         bad = bad || wr(r, v, sizeof v);
      }
      GET();
      char t;
      t = b[3], b[3] = '\0', i = atoi(b), b[3] = t;
      t = b[6], b[6] = '\0', j = atoi(b + 3), b[6] = t;
      t = b[9], b[9] = '\0', k = atoi(b + 6), b[9] = t;
      const char c[3] = {(char) i, (char) j, (char) k};
      bad = bad || wr(r, c, 3);
      GET();
      GET();
      if (b[10] == 'x' || b[10] == 'X') {   /* Okeechobee datums */
         t = b[10], b[10] = '\0', ht1 = (float) atof(b + 5), b[10] = t;
         t = b[16], b[16] = '\0', ht2 = (float) atof(b + 11), b[16] = t;
      } else {
         t = b[5], b[5] = '\0', ht1 = (float) atof(b), b[5] = t;
         t = b[10], b[10] = '\0', ht2 = (float) atof(b + 5), b[10] = t;
      }
   }
#undef GET
   fclose(fp);
   sl_datum_override(&ht1, &ht2);
   bad = bad || wr(r, &ht1, 4) || wr(r, &ht2, 4);
   if (bad) return -1;
   r->off += 4 + 600 * 4 + 3 + 2 * 4;
   return 0;
}

/* RexSaveEnv (rex.c:505-650): the track file's record and the envelope
 * frame's bit stream; then RexSetMinMaxHt, which writes only the minimum
 * height (0) into the header (rex.c:232-256). */
int slg_rex_env(slg_rex *r, const char *trk_path, const unsigned char *bits, int nbytes)
{
   if (!r || !trk_path || nbytes < 0 || (nbytes && !bits)) return SLG_EINPUT;
   if (link_prev(r)) return SLG_EINPUT;
   if (write_trk(r, trk_path)) return SLG_EINPUT;
   if (nbytes && wr(r, bits, (size_t) nbytes)) return SLG_EINPUT;
   r->off += nbytes;
   if (fseek(r->fp, r->minft_off, SEEK_SET) || wr_u16(r, 0)) return SLG_EINPUT;
   return 0;
}

void slg_rex_close(slg_rex *r)
{
   if (!r) return;
   if (r->fp) fclose(r->fp);
   free(r);
}

/* Point series in rexout's style 0 (sorc/rexout.cd/rexout.c:631-639 header,
 * 736-742 rows): the values a rex file's frames hold at the listed squares,
 * written without the frames.  The basin field is the rex header's 4-byte
 * abbreviation as rexout prints it (a leading blank dropped, a trailing blank
 * cut, rexout.c:534-543); each row gives the frame's date from the model clock
 * (as slg_rex_frame stores it) and value / 10 in float with one decimal. */
struct slg_pnt {
   FILE *fp;
   int npts;
};

int slg_pnt_open(const char *path, const char *abbrev, const char *storm, const char *rexfile, int npts,
                 const int *xy, slg_pnt **out)
{
   *out = NULL;
   if (!path || npts < 0 || npts > SLG_MAX_POINTS || (npts && !xy)) return SLG_EINPUT;
   char abv[5] = {0, 0, 0, 0, 0};
   if (abbrev) {
      const size_t la = strlen(abbrev);
      memcpy(abv, abbrev, la < 4 ? la : 4);
   }
   if (abv[0] == ' ') {
      abv[0] = abv[1];
      abv[1] = abv[2];
      abv[2] = abv[3];
      abv[3] = ' ';
   }
   if (abv[3] == ' ') abv[3] = '\0';
   slg_pnt *p = (slg_pnt *) calloc(1, sizeof *p);
   if (!p) return SLG_EINPUT;
   p->fp = fopen(path, "w");
   p->npts = npts;
   if (!p->fp) {
      free(p);
      return SLG_EINPUT;
   }
   fprintf(p->fp, "Basin=%s\nStorm=%s\nRexFile=%s\nDate, Time", abv, storm ? storm : "", rexfile ? rexfile : "");
   for (int k = 0; k < npts; k++) fprintf(p->fp, ", (%d %d)", xy[2 * k], xy[2 * k + 1]);
   fprintf(p->fp, "\n");
   *out = p;
   return ferror(p->fp) ? SLG_EINPUT : 0;
}

int slg_pnt_row(slg_pnt *p, double clock, const short *vals)
{
   if (!p || (p->npts && !vals)) return SLG_EINPUT;
   int year, mon, day, hr, mn;
   double sec;
   clock_date(clock, &year, &mon, &day, &hr, &mn, &sec);
   fprintf(p->fp, "%02d/%02d/%d, %02d:%02d:%02d", mon, day, year, hr, mn, (int) (unsigned char) sec);
   for (int k = 0; k < p->npts; k++) {
      const float g = (float) (vals[k] / 10.0);
      fprintf(p->fp, ", %.1f", g);
   }
   fprintf(p->fp, "\n");
   return ferror(p->fp) ? SLG_EINPUT : 0;
}

void slg_pnt_close(slg_pnt *p)
{
   if (!p) return;
   if (p->fp) fclose(p->fp);
   free(p);
}
