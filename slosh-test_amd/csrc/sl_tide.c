/* sl_tide.c -- the gridded astronomical tide of a basin, read as the
 * reference's driver reads it for its -f_tide V2 modes.
 *
 * Restates, in double precision:
 *   InitTideGrid / LoadYearlyConstit   sorc/slosh.fd/tideutil.c:148-352
 *     (constituent speeds; the year's nodal factors and equilibrium
 *     arguments from ft03, reordered to "Tides and Currents" order)
 *   ReadHarmonicHeader / Body          tideutil.c:470-645 (the basin's
 *     harmonic constants, one record of 37 amplitudes and phases per square)
 *   ReadAdjDatum                       tideutil.c:647-734 (MTL -> NAVD88)
 *   Clock_ScanDate                     clock.c:1219-1285, seconds from 1970
 *                                      to a date's 00:00
 * The constituent tables must be the reference's (they are data); the
 * readers follow its field order and scale factors so the values are the same
 * doubles; the sine and cosine of a phase come from one sincos call, as the
 * reference's build takes them (glibc's sincos and its separate sin and cos
 * differ in the last bit for about 0.1 % of arguments).  The harmonic
 * constants are kept as read (integers) and combined
 * with a year's constituents on demand, so one basin serves runs of several
 * years.
 */
#define _GNU_SOURCE   /* sincos */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sl_host.h"

#define TD_PI_180 0.0174532925199432957692

/* speeds in degrees per hour, "Tides and Currents" order (tideutil.c:290-305) */
static const double td_speed[SL_NUMT] = {
   28.9841042, 30.0000000, 28.4397295, 15.0410686, 57.9682084, 13.9430356, 86.9523127, 44.0251729,
   60.0000000, 57.4238337, 28.5125831, 90.0000000, 27.9682084, 27.8953548, 16.1391017, 29.4556253,
   15.0000000, 14.4966939, 15.5854433, 0.5443747,  0.0821373,  0.0410686,  1.0158958,  1.0980331,
   13.4715145, 13.3986609, 29.9589333, 30.0410667, 12.8542862, 14.9589314, 31.0158958, 43.4761563,
   29.5284789, 42.9271398, 30.0821373, 115.9364166, 58.9841042};
/* where each of those sits in an ft03 year's record (tideutil.c:162-173) */
static const int td_ft03[SL_NUMT] = {0,  1,  2,  3,  11, 4,  12, 29, 23, 30, 9,  31, 13, 8,  25, 18, 17, 19, 24,
                                     32, 16, 14, 33, 34, 26, 10, 15, 27, 28, 5,  35, 22, 7,  36, 6,  20, 21};

/* days from 1970-01-01 to y-m-d (proleptic Gregorian) */
static long days_from_civil(long y, int m, int d)
{
   long era, yoe, doy, doe;
   y -= m <= 2;
   era = (y >= 0 ? y : y - 399) / 400;
   yoe = y - era * 400;
   doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
   doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
   return era * 146097 + doe - 719468;
}

double sl_clock_scan_date(int year, int month, int day)
{
   return (double) days_from_civil(year, month, day) * 86400.;
}

double sl_tide_speed(int k)
{
   return td_speed[k] * TD_PI_180;
}

/* one 8-field group of an ft03 line: 4 columns of nodal factor x 1000, 4 of
 * equilibrium argument x 10, after an 8-column year field (tideutil.c:82-108) */
static void ft03_fields(char *p, int beg, int end, double *xode, double *vpu)
{
   char c;
   int i;
   p += 8;
   for (i = beg; i <= end; i++) {
      c = p[4];
      p[4] = '\0';
      xode[i - 1] = atoi(p) / 1000.;
      p[4] = c;
      c = p[8];
      p[8] = '\0';
      vpu[i - 1] = atoi(p + 4) / 10.;
      p[8] = c;
      p += 8;
   }
}

/* the year's nodal factors and equilibrium arguments, in "Tides and
 * Currents" order (LoadYearlyConstit, tideutil.c:180-286) */
static int load_year(const char *ft03, int year, double *xo, double *vp, char err[256])
{
   char b[1000];
   int iyr1, i;
   double xode[SL_NUMT], vpu[SL_NUMT];
   static const int groups[5][2] = {{1, 8}, {9, 16}, {17, 24}, {25, 32}, {33, SL_NUMT}};
   FILE *fp = fopen(ft03, "rt");
   if (!fp) {
      snprintf(err, 256, "cannot open the constituent file %s", ft03);
      return -1;
   }
   if (!fgets(b, sizeof b, fp) || sscanf(b, "%d", &iyr1) != 1 || year < iyr1) {
      fclose(fp);
      snprintf(err, 256, "year %d is not in the constituent file", year);
      return -1;
   }
   /* five lines per year; the first year's first line is already read */
   for (i = 1; i < (year - iyr1) * 5; i++)
      if (!fgets(b, sizeof b, fp)) break;
   if (i != 1 && !fgets(b, sizeof b, fp)) {
      fclose(fp);
      snprintf(err, 256, "year %d is not in the constituent file", year);
      return -1;
   }
   {
      char y[5];
      memcpy(y, b, 4);
      y[4] = 0;
      if (atoi(y) != year) {
         fclose(fp);
         snprintf(err, 256, "the constituent file has no record for %d", year);
         return -1;
      }
   }
   for (i = 0; i < 5; i++) {
      if (i > 0 && !fgets(b, sizeof b, fp)) {
         fclose(fp);
         snprintf(err, 256, "the constituent file ends inside %d's record", year);
         return -1;
      }
      ft03_fields(b, groups[i][0], groups[i][1], xode, vpu);
   }
   fclose(fp);
   for (i = 0; i < SL_NUMT; i++) {
      xo[i] = xode[td_ft03[i]];
      vp[i] = vpu[td_ft03[i]];
   }
   return 0;
}

static int rd(FILE *fp, void *p, size_t n)
{
   return fread(p, 1, n, fp) == n ? 0 : -1;
}

int sl_tide_open(sl_tide *t, const char *bhc, const char *adj, int imxb, int jmxb, char err[256])
{
   char id[12], name[5];
   short ver, hi, hj;
   int i, j;
   FILE *fp;
   memset(t, 0, sizeof *t);
   err[0] = 0;
   fp = fopen(bhc, "rb");
   if (!fp) {
      snprintf(err, 256, "cannot open the harmonic-constant file %s", bhc);
      return -1;
   }
   if (rd(fp, id, 12) || memcmp(id, "HarmConstit", 12) || rd(fp, &ver, 2) || ver != 1 ||
       rd(fp, t->abbrev, 5) || rd(fp, &hi, 2) || rd(fp, &hj, 2)) {
      fclose(fp);
      snprintf(err, 256, "%s is not a version 1 HarmConstit file", bhc);
      return -1;
   }
   t->abbrev[5] = 0;
   for (i = 0; i < SL_NUMT; i++)
      if (rd(fp, name, 5)) break;
   if (hi + 1 != imxb || hj + 1 != jmxb) {
      fclose(fp);
      snprintf(err, 256, "harmonic constants for %d x %d squares, the basin has %d x %d", hi, hj, imxb - 1,
               jmxb - 1);
      return -1;
   }
   t->ni = hi;
   t->nj = hj;
   t->ncell = hi * hj;
   t->amp = (int32_t *) calloc((size_t) t->ncell * SL_NUMT, sizeof(int32_t));
   t->phs = (int32_t *) calloc((size_t) t->ncell * SL_NUMT, sizeof(int32_t));
   t->mtl = (double *) calloc((size_t) t->ncell, sizeof(double));
   t->lon = (double *) calloc((size_t) t->ncell, sizeof(double));
   t->lat = (double *) calloc((size_t) t->ncell, sizeof(double));
   if (!t->amp || !t->phs || !t->mtl || !t->lon || !t->lat) {
      fclose(fp);
      sl_tide_free(t);
      snprintf(err, 256, "out of host memory");
      return -1;
   }
   for (i = 0; i < t->ncell; i++) {
      int32_t v;
      if (rd(fp, &v, 4)) break;
      t->lon[i] = v / 1000000.;
      if (rd(fp, &v, 4)) break;
      t->lat[i] = v / 1000000.;
      for (j = 0; j < SL_NUMT; j++)
         if (rd(fp, &t->amp[(size_t) i * SL_NUMT + j], 4) || rd(fp, &t->phs[(size_t) i * SL_NUMT + j], 4)) break;
      if (j < SL_NUMT) break;
   }
   fclose(fp);
   if (i < t->ncell) {
      sl_tide_free(t);
      snprintf(err, 256, "%s ends before its %d squares", bhc, hi * hj);
      return -1;
   }
   /* the datum adjustment, column 5 of each line after the header, one line
    * per square in file order; 0 where there is no file (slosh2.c:1826-1836) */
   fp = adj ? fopen(adj, "rt") : NULL;
   if (fp) {
      char b[1000];
      int line = 0;
      while (fgets(b, sizeof b, fp)) {
         char *c, *f[6];
         int n = 0;
         if (++line == 1) continue;
         if (line - 2 >= t->ncell) break;
         for (c = b; n < 6;) {
            f[n++] = c;
            c = strchr(c, ',');
            if (!c) break;
            *c++ = 0;
         }
         if (n >= 5) t->mtl[line - 2] = atof(f[4]);
      }
      fclose(fp);
   }
   return 0;
}

int sl_tide_year(const sl_tide *t, const char *ft03, int year, double *xamp, double *cosv, double *sinv,
                 char err[256])
{
   double xode[SL_NUMT], vpu[SL_NUMT];
   const size_t n = (size_t) t->ncell;
   size_t c;
   int k;
   if (load_year(ft03, year, xode, vpu, err)) return -1;
   for (c = 0; c < n; c++)
      for (k = 0; k < SL_NUMT; k++) {
         const double ph = (vpu[k] - t->phs[c * SL_NUMT + k] / 1000000.) * TD_PI_180;
         xamp[(size_t) k * n + c] = xode[k] * t->amp[c * SL_NUMT + k] / 1000000.;
         sincos(ph, &sinv[(size_t) k * n + c], &cosv[(size_t) k * n + c]);
      }
   return 0;
}

void sl_tide_free(sl_tide *t)
{
   free(t->amp);
   free(t->phs);
   free(t->mtl);
   free(t->lon);
   free(t->lat);
   memset(t, 0, sizeof *t);
}
