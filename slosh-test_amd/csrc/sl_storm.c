/* sl_storm.c -- track ingest and the per-storm state machine.
 *
 * Restates, in double precision, the reference's storm side:
 *   INPUTN2 trk reader (2013 format)  sorc/slosh.fd/inputn2.f:110-605
 *   spline / splint                   sorc/slosh.fd/spline.f, splint.f
 *   STERON  track -> grid (x, y)      sorc/slosh.fd/steron.f:36-72
 *   INTDLT / INTVAL time setup        sorc/slosh.fd/intdlt.f:40-62, intval.f:52-104
 *   SETCMP  first-hour storm          sorc/slosh.fd/setcmp.f:77-180
 *   INTWND  max-wind first guess      sorc/slosh.fd/intwnd.f:43-82
 *   FORCE1  radial wind/pressure      sorc/slosh.fd/force1.f:55-183
 *   TTIMER  step count + ramp         sorc/slosh.fd/ttimer.f:33-42
 *   STMVAL  hourly/step storm advance sorc/slosh.fd/stmval.f:95-236
 *   CNSDLT  dt-dependent constants    sorc/slosh.fd/cnsdlt.f:27-44
 *   XY2LLX  storm centre lat/lon      sorc/slosh.fd/xy2llx.f
 *   INTLHT  initial water             sorc/slosh.fd/intlht.f:53-185
 * Expressions keep the reference's operand order so the double results match
 * the reference's default FP64 build exactly.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "sl_host.h"

#define PIX(b, i, j) ((size_t) ((j) - 1) * (size_t) (b)->imxb + (size_t) ((i) - 1))
#define RADF 1.74532925199433E-2
#define PI_F 3.14159265358979323846

/* ---- spline / splint (spline.f, splint.f) ---- */
static void spline(const double *x, const double *y, int n, double yp1, double ypn, double *y2)
{
   double u[SL_NT + 2], sig, p, qn, un;
   int i, k;
   if (yp1 > .99e30) {
      y2[1] = 0.;
      u[1] = 0.;
   } else {
      y2[1] = -0.5;
      u[1] = (3. / (x[2] - x[1])) * ((y[2] - y[1]) / (x[2] - x[1]) - yp1);
   }
   for (i = 2; i <= n - 1; i++) {
      sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
      p = sig * y2[i - 1] + 2.;
      y2[i] = (sig - 1.) / p;
      u[i] = (6. * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1])
                    / (x[i] - x[i - 1])) / (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
   }
   if (ypn > .99e30) {
      qn = 0.;
      un = 0.;
   } else {
      qn = 0.5;
      un = (3. / (x[n] - x[n - 1])) * (ypn - (y[n] - y[n - 1]) / (x[n] - x[n - 1]));
   }
   y2[n] = (un - qn * u[n - 1]) / (qn * y2[n - 1] + 1.);
   for (k = n - 1; k >= 1; k--) y2[k] = y2[k] * y2[k + 1] + u[k];
}

static double splint(const double *xa, const double *ya, const double *y2a, int n, double x)
{
   int klo = 1, khi = n, k;
   double h, a, b;
   while (khi - klo > 1) {
      k = (khi + klo) / 2;
      if (xa[k] > x) khi = k;
      else klo = k;
   }
   h = xa[khi] - xa[klo];
   a = (xa[khi] - x) / h;
   b = (x - xa[klo]) / h;
   return a * ya[klo] + b * ya[khi] +
          ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.;
}

static int starts(const char *s, const char *p)
{
   return strncmp(s, p, strlen(p)) == 0;
}

/* One track row (inputn2.f:560-605): bounds checked as INPUTN2 checks them;
 * the hour counted from the first row's hour 1, the longitude negated (west
 * positive inside the model, inputn2.f:439).  dp: the pressure deficit after
 * the PressureType conversion. */
int sl_trk_row(sl_trk *t, double hour, double lat, double lon, double dp, double rmw, char err[256])
{
   int n;
   if (hour < 0 || hour > 999 || lat < 0. || lat > 90. || lon < -180. || lon > 180. || rmw < 1. ||
       rmw > 500. || dp > 213. || dp < 0.1) {
      snprintf(err, 256, "track row out of bounds (hour %d)", (int) hour);
      return -1;
   }
   if (t->n >= SL_NT) return 1;   /* rows beyond NT_ are not read */
   n = ++t->n;
   if (n == 1) t->adj = 1 - (int) hour;
   t->hrs[n] = (int) hour + t->adj;
   t->fla[n] = lat;
   t->flo[n] = lon * (-1.0);
   t->fdp[n] = dp;
   t->frm[n] = rmw;
   if (n > 1 && t->hrs[n] <= t->hrs[n - 1]) {
      snprintf(err, 256, "track hours not increasing");
      return -1;
   }
   return 0;
}

void sl_trk_init(sl_trk *t)
{
   memset(t, 0, sizeof *t);
   t->beg = t->end = t->land = t->dhr = -999;
   t->ocean = t->lake = -999.9;
   t->windf = t->rmwf = 1.0;
}

/* A Fortran Fw.d input field: columns [c0, c0 + w) of line (short lines read
 * as blanks), blanks ignored, an all-blank field 0, and d implied decimals
 * when the digits carry no decimal point. */
static double fort_f(const char *line, int c0, int w, int d)
{
   char b[64];
   int k, n = 0, dot = 0;
   const int len = (int) strlen(line);
   for (k = c0; k < c0 + w && k < len && n < 62; k++) {
      const char c = line[k];
      if (c == ' ' || c == '\n' || c == '\r') continue;
      if (c == '.' || c == 'e' || c == 'E' || c == 'd' || c == 'D') dot = 1;
      b[n++] = (c == 'd' || c == 'D') ? 'e' : c;
   }
   b[n] = 0;
   if (!n) return 0.;
   return dot ? strtod(b, NULL) : strtod(b, NULL) / (d == 1 ? 10. : d == 2 ? 100. : pow(10., d));
}

/* a Fortran Iw field of columns [c0, c0 + w) */
static int fort_i(const char *line, int c0, int w)
{
   char b[32];
   int k, n = 0;
   const int len = (int) strlen(line);
   for (k = c0; k < c0 + w && k < len && n < 30; k++)
      if (line[k] != ' ' && line[k] != '\n' && line[k] != '\r') b[n++] = line[k];
   b[n] = 0;
   return n ? atoi(b) : 0;
}

/* INPUTN2's 1992 branch (inputn2.f:610-627): two title lines, 100 hourly rows
 * (15X,I5,8F8.2: hour, latitude, longitude west positive, speed, direction,
 * pressure deficit, RMW), IBGNT ITEND JHR (3I3), the landfall time INITAL
 * reads (inital.f:206-215: 2X,2I2.2,1X,I2.2,1X,A3,1X,I4) and the datums
 * (2F5.1,A1,F5.1).  No interpolation and no wind / RMW factors. */
static int trk_parse_1992(FILE *fp, sl_trk *t, char err[256])
{
   static const char *mon[12] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN",
                                 "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
   char line[512];
   int i, m;
   for (i = 0; i < 2; i++)
      if (!fgets(line, sizeof line, fp)) goto short_file;
   t->v1992 = 1;
   t->n = 100;
   for (i = 1; i <= 100; i++) {
      if (!fgets(line, sizeof line, fp)) goto short_file;
      t->hrs[i] = i;
      t->fla[i] = fort_f(line, 20, 8, 2);
      t->flo[i] = fort_f(line, 28, 8, 2);
      t->fdp[i] = fort_f(line, 52, 8, 2);
      t->frm[i] = fort_f(line, 60, 8, 2);
   }
   if (!fgets(line, sizeof line, fp)) goto short_file;
   t->beg = fort_i(line, 0, 3);
   t->end = fort_i(line, 3, 3);
   t->land = fort_i(line, 6, 3);
   if (!fgets(line, sizeof line, fp)) goto short_file;
   t->date_hr = fort_i(line, 2, 2);
   t->date_mn = fort_i(line, 4, 2);
   t->date_dy = fort_i(line, 7, 2);
   t->date_yr = fort_i(line, 14, 4);
   for (m = 0; m < 12; m++)
      if ((int) strlen(line) >= 13 && !strncmp(line + 10, mon[m], 3)) break;
   if (m == 12) {
      snprintf(err, 256, "track landfall time: no month in '%.20s'", line);
      return -1;
   }
   t->date_mo = m + 1;
   if (!fgets(line, sizeof line, fp)) goto short_file;
   t->ocean = fort_f(line, 0, 5, 1);
   t->lake = fort_f(line, 5, 5, 1);
   if ((int) strlen(line) > 10 && (line[10] == 'X' || line[10] == 'x')) {
      snprintf(err, 256, "Okeechobee canal datums (XOKE) are not supported");
      return -1;
   }
   t->dhr = 0;
   t->adj = 0;
   return 0;
short_file:
   snprintf(err, 256, "the track file ends early (1992 format: 2 titles, 100 rows, 3 lines)");
   return -1;
}

/* INPUTN2 (inputn2.f:84-627): the trk file into t; the first line says the
 * format (FileVersion=2013, else the 1992 format) */
int sl_trk_parse(const char *path, sl_trk *t, char err[256])
{
   FILE *fp = fopen(path, "r");
   char line[512];
   int press = 0, in_data = 0, rc;
   double amb = -999.9;

   sl_trk_init(t);
   if (!fp) {
      snprintf(err, 256, "cannot open track file %s", path);
      return -1;
   }
   if (!fgets(line, sizeof line, fp)) {
      fclose(fp);
      snprintf(err, 256, "empty track file %s", path);
      return -1;
   }
   {
      const char *v = starts(line, "FileVersion=") ? line + 12 : starts(line, " FileVersion=") ? line + 13
                    : starts(line, " FileVersion =") ? line + 14 : NULL;
      rewind(fp);
      if (!v) {
         rc = trk_parse_1992(fp, t, err);
         fclose(fp);
         return rc;
      }
      if (atoi(v) != 2013) {
         fclose(fp);
         snprintf(err, 256, "track file version %d: 1992 or 2013", atoi(v));
         return -1;
      }
   }
   while (fgets(line, sizeof line, fp)) {
      char *l = line;
      line[strcspn(line, "\r\n")] = 0;
      if (!in_data) {
         while (*l == ' ') l++;
         if (*l == '#') continue;
         if (starts(line, "Hour,") || starts(line, " Hour,") || starts(line, "Hour ,") ||
             starts(line, " Hour ,")) {
            in_data = 1;
            continue;
         }
         if (starts(line, "DateTime=")) {
            sscanf(line + 9, "%4d", &t->date_yr);
            sscanf(line + 14, "%2d", &t->date_mo);
            sscanf(line + 17, "%2d", &t->date_dy);
            sscanf(line + 20, "%2d", &t->date_hr);
            sscanf(line + 23, "%2d", &t->date_mn);
         } else if (starts(line, "DateHour=")) t->dhr = atoi(line + 9);
         else if (starts(line, "LandfallHour=")) t->land = atoi(line + 13);
         else if (starts(line, "BeginHour=")) t->beg = atoi(line + 10);
         else if (starts(line, "EndHour=")) t->end = atoi(line + 8);
         else if (starts(line, "AmbientPressure=")) amb = strtod(line + 16, NULL);
         else if (starts(line, "OceanDatum=")) t->ocean = strtod(line + 11, NULL);
         else if (starts(line, "LakeDatum=")) t->lake = strtod(line + 10, NULL);
         else if (starts(line, "WindFactor=")) t->windf = strtod(line + 11, NULL);
         else if (starts(line, "RMWFactor=")) t->rmwf = strtod(line + 10, NULL);
         else if (starts(line, "InterpolationType=")) {
            t->spline = (starts(line + 18, "Spline") || starts(line + 18, "spline"));
            if (!t->spline && !starts(line + 18, "Linear") && !starts(line + 18, "linear")) {
               fclose(fp);
               snprintf(err, 256, "unknown InterpolationType");
               return -1;
            }
         } else if (starts(line, "PressureType=")) press = atoi(line + 13);
         else if (starts(line, "CanalDatum=")) {
            if (strtod(line + 11, NULL) > 0.) {
               fclose(fp);
               snprintf(err, 256, "CanalDatum (Okeechobee basins) is not supported");
               return -1;
            }
         }
      } else {
         double v[6] = {0, 0, 0, 0, 0, 0}, dp;
         int nv = 0;
         char *p = line, *q;
         while (nv < 6) {
            while (*p == ' ' || *p == ',' || *p == '\t') p++;
            if (!*p) break;
            v[nv] = strtod(p, &q);
            if (q == p) break;
            nv++;
            p = q;
         }
         if (nv < 5) break;
         if ((int) v[0] == 999) break;
         if (press == 0) dp = v[3];
         else if (press == 1) dp = amb - v[3];
         else dp = v[5] - v[3];
         rc = sl_trk_row(t, v[0], v[1], v[2], dp, v[4], err);
         if (rc < 0) {
            fclose(fp);
            return -1;
         }
         if (rc > 0) break;
      }
   }
   fclose(fp);
   return 0;
}

/* INPUTN2 after the rows (inputn2.f:440-575): the header checked, the track
 * interpolated to every hour (linearly, or by the cubic spline), the wind and
 * RMW factors and the datums applied. */
int sl_trk_build(const sl_trk *t, sl_storm *s, char err[256])
{
   const int ntrk = t->n;
   int i, jj, ihour = 0;
   /* per call (tracks may be built on several threads at once) */
   int hrs[SL_NT + 2];
   double hla[SL_NT + 2], hlo[SL_NT + 2], hdp[SL_NT + 2], hrm[SL_NT + 2];
   const double *fla = t->fla, *flo = t->flo, *fdp = t->fdp, *frm = t->frm;

   if (t->beg < 0 || t->beg > 999 || t->end < 0 || t->end > 999 || t->dhr < 0 || t->land < 0 ||
       t->land > 999 || (!t->v1992 && (t->ocean < -50. || t->ocean > 50. || t->lake < -50. || t->lake > 50.)) ||
       t->beg >= t->end || ntrk < 2) {
      snprintf(err, 256, "track header incomplete or out of bounds");
      return -1;
   }
   s->date_yr = t->date_yr;
   s->date_mo = t->date_mo;
   s->date_dy = t->date_dy;
   s->date_hr = t->date_hr;
   s->date_mn = t->date_mn;
   memset(hla, 0, sizeof hla);
   memset(hlo, 0, sizeof hlo);
   memset(hdp, 0, sizeof hdp);
   memset(hrm, 0, sizeof hrm);
   memcpy(hrs, t->hrs, sizeof(int) * (size_t) (ntrk + 1));
   hrs[ntrk + 1] = 0;
   if (!t->spline) {
      for (i = 1; i <= ntrk; i++) {
         ihour = hrs[i];
         hla[ihour] = fla[i];
         hlo[ihour] = flo[i];
         hdp[ihour] = fdp[i];
         hrm[ihour] = frm[i];
         if ((hrs[i + 1] - hrs[i]) > 1 && i < ntrk) {
            int nt = hrs[i + 1] - hrs[i];
            for (jj = 1; jj <= nt - 1; jj++) {
               double ff = (double) jj / (double) nt;
               ihour = hrs[i] + jj;
               hla[ihour] = fla[i] + (ff * (fla[i + 1] - fla[i]));
               hlo[ihour] = flo[i] + (ff * (flo[i + 1] - flo[i]));
               hdp[ihour] = fdp[i] + (ff * (fdp[i + 1] - fdp[i]));
               hrm[ihour] = frm[i] + (ff * (frm[i + 1] - frm[i]));
            }
         }
      }
   } else {
      double xh[SL_NT + 2], y2a[SL_NT + 2], y2o[SL_NT + 2], y2d[SL_NT + 2], y2r[SL_NT + 2];
      for (i = 1; i <= ntrk; i++) xh[i] = hrs[i];
      spline(xh, fla, ntrk, 1.0E+30, 1.0E+30, y2a);
      spline(xh, flo, ntrk, 1.0E+30, 1.0E+30, y2o);
      spline(xh, fdp, ntrk, 1.0E+30, 1.0E+30, y2d);
      spline(xh, frm, ntrk, 1.0E+30, 1.0E+30, y2r);
      for (i = 1; i <= ntrk; i++) {
         ihour = hrs[i];
         hla[ihour] = fla[i];
         hlo[ihour] = flo[i];
         hdp[ihour] = fdp[i];
         hrm[ihour] = frm[i];
         if ((hrs[i + 1] - hrs[i]) > 1 && i < ntrk) {
            int nt = hrs[i + 1] - hrs[i];
            for (jj = 1; jj <= nt - 1; jj++) {
               ihour = hrs[i] + jj;
               hla[ihour] = splint(xh, fla, y2a, ntrk, (double) ihour);
               hlo[ihour] = splint(xh, flo, y2o, ntrk, (double) ihour);
               hdp[ihour] = splint(xh, fdp, y2d, ntrk, (double) ihour);
               hrm[ihour] = splint(xh, frm, y2r, ntrk, (double) ihour);
            }
         }
      }
   }
   s->seadtm = t->ocean;
   s->dtmlak = t->lake;
   s->ibgnt = t->beg + t->adj;
   s->itend = t->end + t->adj;
   s->jhr = t->land + t->adj;
   s->trklen = ihour;
   for (i = 1; i <= s->trklen; i++) {
      s->XLAT[i] = hla[i];
      s->YLONG[i] = hlo[i];
      s->PT[i] = hdp[i] * t->windf;
      s->R[i] = hrm[i] * t->rmwf;
   }
   if ((int) (s->seadtm * 10 + .5) == 999) {
      s->seadtm = 0;
      s->dtmlak = 0;
   }
   if ((s->seadtm >= 150) || (s->seadtm <= -250)) s->seadtm = fmod((s->seadtm + 50), 100.) - 50;
   if (s->itend + 1 > s->trklen || s->ibgnt < 1 || s->jhr < 4) {
      snprintf(err, 256, "track does not cover hours %d..%d", t->beg, t->end + 1);
      return -1;
   }
   return 0;
}

/* STERON, steron.f:36-72 */
static void steron(sl_storm *s, const sl_basin *b)
{
   const double erad = 3437.87, ftpnm = 6080.21;
   double so = sin(b->alt1c), co = cos(b->alt1c), erad2 = 2. * erad, co2 = 2. * co;
   int i;
   for (i = 1; i <= s->trklen; i++) {
      double altii = RADF * (90. - s->XLAT[i]);
      double alnii = RADF * s->YLONG[i];
      double dlat = b->alt1c - altii, dlat2 = .5 * dlat;
      double dlon = b->aln1 - alnii, dlon2 = .5 * dlon;
      double t = sin(dlon2), s2 = t * t, u = sin(dlat2);
      double ss = u * u + so * sin(altii) * s2;
      double rlz = erad2 * sqrt(ss / (1. - ss));
      double an = sin(altii) * sin(dlon);
      double ad = sin(dlat) + co2 * sin(altii) * s2;
      double ang;
      if (an * an + ad * ad < 1.E-10) ang = PI_F - b->phi1;
      else ang = PI_F + atan2(-an, -ad) - b->phi1;
      s->X[i] = (rlz * sin(ang)) * ftpnm / 5280. + b->xmouth;
      s->Y[i] = (rlz * cos(ang)) * ftpnm / 5280. + b->ymouth;
      if (i != 1) {
         double dx = s->X[i - 1] - s->X[i], dy = s->Y[i - 1] - s->Y[i];
         s->SP[i - 1] = sqrt(dx * dx + dy * dy);
         if (s->SP[i - 1] < 1.E-5) s->DIR[i - 1] = PI_F;
         else s->DIR[i - 1] = PI_F + atan2(s->Y[i - 1] - s->Y[i], s->X[i - 1] - s->X[i]);
      }
   }
   s->SP[s->trklen] = s->SP[s->trklen - 1];
   s->DIR[s->trklen] = s->DIR[s->trklen - 1];
   {
      double dx = s->X[s->jhr] - b->xmouth, dy = s->Y[s->jhr] - b->ymouth;
      s->rlngth = sqrt(dx * dx + dy * dy);
   }
}

/* INTWND, intwnd.f:43-82; the 150-entry tables are the reference's data. */
static const double WM[150] = {
   .10, .14, .17, .20, .23, .25, .28, .30, .32, .33, .35, .37, .38, .39,
   .41, .42, .43, .44, .45, .46, .46, .47, .48, .49, .49, .50, .51, .52, .52, .53,
   .53, .54, .54, .55, .56, .56, .57, .57, .58, .58, .59, .59, .60, .60, .61, .62,
   .62, .63, .63, .64, .65, .65, .66, .66, .67, .68, .69, .69, .70, .71, .71, .72,
   .72, .73, .73, .74, .74, .75, .75, .76, .76, .76, .77, .77, .77, .77, .78, .78,
   .78, .78, .79, .79, .79, .79, .80, .80, .80, .80, .80, .81, .81, .81, .81, .81,
   .82, .82, .82, .82, .82, .82, .83, .83, .83, .83, .84, .84, .84, .84, .84, .84,
   .84, .85, .85, .85, .85, .85, .85, .85, .86, .86, .86, .86, .86, .86, .86, .86,
   .86, .86, .87, .87, .87, .87, .87, .87, .87, .87, .87, .87, .87, .87, .87, .87,
   .87, .87, .87, .87, .87, .87, .87, .87};
static const double WB[150] = {
   12., 23., 29., 33., 37., 39., 41., 43., 45., 47., 49., 51., 53., 55.,
   57., 58., 60., 62., 63., 65., 67., 68., 70., 71., 73., 74., 76., 77., 78., 80.,
   81., 82., 84., 85., 86., 88., 89., 90., 92., 93., 94., 95., 96., 98., 99., 100.,
   101., 102., 104., 105., 106., 107., 108., 109., 110., 112., 113., 114., 115.,
   116., 117., 118., 119., 120., 121., 122., 123., 123., 124., 125., 126., 127.,
   128., 129., 130., 131., 132., 132., 133., 134., 135., 136., 137., 137., 138.,
   139., 140., 141., 141., 142., 143., 144., 145., 145., 146., 147., 148., 148.,
   149., 150., 151., 151., 152., 153., 154., 154., 155., 156., 157., 157., 158.,
   158., 159., 160., 160., 161., 162., 162., 163., 164., 164., 165., 166., 166.,
   167., 167., 168., 169., 169., 170., 171., 171., 172., 172., 173., 174., 174.,
   175., 175., 176., 176., 177., 177., 178., 179., 179., 180., 180., 181., 182.};

double sl_intwnd(double pdrop, double rmax, double alto)
{
   double q = pdrop / 67.7, pf = .3 * (30. - alto) * (q * q), vmax;
   int kpres = (int) (pdrop + pf + 0.5);
   if (kpres >= 151) kpres = 150;
   if (kpres < 1) kpres = 1;
   vmax = -WM[kpres - 1] * rmax + WB[kpres - 1];
   return vmax > 10. ? vmax : 10.;
}

/* FORCE1, force1.f:55-183.  Returns 0, or -1 where the reference STOPs. */
int sl_force1(double pdrop, double rmax, double *vmaxp, double alto, int ncatg,
              double S[SL_NTAB + 1], double C[SL_NTAB + 1], double P[SL_NTAB + 1],
              double DELP[SL_NTAB + 1])
{
   const double rhoa = 2.298E-3, rhow = 4905., rhowg = 29.89, epsln = 1.E-5;
   const int ic12 = 795;
   double vmax = *vmaxp, cor, corhr, cfric, xks, xkn, r0, delr, veer0, cotphi, sinphi, cosphi;
   double osinfi, w0, press, dpdr0, fact, slp0, r1, veer1, w1, slp1, padd, dgdw, delw, dosinf;
   double cos1, dpdr1, term, pinft, prs;
   int jset, j2, mmax, i, kk, k;
   int t12 = (int) (12. * rmax);

#define VEE(R) (2. * vmax * rmax * (R) / (rmax * rmax + (R) * (R)))
#define DERLNV(R) (1. / (R) - 2. * (R) / (rmax * rmax + (R) * (R)))
#define PRGRAD(R, V) (rhoa * (V) * (V) * (xks / sinphi - DERLNV(R)))
#define PYTHAG(X) (sqrt(1. - (X) * (X)))
   jset = ic12 - 51;
   if ((150 > t12 ? 150 : t12) < jset) jset = (150 > t12 ? 150 : t12);
   cor = 2. * (7.292116E-5) * sin(fabs(alto) * 1.74532925199433E-2);
   corhr = 3600. * cor;
   cfric = 1.;
   if (ncatg == 1) cfric = 4. * sqrt(22. / rmax);
   for (;;) {
      xks = .01 * sqrt(rmax / (.3 * vmax + 60.));
      xkn = .01 * sqrt(rmax / (.4 * vmax + 80.));
      xks = xks * cfric;
      xkn = xkn * cfric;
      {
         double t = 3. * rmax;
         j2 = (int) (t < 50. ? t : 50.);
      }
      r0 = jset + 50;
      mmax = jset + 9 * j2 + 49;
      delr = 1.;
      veer0 = VEE(r0) * r0;
      cotphi = (corhr * r0 / veer0 + xkn) / xks;
      sinphi = 1. / sqrt(1. + cotphi * cotphi);
      cosphi = PYTHAG(sinphi);
      osinfi = 1. / sinphi;
      w0 = veer0 * cosphi;
      press = 0.;
      dpdr0 = PRGRAD(r0, veer0 / r0);
      for (i = 1; i <= mmax; i++) {
         int conv = 0;
         fact = 2. / delr;
         slp0 = w0 * xks / sinphi - corhr * r0 - xkn * veer0;
         r1 = r0 - delr;
         veer1 = VEE(r1) * r1;
         w1 = veer1 * cosphi;
         for (kk = 1; kk <= 75; kk++) {
            slp1 = w1 * xks / sinphi - corhr * r1 - xkn * veer1;
            padd = fact * (w0 - w1) - slp0;
            dgdw = xks / (sinphi * sinphi * sinphi);
            delw = (slp1 - padd) / (dgdw + fact);
            dosinf = -delw * cosphi / veer1 / sinphi / sinphi / sinphi;
            osinfi = osinfi + dosinf;
            sinphi = 1. / osinfi;
            cos1 = cosphi;
            cosphi = PYTHAG(sinphi);
            w1 = veer1 * cosphi;
            if (w1 <= 0) return -1;                       /* STOP, force1.f:177 */
            if (fabs((cos1 - cosphi) / cos1) - epsln <= 0) {
               conv = 1;
               break;
            }
         }
         if (!conv) return -1;                            /* STOP 999 */
         r0 = r1;
         k = (int) (r1 + 1.01);
         veer0 = veer1;
         w0 = w1;
         S[k] = sinphi;
         C[k] = cosphi;
         dpdr1 = PRGRAD(r1, veer1 / r1);
         P[k] = dpdr1;
         press = press - (dpdr0 + dpdr1) / fact;
         DELP[k] = press;
         dpdr0 = dpdr1;
         if (k - j2 - 1 <= 0) delr = .1;
      }
      DELP[1] = press - dpdr0 / fact;
      P[1] = 0.;
      S[1] = 0.;
      C[1] = 1.;
      r0 = jset - 1;
      term = r0 / (rmax * rmax + r0 * r0);
      {
         double vr = vmax * rmax;
         padd = 2. * rhoa * (vr * vr) * (xks * (atan(rmax / r0) / rmax + term) / S[jset] +
                                         term * term);
      }
      pinft = padd + DELP[jset] - DELP[1];
      prs = pdrop;
      if (fabs(pinft - prs) - .2 <= 0) break;
      vmax = vmax * sqrt(prs / pinft);
   }
   sinphi = S[jset - 1];
   cosphi = C[jset - 1];
   for (k = jset + 1; k <= ic12; k++) {
      r0 = k + 1;
      S[k] = sinphi;
      C[k] = cosphi;
      P[k] = PRGRAD(r0, VEE(r0));
      DELP[k] = DELP[k - 1] + .5 * (P[k] + P[k - 1]);
   }
   press = padd + DELP[jset];
   for (i = 1; i <= ic12; i++) {
      DELP[i] = (press - DELP[i]) / rhowg;
      P[i] = P[i] / rhow;
   }
   for (i = 2; i <= ic12; i++) P[i] = P[i] / ((i - 1.) * 5280.);
   *vmaxp = vmax;
#undef VEE
#undef DERLNV
#undef PRGRAD
#undef PYTHAG
   return 0;
}

/* CNSDLT, cnsdlt.f:27-44 */
static void cnsdlt(sl_storm *s, const sl_basin *b)
{
   double g = 32.2;
   s->x1b[0] = g * s->delt / b->dela;
   s->x1b[1] = s->delt / b->dela;
   s->x1b[3] = s->delt * b->cor;
   s->x1b[4] = .5 * s->x1b[1];
   s->x1b[5] = .5 * s->x1b[0];
   s->x1b[6] = .6 * s->delt * sqrt(2. * g / 3.) / 3.;
}

int sl_storm_load(const char *trk_path, const sl_basin *b, sl_storm *s, char err[256])
{
   sl_trk *t = (sl_trk *) malloc(sizeof *t);
   int rc;
   err[0] = 0;
   if (!t) {
      snprintf(err, 256, "out of host memory");
      return -1;
   }
   rc = sl_trk_parse(trk_path, t, err);
   if (!rc) rc = sl_storm_init(t, b, s, err);
   free(t);
   return rc;
}

int sl_storm_init(const sl_trk *trk, const sl_basin *b, sl_storm *s, char err[256])
{
   int i, ncatg, igrwth;
   double t, vmax;
   memset(s, 0, sizeof *s);
   err[0] = 0;
   if (sl_trk_build(trk, s, err)) return -1;
   s->dtmlak = s->dtmlak + b->dtmlak_add;   /* DEPMSY, depmsy.f:24 ('2$' basins) */
   steron(s, b);
   /* INTDLT, intdlt.f:40-62 */
   s->dltb = b->dltin[1];
   if (!(s->rlngth > 90.)) {
      double m = s->PT[s->jhr] > s->PT[s->jhr - 3] ? s->PT[s->jhr] : s->PT[s->jhr - 3];
      int maxp = (int) m;
      if (maxp > 60) s->dltb = b->dltin[2];
   }
   s->delt = b->dltin[0];
   memcpy(s->x1b, b->x1b, sizeof s->x1b);
   /* INTVAL, intval.f:52-104 */
   s->ndlthr = (int) (3600. / (s->delt - .001));
   s->ndlth2 = s->ndlthr / 2;
   s->dlthr = s->ndlthr;
   s->dlth2 = s->ndlth2;
   s->iprtad = s->ndlthr / 4;
   if (s->ndlthr % 6 == 0) s->iprtad = s->ndlthr / 6;
   s->iprtsv = s->iprtad;
   t = s->itend - s->ibgnt;
   s->mhalt = (int) ((t * 3600. + 1.) / s->delt);
   s->ic19 = 2 * s->ndlthr;
   if (s->SP[s->ibgnt] < 1.E-5) igrwth = 4 * s->ndlthr + 4;
   else igrwth = (int) ((2. * s->R[s->ibgnt] / s->SP[s->ibgnt]) * s->dlthr);
   {
      int a = 4 * s->ndlthr + 4 < igrwth ? 4 * s->ndlthr + 4 : igrwth;
      if (a > s->ic19) s->ic19 = a;
   }
   for (i = 1; i <= 5; i++) {
      int j = s->ibgnt + i - 1;
      if (fabs(s->X[j] - 1.) < (.25 * s->R[s->ibgnt])) s->ic19 = 4 * s->ndlthr;
   }
   s->ic12 = 0;
   s->key1 = 1;
   s->jumpr = 1;
   /* SETCMP, setcmp.f:77-180 */
   s->itime = 0;
   s->treal = 0.;
   s->nhrad = s->ndlthr;
   s->ax = 0.;
   s->ay = 0.;
   s->c1 = s->X[s->ibgnt] * 5280.;
   s->c2 = s->Y[s->ibgnt] * 5280.;
   s->zlato = s->XLAT[s->ibgnt];
   s->c21 = s->SP[s->ibgnt];
   s->c22 = s->DIR[s->ibgnt];
   s->ydelp = s->PT[s->ibgnt];
   s->zdelp = s->PT[s->ibgnt];
   s->yc24 = s->R[s->ibgnt];
   s->zc24 = s->R[s->ibgnt];
   s->x12[20] = cos(s->c22);
   s->x12[21] = sin(s->c22);
   s->x12[7] = s->c21 * 44. / 30.;
   s->x12[9] = s->x12[7] * s->x12[20];
   s->x12[10] = s->x12[7] * s->x12[21];
   s->x12[11] = s->delt * s->x12[9];
   s->x12[12] = s->delt * s->x12[10];
   s->x12[23] = s->x12[7] * cos(s->c22 + 2. * PI_F / 3.);
   s->x12[24] = s->x12[7] * sin(s->c22 + 2. * PI_F / 3.);
   s->ptency = (s->PT[s->ibgnt + 1] - s->PT[s->ibgnt]) / s->ndlthr;
   s->rtency = (s->R[s->ibgnt + 1] - s->R[s->ibgnt]) / s->ndlthr;
   s->itmadv = s->ibgnt + 1;
   s->x12[49] = s->c1 + s->ax;
   s->x12[50] = s->c2 + s->ay;
   s->pnn = s->ydelp;
   vmax = sl_intwnd(s->ydelp, s->yc24, s->zlato);
   s->wmax[1] = vmax;
   s->wmax[0] = vmax * .85;
   for (ncatg = 1; ncatg <= 2; ncatg++) {
      vmax = s->wmax[ncatg - 1];
      if (sl_force1(s->ydelp, s->yc24, &vmax, s->zlato, ncatg, s->S, s->C, s->P, s->DELP)) {
         snprintf(err, 256, "FORCE1 did not converge for the first-hour storm");
         return -1;
      }
      s->wmax[ncatg - 1] = vmax;
      s->w1218[ncatg - 1] = 2. * vmax * 44. / 30.;
      if (ncatg == 1)
         for (i = 1; i <= 800; i++) {
            s->SW[i] = s->S[i];
            s->CW[i] = s->C[i];
         }
   }
   s->x12[4] = s->yc24 * 5280.;
   s->x12[15] = s->x12[4] * s->x12[4];
   s->x12[17] = 3.E-6 * s->x12[15];
   s->table_gen = 1;
   return 0;
}

/* TTIMER, ttimer.f:33-42 */
void sl_storm_ttimer(sl_storm *s)
{
   s->itime = s->itime + 1;
   s->bt = 1.;
   if (s->itime < s->ic19) {
      double time = s->itime, ang = time / s->ic19;
      s->bt = .5 * (1. - cos(ang * PI_F));
   }
}

/* STMVAL, stmval.f:95-236.  Returns 0, -1 where FORCE1 would STOP. */
int sl_storm_stmval(sl_storm *s, const sl_basin *b)
{
   int ncatg, i, np;
   double pn, cz, vmax;
   if (s->itime - 1 == s->nhrad) {
      if (s->dltb != b->dltin[0]) {
         double delt1 = 0;
         int sw = 0;
         if (s->itmadv == s->jhr - 6) {
            delt1 = b->dltin[1];
            sw = 1;
         } else if (s->itmadv == s->jhr && s->dltb != b->dltin[1]) {
            delt1 = b->dltin[2];
            sw = 1;
         }
         if (sw) {
            s->mhalt = (int) ((s->itime - 1) + (s->mhalt - (s->itime - 1)) * s->delt / delt1 + .01);
            s->delt = delt1;
            s->ndlthr = (int) (3600. / (s->delt - .001));
            s->iprtsv = s->iprtsv - s->iprtad;
            s->iprtad = s->ndlthr / 4;
            if (s->ndlthr % 6 == 0) s->iprtad = s->ndlthr / 6;
            s->iprtsv = s->iprtsv + s->iprtad;
            cnsdlt(s, b);
         }
      }
      s->nhrad = s->nhrad + s->ndlthr;
      s->zlato = s->XLAT[s->itmadv];
      s->c22 = s->DIR[s->itmadv];
      s->x12[20] = cos(s->c22);
      s->x12[21] = sin(s->c22);
      s->c21 = s->SP[s->itmadv];
      s->x12[7] = s->c21 * 44. / 30.;
      s->x12[9] = s->x12[7] * s->x12[20];
      s->x12[10] = s->x12[7] * s->x12[21];
      s->x12[11] = s->delt * s->x12[9];
      s->x12[12] = s->delt * s->x12[10];
      s->x12[23] = s->x12[7] * cos(s->c22 + 2. * PI_F / 3.);
      s->x12[24] = s->x12[7] * sin(s->c22 + 2. * PI_F / 3.);
      s->rtency = (s->R[s->itmadv + 1] - s->R[s->itmadv]) / s->ndlthr;
      s->ptency = (s->PT[s->itmadv + 1] - s->PT[s->itmadv]) / s->ndlthr;
      s->yc24 = s->R[s->itmadv];
      s->ydelp = s->PT[s->itmadv];
      s->itmadv = s->itmadv + 1;
   }
   s->ax = s->ax + s->x12[11];
   s->ay = s->ay + s->x12[12];
   s->treal = s->treal + s->delt;
   s->ydelp = s->ydelp + s->ptency;
   s->yc24 = s->yc24 + s->rtency;
   if (s->jumpr < 2 && s->itmadv == s->jhr + 1) {
      s->jumpr = 3;                        /* IOPERL(1) = 2: operational mode */
      s->iprt = 1;
   } else {
      if (s->key1 == 0) return 0;
      if (fabs(s->ydelp - s->zdelp) < 1. && fabs(s->yc24 - s->zc24) < 1.) return 0;
   }
   np = (int) (s->ydelp * 10. + .5);
   pn = np / 10.;
   cz = sqrt(pn / s->pnn);
   for (ncatg = 1; ncatg <= 2; ncatg++) {
      vmax = s->wmax[ncatg - 1] * cz;
      if (sl_force1(pn, s->yc24, &vmax, s->zlato, ncatg, s->S, s->C, s->P, s->DELP)) return -1;
      s->wmax[ncatg - 1] = vmax;
      s->w1218[ncatg - 1] = 2. * vmax * 44. / 30.;
      if (ncatg == 1)
         for (i = 1; i <= 800; i++) {
            s->SW[i] = s->S[i];
            s->CW[i] = s->C[i];
         }
   }
   s->pnn = pn;
   s->zdelp = s->ydelp;
   s->zc24 = s->yc24;
   s->x12[4] = s->yc24 * 5280.;
   s->x12[15] = s->x12[4] * s->x12[4];
   s->x12[17] = 3.E-6 * s->x12[15];
   s->table_gen++;
   return 0;
}

/* XY2LLX, xy2llx.f (storm centre for the rex track record) */
void sl_storm_latlon(const sl_storm *s, const sl_basin *b, double *lat, double *lon)
{
   double xa = (s->c1 + s->ax) / 5280., ya = (s->c2 + s->ay) / 5280.;
   double clto = cos(b->alt1c), slto = sin(b->alt1c);
   double xc = (xa - b->xmouth) * 5280. / 6080.21, yc = (ya - b->ymouth) * 5280. / 6080.21;
   double rll = sqrt(xc * xc + yc * yc), enu, psi, spsi, cpsi, cnu, dlat, dlong;
   enu = (rll < 1.E-5) ? b->phi1 : atan2(xc, yc) + b->phi1;
   psi = 2. * atan(.5 * rll / 3437.87);
   spsi = sin(psi);
   cpsi = cos(psi);
   cnu = cos(enu);
   dlat = acos(clto * cpsi + spsi * slto * cnu);
   dlong = b->aln1 - atan2(spsi * sin(enu), cpsi * slto - spsi * cnu * clto);
   *lat = 90. - dlat / 1.74532925199433E-2;
   *lon = dlong / 1.74532925199433E-2;
}

/* MSY's first lake column of each row (LDTMG, flw1dm_block_c.f:67-69) */
static int ldtmg(int j)
{
   static const int t[32] = {33, 33, 33, 33, 33, 35, 35, 35, 35, 35, 35, 35, 35, 34, 34, 34,
                             34, 34, 30, 29, 29, 28, 27, 26, 26, 24, 22, 20, 21, 22, 22, 21};
   return (j >= 1 && j <= 32) ? t[j - 1] : 1;
}

int sl_intlht_class(const sl_basin *b, int i, int j)
{
   const char *it = b->itree;
   char a = it[PIX(b, i, j)], c = it[PIX(b, i + 1, j)], d = it[PIX(b, i, j + 1)], e = it[PIX(b, i + 1, j + 1)];
   int k, cls;
   if (!strcmp(b->dollar, "2$")) {   /* the MSY branch, intlht.f:96-121 */
      if (i <= ldtmg(j)) cls = 0;
      else if (a == '2' || d == '2' || c == '2' || e == '2') cls = SL_IW_STORM;
      else cls = SL_IW_SEA;
   } else {                          /* intlht.f:55-78 */
      cls = (a == '2' || a == '5' || c == '2' || c == '5' || d == '2' || d == '5' || e == '2' || e == '5')
               ? SL_IW_STORM : 0;
   }
   /* '1$': the listed squares dry out (intlht.f:81-92); MSY's list is empty
    * once INITALCOMMON has cleared COMMON /LDRY/ (initalcommon.f:129-131) */
   if (!strcmp(b->dollar, "1$"))
      for (k = 0; k < b->nodry; k++)
         if (b->idry[k] == i && b->jdry[k] == j) cls |= SL_IW_DRY;
   return cls;
}

/* INTLHT, intlht.f:53-185 without the Okeechobee branch (+ the J=JMXB copy :161) */
void sl_initial_water(const sl_basin *b, const sl_storm *s, double *hb)
{
   int i, j;
   memset(hb, 0, sizeof(double) * (size_t) b->imxb * b->jmxb);
   for (j = 1; j <= b->jmxb1; j++)
      for (i = 1; i <= b->imxb1; i++) {
         const int cls = sl_intlht_class(b, i, j);
         double zb = b->zb[PIX(b, i, j)], h;
         if (cls & SL_IW_STORM) {
            double xr = b->elpcl[i - 1] * b->cosl[j - 1], yr = b->elpdl[i - 1] * b->sinl[j - 1];
            double x = xr - s->c1 - s->ax, y = yr - s->c2 - s->ay;
            double rsq = x * x + y * y, r1 = sqrt(rsq) / 5280. + 1., r2, dr;
            int k = (int) r1;
            r2 = k;
            dr = r1 - r2;
            if (k > 790) k = 790;
            h = s->DELP[k] + dr * (s->DELP[k + 1] - s->DELP[k]) + s->seadtm;
            hb[PIX(b, i, j)] = h > -zb ? h : -zb;
         } else {
            const double dtm = (cls & SL_IW_SEA) ? s->seadtm : s->dtmlak;
            double t = dtm > -zb ? dtm : -zb;
            hb[PIX(b, i, j)] = -zb > t ? -zb : t;
         }
      }
   if (b->nodry > 0)
      for (j = 1; j <= b->jmxb1; j++)
         for (i = 1; i <= b->imxb1; i++)
            if (sl_intlht_class(b, i, j) & SL_IW_DRY) hb[PIX(b, i, j)] = -b->zb[PIX(b, i, j)];
   for (i = 1; i <= b->imxb1; i++) hb[PIX(b, i, b->jmxb)] = hb[PIX(b, i, 1)];
}
