/* meow_ref.c -- test infrastructure: MEOW seed points over the reference's own
 * projection code.
 *
 * Links the reference's sorc/rexout.cd/convert.c (load_bsndta, cl2cnf,
 * cnf2cl, sll2xy, sxy2ll, pq2xy; compiled from its own source by
 * oracle/Makefile, output under oracle/_ref/) and walks the seed points as the
 * reference's MEOW generator does (conMeowGenCmd, convert.c:1181-1288, with
 * convertShadow, :972-1009, and Earth_Jump, :1127-1151).  Those three exist in
 * the reference only inside its Tcl GUI build (INCLUDE_TCL), which needs Tcl
 * headers this image lacks, so their steps are written out here; everything
 * they call is the reference's compiled code.
 *
 * usage: meowref <bnt file> <basin abbrev> <xyflag> <lat> <lon> <dir> <apart> <num> <r34>
 * prints one line per seed: "<index> <lat> <lon>" with %.17g.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "convert.h"

#define H_PI 3.1415926535     /* convert.c:970 */
#define H_PI_180 0.01745329251994
#define H_C180_PI 57.29577951308

static bsndta_type bsn;
static tanplane_type slsh;

static void jump(LatLon p, double dist, double bear, LatLon *q)
{
   double R, A, B, C;
   p.lat *= H_PI_180;
   p.lon *= H_PI_180;
   dist = dist / 1.151;
   R = 60. * H_C180_PI;
   dist = dist / R;
   bear = H_PI - (bear * H_PI_180);
   q->lat = asin(cos(dist) * sin(p.lat) - sin(dist) * cos(bear) * cos(p.lat));
   A = cos(dist) * cos(p.lat);
   B = sin(dist) * cos(bear) * sin(p.lat);
   C = sin(dist) * sin(bear);
   q->lon = atan2(A * sin(p.lon) + B * sin(p.lon) + C * cos(p.lon),
                  A * cos(p.lon) + B * cos(p.lon) - C * sin(p.lon));
   q->lat = q->lat * H_C180_PI;
   q->lon = q->lon * H_C180_PI;
}

static void extent(double dir, double x0, double y0, double *lo, double *hi)
{
   double x1, y1, d;
   int i, j;
   for (i = 0; i < bsn.ximax; i++)
      for (j = 0; j < bsn.yjmax; j += bsn.yjmax - 1) {
         pq2xy(i + 1, j + 1, &x1, &y1, &bsn);
         d = (x1 - x0) * sin(dir) + (y1 - y0) * cos(dir);
         if (i == 0 && j == 0) {
            *lo = d;
            *hi = d;
         } else {
            if (*lo > d) *lo = d;
            if (*hi < d) *hi = d;
         }
      }
   for (i = 0; i < bsn.ximax; i += bsn.ximax - 1)
      for (j = 1; j < bsn.yjmax - 1; j++) {
         pq2xy(i + 1, j + 1, &x1, &y1, &bsn);
         d = (x1 - x0) * sin(dir) + (y1 - y0) * cos(dir);
         if (*lo > d) *lo = d;
         if (*hi < d) *hi = d;
      }
}

static void point(double x0, double y0, double cur, double dir, LatLon *s)
{
   double t;
   sxy2ll(x0 + cur * sin(dir), y0 + cur * cos(dir), s, &slsh);
   cnf2cl(s->lat, &s->lat, &t);
}

int main(int argc, char **argv)
{
   LatLon pt, seed;
   double dir, Dir, r34, lo = 0, hi = 0, t, x0, y0, apart, start, cur;
   int num, cnt, cnt2, xy;
   if (argc != 10) {
      fprintf(stderr, "usage: meowref bnt abbrev xyflag lat lon dir apart num r34\n");
      return 2;
   }
   if (load_bsndta(argv[1], argv[2], &bsn, &slsh)) return 1;
   xy = atoi(argv[3]);
   pt.lat = atof(argv[4]);
   pt.lon = atof(argv[5]);
   dir = atof(argv[6]) + 90;
   if (dir > 360) dir -= 360;
   apart = atof(argv[7]);
   num = atoi(argv[8]);
   r34 = atof(argv[9]);
   Dir = dir;
   dir = dir * H_PI / 180.;
   cl2cnf(pt.lat, &pt.lat, &t);
   sll2xy(pt, &x0, &y0, &slsh);
   extent(dir, x0, y0, &lo, &hi);
   lo -= r34;
   hi += r34;
   start = 0;
   while (start < lo) start += apart;
   while (start > hi) start -= apart;
   cnt = 0;
   cur = start;
   point(x0, y0, cur, dir, &seed);
   printf("%d %.17g %.17g\n", cnt, seed.lat, seed.lon);
   for (;;) {
      cnt++;
      cur += apart;
      if (cur > hi || (num != -1 && cnt >= num / 2 + 1)) break;
      if (xy == 0) point(x0, y0, cur, dir, &seed);
      else jump(seed, apart, Dir, &seed);
      printf("%d %.17g %.17g\n", cnt, seed.lat, seed.lon);
   }
   cur = start;
   if (xy == 1) point(x0, y0, cur, dir, &seed);
   cnt2 = 0;
   for (;;) {
      cnt++;
      cnt2++;
      cur -= apart;
      if (cur < lo || (num != -1 && cnt >= num)) break;
      if (xy == 0) point(x0, y0, cur, dir, &seed);
      else jump(seed, -apart, Dir, &seed);
      printf("-%d %.17g %.17g\n", cnt2, seed.lat, seed.lon);
   }
   return 0;
}
