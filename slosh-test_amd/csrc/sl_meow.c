/* sl_meow.c -- MEOW track sets without files: the landfall points of a family
 * of parallel tracks across a basin, and the straight hourly tracks through
 * them.
 *
 * Restates, in double precision:
 *   conMeowGenCmd  seed points          sorc/rexout.cd/convert.c:1181-1288
 *   convertShadow  the basin's extent   sorc/rexout.cd/convert.c:972-1009
 *                  across a direction
 *   Earth_Jump     great-circle step    sorc/rexout.cd/convert.c:1127-1151
 * over the tangent-plane projection of the basin (sl_basin.c, savellx.c).
 * The reference runs these inside its Tcl GUI (halo_conMeowGen); the steps
 * and constants are kept as written there so the points come out the same.
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "sl_host.h"

#define MW_PI 3.1415926535        /* the value convert.c:970 defines */
#define MW_PI_180 0.01745329251994
#define MW_C180_PI 57.29577951308

/* Earth_Jump: dist in statute miles, bearing in degrees from north */
void sl_earth_jump(double lat, double lon, double dist, double bear, double *lat2, double *lon2)
{
   double R, A, B, C, la, lo;
   la = lat * MW_PI_180;
   lo = lon * MW_PI_180;
   dist = dist / 1.151;            /* nautical miles */
   R = 60. * MW_C180_PI;
   dist = dist / R;
   bear = MW_PI - (bear * MW_PI_180);
   *lat2 = asin(cos(dist) * sin(la) - sin(dist) * cos(bear) * cos(la));
   A = cos(dist) * cos(la);
   B = sin(dist) * cos(bear) * sin(la);
   C = sin(dist) * sin(bear);
   *lon2 = atan2(A * sin(lo) + B * sin(lo) + C * cos(lo), A * cos(lo) + B * cos(lo) - C * sin(lo));
   *lat2 = *lat2 * MW_C180_PI;
   *lon2 = *lon2 * MW_C180_PI;
}

/* convertShadow: the least and greatest distance along dir (radians, compass)
 * of the basin's outer ring of grid points from (x0, y0) */
static void shadow(const sl_basin *b, double dir, double x0, double y0, double *mn, double *mx)
{
   const double ximax = b->proj.ximax, yjmax = b->proj.yjmax;
   double x1, y1, dist;
   int i, j;
   for (i = 0; i < ximax; i++)
      for (j = 0; j < yjmax; j += (int) (yjmax - 1)) {
         sl_proj_pq2xy(b, i + 1, j + 1, &x1, &y1);
         dist = (x1 - x0) * sin(dir) + (y1 - y0) * cos(dir);
         if (i == 0 && j == 0) {
            *mn = dist;
            *mx = dist;
         } else {
            if (*mn > dist) *mn = dist;
            if (*mx < dist) *mx = dist;
         }
         if ((int) (yjmax - 1) <= 0) break;
      }
   for (i = 0; i < ximax; i += (int) (ximax - 1)) {
      for (j = 1; j < yjmax - 1; j++) {
         sl_proj_pq2xy(b, i + 1, j + 1, &x1, &y1);
         dist = (x1 - x0) * sin(dir) + (y1 - y0) * cos(dir);
         if (*mn > dist) *mn = dist;
         if (*mx < dist) *mx = dist;
      }
      if ((int) (ximax - 1) <= 0) break;
   }
}

/* the point cur miles from (x0, y0) along dir, as a Clarke lat/lon */
static void along(const sl_basin *b, double x0, double y0, double cur, double dir, double *lat, double *lon)
{
   double x1 = x0 + cur * sin(dir), y1 = y0 + cur * cos(dir), la;
   sl_proj_sxy2ll(b, x1, y1, &la, lon);
   sl_cnf2cl(la, lat);
}

int sl_meow_seeds(const sl_basin *b, int xyflag, double lat, double lon, double dir_deg, double apart, int num,
                  double r34, int cap, int *idx, double *slat, double *slon)
{
   double dir, Dir, mn = 0, mx = 0, temp, x0, y0, start, cur, la, lo;
   int cnt, cnt2, n = 0;
   if (!b->proj.have || !(apart > 0.) || cap < 1 || b->proj.ximax < 2 || b->proj.yjmax < 2) return -1;
#define EMIT(k, a, o)                                                                                  \
   do {                                                                                                \
      if (n >= cap) return -1;                                                                         \
      idx[n] = (k);                                                                                    \
      slat[n] = (a);                                                                                   \
      slon[n] = (o);                                                                                   \
      n++;                                                                                             \
   } while (0)
   dir = dir_deg + 90;
   if (dir > 360) dir -= 360;
   Dir = dir;
   dir = dir * MW_PI / 180.;
   sl_cl2cnf(lat, &lat, &temp);
   sl_proj_sll2xy(b, lat, lon, &x0, &y0);
   shadow(b, dir, x0, y0, &mn, &mx);
   mn -= r34;
   mx += r34;
   start = 0;
   while (start < mn) start += apart;
   while (start > mx) start -= apart;
   /* min <= start <= max */
   cnt = 0;
   cur = start;
   along(b, x0, y0, cur, dir, &la, &lo);
   EMIT(0, la, lo);
   for (;;) {
      cnt++;
      cur += apart;
      if ((cur > mx) || ((num != -1) && (cnt >= (num / 2) + 1))) break;
      if (xyflag == 0) along(b, x0, y0, cur, dir, &la, &lo);
      else sl_earth_jump(la, lo, apart, Dir, &la, &lo);
      EMIT(cnt, la, lo);
   }
   /* back to the first point, then the other way */
   cur = start;
   if (xyflag == 1) along(b, x0, y0, cur, dir, &la, &lo);
   cnt2 = 0;
   for (;;) {
      cnt++;
      cnt2++;
      cur -= apart;
      if ((cur < mn) || ((num != -1) && (cnt >= num))) break;
      if (xyflag == 0) along(b, x0, y0, cur, dir, &la, &lo);
      else sl_earth_jump(la, lo, -apart, Dir, &la, &lo);
      EMIT(-cnt2, la, lo);
   }
#undef EMIT
   return n;
}
