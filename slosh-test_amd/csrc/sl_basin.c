/* sl_basin.c -- basin ingest and static fields of the SLOSH step.
 *
 * Restates, in double precision, the reference's basin initialisation:
 *   CRDRD1  header           sorc/slosh.fd/crdrd1.f:102-213
 *   CRDRD2  body             sorc/slosh.fd/crdrd2.f:65-489
 *   DEPSFC  masks / bounds   sorc/slosh.fd/depsfc.f:62-451
 *   SETDST  grid metrics     sorc/slosh.fd/setdst.f:47-87
 *   BTMSTR  stress tables    sorc/slosh.fd/btmstr.f:79-214 (COMPLEX_p16 branch)
 *   CNSTNT  constants        sorc/slosh.fd/cnstnt.f:36-69
 *   memSetLLx cell latitude  sorc/slosh.fd/savellx.c:40-227, 396-431, 704-765
 * for the single-process case (FRANK: ISTAR=JSTAR=1, so MAPSHIFT is the
 * identity).  Sub-grid 1-D flow, weirs, channels, cuts, Manning friction and
 * the MSY / Okeechobee / dry-cell basin variants are detected and rejected.
 */
#define _GNU_SOURCE 1   /* sincos */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <ctype.h>
#include "sl_host.h"

#define PIX(b, i, j) ((size_t) ((j) - 1) * (size_t) (b)->imxb + (size_t) ((i) - 1))
#define RADF 1.74532925199433E-2

/* ------------------------------------------------------------------------ */
/* Fortran formatted READ of one record: fixed-width fields, blanks ignored,
 * implied decimal for F fields without a point, short records padded.       */
typedef struct {
   FILE *fp;
   char line[4096];
   int len, col, eof;
} frec;

static int fr_next(frec *r)
{
   if (!fgets(r->line, sizeof r->line, r->fp)) {
      r->line[0] = 0;
      r->len = 0;
      r->col = 0;
      r->eof = 1;
      return -1;
   }
   r->len = (int) strcspn(r->line, "\r\n");
   r->line[r->len] = 0;
   r->col = 0;
   return 0;
}

static void fr_field(frec *r, int w, char *out)
{
   int k;
   for (k = 0; k < w; k++) {
      int c = r->col + k;
      out[k] = (c < r->len) ? r->line[c] : ' ';
   }
   out[w] = 0;
   r->col += w;
}

static long fr_int(frec *r, int w)
{
   char f[64], g[64];
   int k, n = 0;
   fr_field(r, w, f);
   for (k = 0; f[k]; k++)
      if (f[k] != ' ') g[n++] = f[k];
   g[n] = 0;
   return n ? strtol(g, NULL, 10) : 0;
}

static double fr_real(frec *r, int w, int d)
{
   char f[64], g[64];
   int k, n = 0, has_pt = 0, has_e = 0;
   fr_field(r, w, f);
   for (k = 0; f[k]; k++) {
      if (f[k] == ' ') continue;
      if (f[k] == '.') has_pt = 1;
      if (f[k] == 'E' || f[k] == 'e' || f[k] == 'D' || f[k] == 'd') {
         has_e = 1;
         g[n++] = 'E';
         continue;
      }
      g[n++] = f[k];
   }
   g[n] = 0;
   if (!n) return 0.0;
   if (has_pt || has_e) return strtod(g, NULL);
   return strtod(g, NULL) / pow(10.0, d);
}

static char fr_char(frec *r)
{
   char f[4];
   fr_field(r, 1, f);
   return f[0];
}

/* ------------------------------------------------------------------------ */
static void *xcalloc(size_t n, size_t sz)
{
   void *p = calloc(n ? n : 1, sz);
   if (!p) {
      fprintf(stderr, "sl_basin: out of host memory\n");
      abort();
   }
   return p;
}

static int thousands(char ac, const char *letters)
{
   const char *p = strchr(letters, ac);
   return (ac != ' ' && ac && p) ? (int) (p - letters + 1) * 1000 : 0;
}

/* CRDRD1, crdrd1.f:102-213 */
static int read_header(frec *r, sl_basin *b, char err[256])
{
   char sta[11], ebsn, dol[3], fl[8];
   int k;
   double tlat, tlong;

   fr_next(r);
   fr_field(r, 10, sta);
   b->iver = 199201;
   b->iprj = 0;
   if (!strcmp(sta, "Ver2014-08") || !strcmp(sta, "Ver2019-03")) {
      char proj[16];
      b->iver = !strcmp(sta, "Ver2014-08") ? 201408 : 201903;
      fr_next(r);
      fr_field(r, 11, proj);
      for (k = (int) strlen(proj) - 1; k >= 0 && proj[k] == ' '; k--) proj[k] = 0;
      if (!strcmp(proj, "Proj=GRS80")) b->iprj = 1;
      else if (strcmp(proj, "Proj=Clarke")) {
         snprintf(err, 256, "unrecognized ellipsoid line '%s'", proj);
         return -1;
      }
      fr_next(r);
      fr_field(r, 10, sta);
   }
   memset(b->sta, ' ', 16);
   memcpy(b->sta, sta, 10);
   b->sta[16] = 0;
   r->col = 10;
   fr_char(r);                          /* 1X */
   ebsn = fr_char(r);
   fr_field(r, 2, dol);
   for (k = 0; k < 8; k++) fl[k] = fr_char(r);
   b->ebsn = ebsn;
   b->dollar[0] = dol[0];
   b->dollar[1] = dol[1];
   b->dollar[2] = 0;
   b->ebsn1 = fl[0];
   b->ebsn2 = fl[1];
   b->isouth = fl[2];
   b->nofld = fl[3];
   b->nof1d = fl[4];
   b->iterrain = fl[5];
   b->ismooth = fl[6];
   b->fsouth = (b->isouth == '&') ? -1.0 : 1.0;
   b->hter = (b->iterrain == '+') ? 56.0 : 35.0;
   {
      const char *codes = "+-@#$*%^&=";
      const char *p = (b->ismooth && b->ismooth != ' ') ? strchr(codes, b->ismooth) : NULL;
      b->ismth = p ? (int) (p - codes) + 2 : 1;
   }

   fr_next(r);
   b->azmth = fr_real(r, 10, 6);
   if (ebsn == '$' || ebsn == '+') {
      fr_next(r);
      tlat = fr_real(r, 10, 6);
      tlong = fr_real(r, 10, 6);
      (void) tlat;
      (void) tlong;
   }
   fr_next(r);
   b->alto = fr_real(r, 10, 6);
   b->alno = fr_real(r, 10, 6);
   if (ebsn != '$' && ebsn != '+') {
      fr_next(r);
      b->degree = fr_real(r, 10, 6);
      b->dela = b->degree * RADF;
   }
   if (ebsn == '$') {
      fr_next(r);
      b->dstnt = fr_real(r, 10, 6);
      b->asmpt = fr_real(r, 10, 6);
   }
   if (ebsn == '+') {
      fr_next(r);
      b->aaxis = fr_real(r, 10, 6);
      b->baxis = fr_real(r, 10, 6);
      b->gsize = fr_real(r, 10, 6);
   }
   fr_next(r);
   b->irb = (int) fr_int(r, 5);
   b->ire = (int) fr_int(r, 5);
   b->ndb = (int) fr_int(r, 5);
   b->nde = (int) fr_int(r, 5);
   b->ndb1 = (int) fr_int(r, 5);
   if (ebsn == '$') {
      int ndbz = (b->ndb != 0) ? b->ndb : b->nde;
      if (b->ebsn1 == '$') ndbz = b->ndb1;
      b->degree = (90. - b->asmpt) / ndbz;
      b->dela = b->degree * RADF;
      b->asmpt = b->asmpt * RADF;
      b->aaxis = 0.;
      b->baxis = b->dstnt / cos(b->asmpt);
   }
   if (ebsn == '+') {
      b->dela = 2. * b->gsize / (b->aaxis + b->baxis);
      b->degree = b->dela / RADF;
      b->xmouth = b->aaxis;
      b->ymouth = 0;
   }
   b->phi = 270. + b->azmth;
   b->phi1 = RADF * b->phi;
   b->alt1 = RADF * b->alto;
   b->cor = 2. * (7.292116E-5) * sin(b->alt1);
   b->alt1c = RADF * (90. - b->alto);
   b->aln1 = RADF * b->alno;
   fr_next(r);
   for (k = 0; k < 3; k++) b->dltin[k] = fr_real(r, 10, 6);
   if (ebsn != '+') {
      fr_next(r);
      b->xmouth = fr_real(r, 10, 6);
      b->ymouth = fr_real(r, 10, 6);
   }
   if (ebsn != '$' && ebsn != '+') {
      b->aaxis = b->xmouth;
      b->baxis = b->aaxis;
   }
   b->rmouth = .5 * (b->aaxis + b->baxis);
   b->abqab = (b->aaxis - b->baxis) / (b->aaxis + b->baxis);
   b->imxb = b->irb + b->ire + 1;
   b->jmxb = b->ndb + b->nde + 1;
   b->imxb1 = b->imxb - 1;
   b->jmxb1 = b->jmxb - 1;
   b->imxb2 = b->imxb1 - 1;
   b->jmxb2 = b->jmxb1 - 1;
   if (b->imxb < 4 || b->jmxb < 4 || b->imxb > 999 || b->jmxb > 1655) {
      snprintf(err, 256, "basin dimensions %d x %d out of range", b->imxb, b->jmxb);
      return -1;
   }
   return 0;
}

/* 1-D flow edge arrays at the reference's capacity (parm.for:16 L_, NCT_),
 * with CRDRD2's defaults (crdrd2.f:105-110: no banks, no weir, DELCUT 1) */
static void edges_alloc(sl_basin *b)
{
   int l;
   b->isqr = xcalloc(SL_L1D, sizeof(int));
   b->jsqr = xcalloc(SL_L1D, sizeof(int));
   b->iside = xcalloc(SL_L1D, sizeof(int));
   b->iang = xcalloc(SL_L1D, sizeof(int));
   b->ktree = xcalloc(SL_L1D, 1);
   memset(b->ktree, '0', SL_L1D);
   b->hweir = xcalloc(SL_L1D, sizeof(double));
   b->zbmin = xcalloc(SL_L1D, sizeof(double));
   b->bank1 = xcalloc(SL_L1D, sizeof(double));
   b->bank2 = xcalloc(SL_L1D, sizeof(double));
   b->delcut = xcalloc(SL_L1D, sizeof(double));
   b->cos1d = xcalloc(SL_L1D, sizeof(double));
   b->sin1d = xcalloc(SL_L1D, sizeof(double));
   b->cutl = xcalloc(SL_NCUT, sizeof(double));
   b->cutli = xcalloc(SL_NCUT, sizeof(double));
   b->cutle = xcalloc(SL_NCUT, sizeof(double));
   for (l = 0; l < SL_L1D; l++) b->delcut[l] = 1.;
}

/* CRDRD2, crdrd2.f:65-489 (channels and Manning-n data rejected) */
static int read_body(frec *r, sl_basin *b, char err[256])
{
   int imxb = b->imxb, jmxb = b->jmxb, i, j, k;
   long n;
   long inwp, nsqrs, nnsqr, nweirs, ncut, mgpt, nsheet;
   char ac, bc, ix60;
   size_t np = (size_t) imxb * jmxb;

   b->zb = xcalloc(np, sizeof(double));
   b->zbm = xcalloc(np, sizeof(double));
   b->ylt = xcalloc(np, sizeof(double));
   b->itree = xcalloc(np, 1);
   b->kskp = xcalloc(np, 1);
   b->imann = xcalloc(np, sizeof(int));
   memset(b->itree, ' ', np);                  /* blank CHARACTER (INITALCOMMON) */
   memset(b->kskp, ' ', np);
   for (k = 0; k < (int) np; k++) b->zbm[k] = -300.;

   /* barriers */
   fr_next(r);
   inwp = fr_int(r, 3);
   ac = fr_char(r);
   inwp += thousands(ac, "XYZWVUT");
   if (inwp != 0) {
      int inwx = (int) ((inwp - 1) / 7 + 1), i7x = (int) (inwp - (inwx - 1) * 7), ip;
      for (ip = 1; ip <= inwx; ip++) {
         int i7 = (ip == inwx) ? i7x : 7;
         int w = (b->iver == 199201) ? 3 : 4;
         fr_next(r);
         for (n = 0; n < i7; n++) {
            int ii = (int) fr_int(r, w), jj = (int) fr_int(r, w);
            double hw = fr_real(r, 3, 0);
            if (ii < 1 || ii > imxb || jj < 1 || jj > jmxb) {
               snprintf(err, 256, "barrier index (%d,%d) outside the basin", ii, jj);
               return -1;
            }
            b->zbm[PIX(b, ii, jj)] = hw;
         }
      }
   }
   /* 1-D flow points (crdrd2.f:104-182): I, J, KTREE, side, angle; six per
    * record, I3 fields in 1992 basins and I4 after */
   fr_next(r);
   nsqrs = fr_int(r, 3);
   ac = fr_char(r);
   bc = fr_char(r);
   nnsqr = fr_int(r, 3);
   nsqrs = nnsqr * 1000 + nsqrs + thousands(ac, "XYZWVUTSRQPONMLKJIHGFEDCBAzyxwvutsrqpon");
   if (nsqrs > SL_L1D) {
      snprintf(err, 256, "number of 1-D flow points (%ld) exceeds %d", nsqrs, SL_L1D);
      return -1;
   }
   edges_alloc(b);
   if (nsqrs != 0) {
      int w = (b->iver == 199201) ? 3 : 4;
      for (n = 0; n < nsqrs; n++) {
         if (n % 6 == 0) fr_next(r);
         b->isqr[n] = (int) fr_int(r, w);
         b->jsqr[n] = (int) fr_int(r, w);
         b->ktree[n] = fr_char(r);
         b->iside[n] = (int) fr_int(r, 1);
         b->iang[n] = (int) fr_int(r, 3);
      }
      /* banks (crdrd2.f:185-255): matched to the 1-D points in list order */
      if (bc == 'X') {
         long nbnks;
         int lll = 0;
         fr_next(r);
         nbnks = fr_int(r, 3);
         ac = fr_char(r);
         nbnks += thousands(ac, "XYZWVUTSRQP");
         for (k = 0; k < nbnks; k++) {
            int ib, jb, kb, l;
            double bki, bke, wchn;
            if (k % 4 == 0) fr_next(r);
            ib = (int) fr_int(r, w);
            jb = (int) fr_int(r, w);
            (void) fr_char(r);
            kb = (int) fr_int(r, 1);
            bki = fr_real(r, 3, 0);
            bke = fr_real(r, 3, 0);
            wchn = fr_real(r, 4, 2);
            for (l = lll; l < nsqrs; l++)
               if (b->isqr[l] == ib && b->jsqr[l] == jb && b->iside[l] == kb) {
                  lll = l + 1;
                  b->bank1[l] = bki;
                  b->bank2[l] = bke;
                  b->delcut[l] = wchn;
                  break;
               }
            /* a bank with no matching 1-D point is reported and skipped by the reference */
         }
      }
   }
   /* raised weirs (crdrd2.f:263-276): FORMAT (5(2I3,A1,I1,F5.1)) */
   fr_next(r);
   nweirs = fr_int(r, 3);
   ac = fr_char(r);
   nweirs += thousands(ac, "XYZ");
   if (nsqrs + nweirs > SL_L1D) {
      snprintf(err, 256, "number of 1-D flow points and weirs (%ld) exceeds %d", nsqrs + nweirs, SL_L1D);
      return -1;
   }
   for (n = 0; n < nweirs; n++) {
      long m = nsqrs + n;
      if (n % 5 == 0) fr_next(r);
      b->isqr[m] = (int) fr_int(r, 3);
      b->jsqr[m] = (int) fr_int(r, 3);
      b->ktree[m] = fr_char(r);
      b->iside[m] = (int) fr_int(r, 1);
      b->hweir[m] = fr_real(r, 5, 1);
      b->iang[m] = 0;
   }
   b->nsqrs = (int) nsqrs;
   b->nsqrw = (int) (nsqrs + nweirs);
   /* history stations: names (6A10 -> 2 records), IPN, JPN (15I3) */
   fr_next(r);
   fr_next(r);
   fr_next(r);
   for (k = 0; k < 10; k++) b->ipn[k] = (int) fr_int(r, 3);
   fr_next(r);
   for (k = 0; k < 10; k++) b->jpn[k] = (int) fr_int(r, 3);
   /* channels */
   fr_next(r);
   b->npss = (int) fr_int(r, 3);
   if (b->npss < 0 || b->npss > 5) {
      snprintf(err, 256, "number of channels (%d) outside 0-5", b->npss);
      return -1;
   }
   if (b->npss != 0) {   /* (5F5.1), (5F6.0), (5F5.0), then four (10I3) records */
      int (*pt[4])[2] = {b->ipt0, b->jpt0, b->iptl, b->jptl};
      fr_next(r);
      for (k = 0; k < b->npss; k++) b->chlnh[k] = fr_real(r, 5, 1);
      fr_next(r);
      for (k = 0; k < b->npss; k++) b->chwth[k] = fr_real(r, 6, 0);
      fr_next(r);
      for (k = 0; k < b->npss; k++) b->chdph[k] = fr_real(r, 5, 0);
      for (n = 0; n < 4; n++) {
         fr_next(r);
         for (k = 0; k < b->npss; k++) {
            pt[n][k][0] = (int) fr_int(r, 3);
            pt[n][k][1] = (int) fr_int(r, 3);
         }
      }
      for (k = 0; k < b->npss; k++)
         for (n = 0; n < 2; n++) {
            const int i0 = b->ipt0[k][n], j0 = b->jpt0[k][n], il = b->iptl[k][n], jl = b->jptl[k][n];
            if (i0 < 1 || j0 < 1 || i0 > imxb - 1 || j0 > jmxb - 1 || il < 1 || jl < 1 || il > imxb - 1 ||
                jl > jmxb - 1) {
               snprintf(err, 256, "channel %d end squares (%d,%d) (%d,%d) must lie inside the basin", k + 1,
                        i0, j0, il, jl);
               return -1;
            }
         }
   }
   /* cuts (crdrd2.f:301-357), one per record */
   fr_next(r);
   ncut = fr_int(r, 3);
   ac = fr_char(r);
   ncut += thousands(ac, "XY");
   if (b->nsqrw + ncut > SL_L1D || ncut > SL_NCUT) {
      snprintf(err, 256, "number of 1-D flow points and cuts (%ld) exceeds the reference's limits",
               b->nsqrw + ncut);
      return -1;
   }
   for (n = 0; n < ncut; n++) {
      long m = b->nsqrw + n;
      int w = (b->iver == 199201) ? 3 : 4;
      fr_next(r);
      b->isqr[m] = (int) fr_int(r, w);
      b->jsqr[m] = (int) fr_int(r, w);
      b->ktree[m] = fr_char(r);
      b->iside[m] = (int) fr_int(r, 1);
      b->bank1[m] = fr_real(r, 4, 0);
      b->bank2[m] = fr_real(r, 4, 0);
      b->cutl[n] = fr_real(r, 5, 2);
      b->cutli[n] = fr_real(r, 5, 2);
      b->cutle[n] = fr_real(r, 5, 2);
      b->iang[m] = (int) fr_int(r, 3);
      b->hweir[m] = fr_real(r, 5, 1);
      if (b->cutl[n] == 0.) b->cutl[n] = 1.;
      if (b->cutli[n] == 0.) b->cutli[n] = 1.;
      if (b->cutle[n] == 0.) b->cutle[n] = 1.;
   }
   b->ncut = (int) ncut;
   b->nsqrwc = (int) (b->nsqrw + ncut);
   /* rotation of the wind stress into each edge's flow direction (crdrd2.f:362-371) */
   for (n = 0; n < b->nsqrwc; n++) {
      if (b->iang[n] != 0) {
         double ang = b->iang[n] * RADF;
         b->cos1d[n] = cos(ang);
         b->sin1d[n] = sin(ang);
      } else {
         b->cos1d[n] = 1.;
         b->sin1d[n] = 0.;
      }
   }
   /* the sides must join two squares inside the basin (the reference does not
    * check the exterior square and then reads outside its arrays) */
   for (n = 0; n < b->nsqrwc; n++) {
      const int i = b->isqr[n], j = b->jsqr[n], s = b->iside[n];
      int ii, jj;
      if (s < 1 || s > 4) {
         snprintf(err, 256, "1-D flow edge %ld has side %d (must be 1-4)", (long) n + 1, s);
         return -1;
      }
      ii = i + (s == 1 ? -1 : s == 2 ? 1 : 0);
      jj = j + (s == 3 ? -1 : s == 4 ? 1 : 0);
      if (i < 1 || j < 1 || i > imxb - 1 || j > jmxb - 1 || ii < 1 || jj < 1 || ii > imxb - 1 ||
          jj > jmxb - 1) {
         snprintf(err, 256, "1-D flow edge %ld (%d,%d side %d) touches the basin's outer row or column, "
                  "which this build does not support", (long) n + 1, i, j, s);
         return -1;
      }
      if (n >= b->nsqrs && n < b->nsqrw) {   /* a weir's two end points become momentum barriers */
         const int i1 = i + (s == 2), i2 = i + (s != 1), j1 = j + (s != 3), j2 = j + (s == 4);
         if (i1 < 2 || i2 > imxb - 1 || j1 < 2 || j2 < 2 || j1 > jmxb - 1 || j2 > jmxb - 1) {
            snprintf(err, 256, "weir %ld (%d,%d side %d) ends on the basin's outer row or column, "
                     "which this build does not support", (long) n + 1, i, j, s);
            return -1;
         }
      }
   }
   /* trees (60A1 per J, format reversion every 60 characters) */
   fr_next(r);
   mgpt = fr_int(r, 3);
   if (mgpt == 0) {
      /* without the tree map the I=1 / IMXB barrier ring is not raised and
       * MOMNTM would start at I=1 (momntm.f:132-140 reads HB(I-1,.)) */
      snprintf(err, 256, "basins without a tree map (MGPT=0) are not supported by this build");
      return -1;
   }
   if (mgpt != 0) {
      for (j = 1; j <= jmxb; j++) {
         fr_next(r);
         for (i = 1; i <= imxb; i++) {
            if (i > 1 && (i - 1) % 60 == 0) fr_next(r);
            b->itree[PIX(b, i, j)] = fr_char(r);
         }
      }
      /* Manning's n (times 1000) of every point, Ver2019-03 basins: a count
       * line, then (20I3) records per J (crdrd2.f:379-385) */
      if (b->iver == 201903) {
         fr_next(r);
         (void) fr_int(r, 3);
         for (j = 1; j <= jmxb; j++) {
            fr_next(r);
            for (i = 1; i <= imxb; i++) {
               if (i > 1 && (i - 1) % 20 == 0) fr_next(r);
               b->imann[PIX(b, i, j)] = fr_int(r, 3);
            }
         }
         if (r->eof) {
            snprintf(err, 256, "the Manning-n map ends early");
            return -1;
         }
      }
      if (strcmp(b->dollar, "$ ") && strcmp(b->dollar, "A ")) {
         for (j = 1; j <= jmxb; j += b->jmxb1)
            for (i = 1; i <= imxb; i++) b->zbm[PIX(b, i, j)] = 100.;
      }
      for (i = 1; i <= imxb; i += b->imxb1)
         for (j = 1; j <= jmxb; j++) b->zbm[PIX(b, i, j)] = 100.;
   }
   /* print sheets */
   fr_next(r);
   nsheet = fr_int(r, 3);
   ix60 = fr_char(r);              /* (I3,2A1) into CHARACTER*2 IX60 */
   if (nsheet > 6) nsheet = 6;
   fr_next(r);
   fr_next(r);
   if (ix60 == 'Z') fr_next(r);
   if (ix60 == 'Y' || ix60 == 'Z') fr_next(r);
   /* depths in 15-column strips */
   {
      int imx = (imxb - 1) / 15 + 1, kk, ist, ifn, w = (b->iver == 199201) ? 4 : 5;
      for (kk = 1; kk <= imx; kk++) {
         ist = (kk - 1) * 15 + 1;
         ifn = ist + 14 < imxb ? ist + 14 : imxb;
         for (j = 1; j <= jmxb; j++) {
            fr_next(r);
            if (r->eof) {
               snprintf(err, 256, "dta file ends inside the depth section");
               return -1;
            }
            for (i = ist; i <= ifn; i++) {
               double z = fr_real(r, w, 0);
               if (z > 600. && !b->full_depth) z = 600.;   /* MFULLDEP = 0 (crdrd2.f:441-444) */
               b->zb[PIX(b, i, j)] = z;
            }
         }
      }
   }
   if (b->dollar[0] == 'A') {
      snprintf(err, 256, "basin variant '%s' (the MPI team's island strip) is not supported by this build",
               b->dollar);
      return -1;
   }
   /* '1$': the squares dried at the start (crdrd2.f:447-461): NODRY (I4),
    * then, when more than one, pairs (9(I3,1X,I3,1X)) or (9(I4,1X,I4,1X)) */
   b->nodry = 0;
   if (!strcmp(b->dollar, "1$")) {
      int n, k, w = (b->iver == 199201) ? 3 : 4;
      fr_next(r);
      n = (int) fr_int(r, 4);
      if (r->eof || n > 6000 || n < 0) {
         snprintf(err, 256, "dry cells: %d listed (at most 6000)", n);
         return -1;
      }
      b->nodry = n;
      b->idry = (int *) calloc((size_t) (n > 0 ? n : 1), sizeof(int));
      b->jdry = (int *) calloc((size_t) (n > 0 ? n : 1), sizeof(int));
      if (n > 1) {
         for (k = 0; k < n; k++) {
            if (k % 9 == 0) fr_next(r);
            b->idry[k] = (int) fr_int(r, w);
            (void) fr_char(r);
            b->jdry[k] = (int) fr_int(r, w);
            (void) fr_char(r);
         }
         if (r->eof) {
            snprintf(err, 256, "the dry-cell list ends early");
            return -1;
         }
      }
   }
   if (!strcmp(b->dollar, "$ ")) {
      for (i = 1; i <= imxb; i++) b->zb[PIX(b, i, jmxb)] = b->zb[PIX(b, i, 1)];
   } else {
      for (i = 1; i <= imxb; i++) b->zb[PIX(b, i, jmxb)] = b->zb[PIX(b, i, b->jmxb1)];
   }
   /* '2$' (the MSY basin): DEPMSY after CRDRD2 (initlz.f:52, depmsy.f:8-24)
    * lowers ZB of rows 1..JSUB1 by ZSUB(J), and the lake datum by ZSUBCE;
    * JSUB = 38, JSUB1 = 50, ZSUBCE = -1 (flw1dm_block_c.f:60) */
   b->dtmlak_add = 0.;
   if (!strcmp(b->dollar, "2$")) {
      const int jsub = 38, jsub1 = 50;
      const double zsubce = -1.;
      for (j = 1; j <= jsub1 && j <= jmxb; j++) {
         double zsub = 0.;
         if (j <= jsub) zsub = zsubce;
         else {
            double z = jsub1 - j;
            int izs;
            z = z / (jsub1 - jsub);
            izs = (int) (-zsubce * z * 10.);
            zsub = -izs * .1;
         }
         for (i = 1; i <= b->imxb1; i++) b->zb[PIX(b, i, j)] = b->zb[PIX(b, i, j)] - zsub;
      }
      b->dtmlak_add = zsubce;
   }
   (void) nsheet;
   return 0;
}

/* DEPSFC's 1-D flow branch, depsfc.f:239-372: bank heights from the barrier
 * heights where the dta gives none, weir heights, weir sides made momentum
 * barriers (ITREE '4', ZBM = HTER+1), the weir/sill height HWEIR of every side
 * and ZBMIN, the lower barrier at the ends of the side.  Every edge joins two
 * interior squares (checked when the edges are read), so the reference's
 * out-of-range branches do not arise. */
static void depsfc_edges(sl_basin *b)
{
   static const int IHH[5] = {0, -1, 1, 0, 0}, JHH[5] = {0, 0, 0, -1, 1};
   static const int IZ1[5] = {0, 0, 1, 0, 0}, IZ2[5] = {0, 0, 1, 1, 1};
   static const int JZ1[5] = {0, 1, 1, 0, 1}, JZ2[5] = {0, 0, 0, 0, 1};
   int l, i, j;
#define ZB(i, j) b->zb[PIX(b, i, j)]
#define ZBM(i, j) b->zbm[PIX(b, i, j)]
#define IT(i, j) b->itree[PIX(b, i, j)]
#define MIN4(a, c, d, e) fmin(fmin(fmin(a, c), d), e)
   for (l = 0; l < b->nsqrwc; l++) {
      if (b->bank1[l] == 0.) {
         i = b->isqr[l];
         j = b->jsqr[l];
         b->bank1[l] = MIN4(ZBM(i, j), ZBM(i, j + 1), ZBM(i + 1, j), ZBM(i + 1, j + 1));
      }
      if (b->bank2[l] != 0.) continue;
      i = b->isqr[l] + IHH[b->iside[l]];
      j = b->jsqr[l] + JHH[b->iside[l]];
      b->bank2[l] = MIN4(ZBM(i, j), ZBM(i, j + 1), ZBM(i + 1, j), ZBM(i + 1, j + 1));
   }
   if (b->nsqrw != b->nsqrs) {
      for (l = b->nsqrs; l < b->nsqrw; l++) {
         const int k = b->iside[l];
         i = b->isqr[l];
         j = b->jsqr[l];
         if (b->hweir[l] == 0.) b->hweir[l] = fmin(ZBM(i + IZ1[k], j + JZ1[k]), ZBM(i + IZ2[k], j + JZ2[k]));
      }
      for (l = b->nsqrs; l < b->nsqrw; l++) {
         const int k = b->iside[l];
         i = b->isqr[l];
         j = b->jsqr[l];
         IT(i + IZ1[k], j + JZ1[k]) = 'Q';
         IT(i + IZ2[k], j + JZ2[k]) = 'Q';
      }
   }
   for (j = 2; j <= b->jmxb1; j++)
      for (i = 2; i <= b->imxb1; i++)
         if (IT(i, j) == 'Q') {
            ZBM(i, j) = b->hter + 1;
            IT(i, j) = '4';
         }
   for (l = 0; l < b->nsqrwc; l++) {
      const int k = b->iside[l];
      i = b->isqr[l];
      j = b->jsqr[l];
      if (b->hweir[l] == 0.) b->hweir[l] = fmax(-ZB(i, j), -ZB(i + IHH[k], j + JHH[k]));
      {
         /* IZBX is an INTEGER (implicit typing): the barrier height is truncated */
         const int i1 = i + IZ1[k], j1 = j + JZ1[k], i2 = i + IZ2[k], j2 = j + JZ2[k];
         const int in1 = i1 > 0 && j1 > 0 && i1 < b->imxb && j1 < b->jmxb;
         const int in2 = i2 > 0 && j2 > 0 && i2 < b->imxb && j2 < b->jmxb;
         double z = 0.;
         if (in1 && in2) z = fmin(ZBM(i1, j1), ZBM(i2, j2));
         else if (in1) z = ZBM(i1, j1);
         else if (in2) z = ZBM(i2, j2);
         b->zbmin[l] = (double) (int) z;
      }
   }
#undef ZB
#undef ZBM
#undef IT
#undef MIN4
}

/* DEPSFC, depsfc.f:62-451 */
static int depsfc(sl_basin *b, char err[256])
{
   int imxb = b->imxb, jmxb = b->jmxb, imxb1 = b->imxb1, jmxb1 = b->jmxb1;
   int i, j, ii, j2, n;
   int isl = !strcmp(b->dollar, "$ ");
   double hter = b->hter;
#define ZB(i, j) b->zb[PIX(b, i, j)]
#define ZBM(i, j) b->zbm[PIX(b, i, j)]
#define IT(i, j) b->itree[PIX(b, i, j)]
#define KS(i, j) b->kskp[PIX(b, i, j)]
   b->is = xcalloc(jmxb + 1, sizeof(int));
   b->ie = xcalloc(jmxb + 1, sizeof(int));
   b->ms = xcalloc(jmxb + 1, sizeof(int));
   b->me = xcalloc(jmxb + 1, sizeof(int));
   b->mf = xcalloc(jmxb + 1, sizeof(int));
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++) {
         if (ZB(i, j) > -hter) {
            int kct = 1;
            KS(i, j) = '1';
            if (IT(i, j) == '6') kct++;
            if (IT(i + 1, j) == '6') kct++;
            if (IT(i, j + 1) == '6') kct++;
            if (IT(i + 1, j + 1) == '6') kct++;
            if (kct != 1 && kct < 5) KS(i, j) = '2';
            else if (kct == 5) {
               ZB(i, j) = -67.;
               KS(i, j) = '0';
            }
         } else {
            KS(i, j) = '0';
         }
      }
   for (j = 1; j <= jmxb; j++) KS(imxb, j) = '0';
   for (i = 1; i <= imxb; i++) KS(i, jmxb) = KS(i, 1);
   j2 = isl ? jmxb : jmxb1;
   for (i = 2; i <= imxb; i++)
      for (j = 2; j <= j2; j++) {
         double zmx = -ZB(i - 1, j - 1);
         if (-ZB(i, j - 1) > zmx) zmx = -ZB(i, j - 1);
         if (-ZB(i - 1, j) > zmx) zmx = -ZB(i - 1, j);
         if (-ZB(i, j) > zmx) zmx = -ZB(i, j);
         if (ZBM(i, j) < zmx) ZBM(i, j) = zmx;
      }
   if (isl)
      for (i = 1; i <= imxb; i++) ZBM(i, 1) = ZBM(i, jmxb);
   for (j = 2; j <= jmxb; j++)
      for (i = 2; i <= imxb; i++) {
         if (IT(i, j) == '5' && ZBM(i, j) >= -15.) IT(i, j) = '3';
         if (hter > 35. && IT(i, j) == '4' && ZBM(i, j) < hter) IT(i, j) = '1';
      }
   for (j = 2; j <= j2; j++)
      for (i = 2; i <= imxb1; i++) {
         if (ZBM(i, j) >= hter) IT(i, j) = '4';
         if (IT(i, j) == '5') IT(i, j) = '2';
      }
   for (j = 2; j <= jmxb; j++) {
      for (i = 1; i <= imxb; i++)
         if (ZBM(i, j) < hter) break;
      if (i > imxb) i = imxb;
      b->ms[j - 1] = i;
   }
   b->ms[0] = isl ? b->ms[jmxb - 1] : b->ms[1];
   for (j = 1; j <= jmxb1; j++) {
      for (i = 1; i <= imxb1; i++)
         if (KS(i, j) == '1') break;
      if (i > imxb1) i = imxb1;
      b->is[j - 1] = i;
   }
   b->is[jmxb - 1] = isl ? b->is[0] : b->is[jmxb1 - 1];
   for (j = 2; j <= jmxb1; j++) {
      int found = 0;
      for (ii = 2; ii <= imxb1; ii++) {
         i = imxb - ii + 1;
         if (ZBM(i, j) >= hter) continue;
         if (IT(i, j) != '6') {
            found = 1;
            break;
         }
      }
      if (!found) i = i + 1;                 /* fell through: last I, then I+1 */
      b->me[j - 1] = i;
   }
   b->me[0] = imxb1;
   b->me[jmxb - 1] = imxb1;
   for (j = 1; j <= jmxb1; j++) {
      int found = 0;
      for (ii = 1; ii <= imxb1; ii++) {
         i = imxb1 - ii + 1;
         if (KS(i, j) == '1') {
            found = 1;
            break;
         }
      }
      if (!found) i = i + 1;                 /* fell through: last I, then I+1 */
      b->ie[j - 1] = i;
   }
   b->ie[jmxb - 1] = b->ie[0];
   for (j = 2; j <= jmxb; j++) {
      for (ii = 1; ii <= imxb1; ii++) {
         i = imxb - ii;
         if (IT(i, j) == '1' || IT(i, j) == '3') break;
      }
      b->mf[j - 1] = i;                      /* exhausted: I keeps its last value */
   }
   b->mf[0] = isl ? b->mf[jmxb - 1] : b->mf[1];
   if (b->nsqrwc != 0) depsfc_edges(b);
   /* static-height boundary squares (depsfc.f:376-390), I outer, J inner */
   n = 0;
   b->ish = xcalloc(SL_NBCPTS, sizeof(int));
   b->jsh = xcalloc(SL_NBCPTS, sizeof(int));
   b->tidesh = xcalloc(SL_NBCPTS, sizeof(double));
   for (i = 1; i <= imxb1; i++)
      for (j = 1; j <= jmxb1; j++)
         if (KS(i, j) == '2') {
            n++;
            if (n > SL_NBCPTS) {
               snprintf(err, 256, "number of boundary points exceeds %d", SL_NBCPTS);
               return -1;
            }
            b->ish[n - 1] = i;
            b->jsh[n - 1] = j;
            KS(i, j) = '0';
         }
   b->nbcpt = n;
   if (b->nofld == '+')
      for (j = 2; j <= jmxb1; j++)
         for (i = 2; i <= imxb1; i++)
            if (ZBM(i, j) > 0.) {
               IT(i, j) = '4';
               ZBM(i, j) = hter + 1;
            }
#undef ZB
#undef ZBM
#undef IT
#undef KS
   return 0;
}

/* SETDST, setdst.f:47-87 (ISTAR = JSTAR = 1) */
static void setdst(sl_basin *b)
{
   int i, j;
   b->elpct = xcalloc(b->imxb, sizeof(double));
   b->elpdt = xcalloc(b->imxb, sizeof(double));
   b->elpcl = xcalloc(b->imxb, sizeof(double));
   b->elpdl = xcalloc(b->imxb, sizeof(double));
   b->elpcl2 = xcalloc(b->imxb, sizeof(double));
   b->elpdl2 = xcalloc(b->imxb, sizeof(double));
   b->sint = xcalloc(b->jmxb, sizeof(double));
   b->cost = xcalloc(b->jmxb, sizeof(double));
   b->sinl = xcalloc(b->jmxb, sizeof(double));
   b->cosl = xcalloc(b->jmxb, sizeof(double));
   b->sinl2 = xcalloc(b->jmxb, sizeof(double));
   b->sint2 = xcalloc(b->jmxb, sizeof(double));
   for (i = 1; i <= b->imxb; i++) {
      double x = i - 1 + 1 - 1 - b->irb;
      double rf = exp(x * b->dela);
      double grid = b->rmouth * rf * 5280.;
      double z1 = 1. / (rf * rf);
      b->elpct[i - 1] = (1. + b->abqab * z1) * grid;
      b->elpdt[i - 1] = (1. - b->abqab * z1) * grid;
   }
   for (j = 1; j <= b->jmxb; j++) {
      double x = j - 1 + 1 - 1 - b->ndb;
      double angl = x * b->dela, angh;
      b->cost[j - 1] = cos(angl);
      b->sint[j - 1] = sin(angl);
      b->sint2[j - 1] = b->sint[j - 1] * b->sint[j - 1];
      angh = angl + .5 * b->dela;
      b->cosl[j - 1] = cos(angh);
      b->sinl[j - 1] = sin(angh);
      b->sinl2[j - 1] = b->sinl[j - 1] * b->sinl[j - 1];
   }
   for (i = 1; i <= b->imxb; i++) {
      double x = i - 1 + 1 - 1 - b->irb;
      double rf, grid, z1;
      x = x + 0.5;
      rf = exp(x * b->dela);
      grid = b->rmouth * rf * 5280.;
      z1 = 1. / (rf * rf);
      b->elpcl[i - 1] = (1. + b->abqab * z1) * grid;
      b->elpdl[i - 1] = (1. - b->abqab * z1) * grid;
      b->elpcl2[i - 1] = b->elpcl[i - 1] * b->elpcl[i - 1];
      b->elpdl2[i - 1] = b->elpdl[i - 1] * b->elpdl[i - 1];
   }
   /* areas of the channels' entrance and exit squares, setdst.f:88-107 (the
    * end squares lie inside the basin, checked when they are read) */
   {
      int np, k;
      for (np = 0; np < b->npss; np++) {
         double z1 = 0., z2 = 0.;
         for (k = 0; k < 2; k++) {
            const int i0 = b->ipt0[np][k], j0 = b->jpt0[np][k], il = b->iptl[np][k], jl = b->jptl[np][k];
            z1 = z1 + b->elpdl2[i0 - 1] + (b->elpcl2[i0 - 1] - b->elpdl2[i0 - 1]) * b->sinl2[j0 - 1];
            z2 = z2 + b->elpdl2[il - 1] + (b->elpcl2[il - 1] - b->elpdl2[il - 1]) * b->sinl2[jl - 1];
         }
         b->aresq0[np] = .5 * z1 * b->dela * b->dela;
         b->aresql[np] = .5 * z2 * b->dela * b->dela;
      }
   }
}

/* BTMSTR, btmstr.f:79-214.  The reference's default build takes the
 * COMPLEX_p16 branch: cos/sin/cosh/sinh of double arguments assembled into
 * quad-precision complex values (DCMPLX is double complex = kind 16 under
 * -fdefault-real-8), including its in-place update of CO2_X before CO2_Y. */
typedef _Complex float __attribute__((mode(TC))) cq;
typedef __float128 fq;

static cq mkq(fq re, fq im)
{
   cq z;
   __real__ z = re;
   __imag__ z = im;
   return z;
}

/* Complex quotient by the textbook formula (a+bi)/(c+di) =
 * ((ac+bd) + (bc-ad)i) / (c^2+d^2), which is how the reference's Fortran
 * compiler (LLVM flang) divides COMPLEX(16) values; libgcc's Smith-style
 * division rounds differently in the last bit. */
static cq cdivq(cq x, cq y)
{
   fq a = __real__ x, bb = __imag__ x, c = __real__ y, d = __imag__ y;
   fq den = c * c + d * d;
   return mkq((a * c + bb * d) / den, (bb * c - a * d) / den);
}

static void btmstr(sl_basin *b)
{
   double zlato = b->alto;
   double cor = 2. * (7.292116E-5) * sin(fabs(zlato) * RADF);
   double c25 = .25, c7;
   int ii, n;
   size_t nt = 600 * 245;
   b->ar2 = xcalloc(nt, sizeof(double));
   b->ai2 = xcalloc(nt, sizeof(double));
   b->br2 = xcalloc(nt, sizeof(double));
   b->bi2 = xcalloc(nt, sizeof(double));
   b->cr2 = xcalloc(nt, sizeof(double));
   b->ci2 = xcalloc(nt, sizeof(double));
   c7 = .006;
   for (ii = 1; ii <= 245; ii++) {
      if (b->iver == 201903) c7 = (6 + (ii - 1)) / 1000.;
      else c7 = 0.006;
      for (n = 1; n <= 600; n++) {
         double a, e, cor1, sx, sy, co1x, co1y, co2x, co2y, den, ssx, csx;
         cq co1, co2, co3, co4, co5, cp, co7, co17, co16, co10;
         size_t k = (size_t) (n - 1) * 245 + (ii - 1);
         if (ii == 1) {
            if (b->iver != 201903 && !strncmp(b->sta, "SOUTH FL", 8)) c7 = (n < 31) ? 0.009 : 0.006;
         } else if (ii == 4) {
            c7 = 0.006;
         } else if (ii == 245) {
            if (b->iver != 201903 && !strncmp(b->sta, "SOUTH FL", 8) && n < 57) c7 = 0.25;
         }
         if (ii > 4 && n >= 57) c7 = 0.006;
         a = n;
         e = a * sqrt(cor / (2. * c25));
         cor1 = cor * a / c7;
         sx = -e;
         sy = e;
         /* The reference build evaluates SIN and COS of one argument with a
          * single sincos() call, whose sine differs from sin() in the last
          * bit for some arguments; the tables follow the reference. */
         sincos(sx, &ssx, &csx);
         co1x = csx * cosh(sy);
         co1y = -1 * ssx * sinh(sy);
         co1 = mkq(co1x, co1y);
         co2x = ssx * cosh(sy);
         co2y = csx * sinh(sy);
         den = (co2x * co2x + co2y * co2y);
         co2x = (sx * co2x + sy * co2y) / den;
         co2y = (-1 * sx * co2x + sy * co2y) / den;
         co2 = mkq(co2x, co2y);
         co1 = co1 * co2;
         co3 = cdivq(mkq(1, 0), co1 + mkq(-1., cor1));
         co4 = co3 * co3;
         co5 = co2 * co2 + co1;
         cp = 1 - co2;
         co7 = ((fq) .5 * co5 - 1) * co4;
         co17 = cdivq(mkq(1, 0), 1 + co7);
         co16 = (1 + co3) * co17;
         b->ar2[k] = (double) __real__ co16;
         b->ai2[k] = (double) __imag__ co16;
         b->br2[k] = (double) __real__ co17;
         b->bi2[k] = (double) __imag__ co17;
         co10 = (1 + cp * co3) * co17;
         b->cr2[k] = (double) __real__ co10;
         b->ci2[k] = (double) __imag__ co10;
      }
   }
}

/* CNSTNT, cnstnt.f:36-69 (DELT = DLTIN(1), INTDLT intdlt.f:62) */
static void cnstnt(sl_basin *b)
{
   int l;
   double g = 32.2, expdla, coshla;
   b->delt = b->dltin[0];
   memset(b->x1b, 0, sizeof b->x1b);
   b->x1b[0] = g * b->delt / b->dela;
   b->x1b[1] = b->delt / b->dela;
   b->x1b[3] = b->delt * b->cor;
   for (l = 1; l <= 90; l++)
      b->corl[l - 1] = 2. * (7.292116E-5) * sin((l - 1) * RADF) / b->cor;
   b->x1b[4] = .5 * b->x1b[1];
   b->x1b[5] = .5 * b->x1b[0];
   b->x1b[6] = .6 * b->delt * sqrt(2. * g / 3.) / 3.;
   expdla = exp(b->dela);
   coshla = .5 * (expdla + 1. / expdla);
   b->gama = expdla / coshla - 1.;
   b->gama1 = -b->gama;
   b->gamfp[0] = 0.;
   b->gamfp[2] = 0.;
   b->gamfp[1] = b->gama;
   b->gamfp[3] = b->gama1;
   /* CHANNL's fixed constants, cnstnt.f:62-67 */
   {
      const double cman = 1.5 / .03;
      b->aman = 3. * (cman * cman) / 13.;
      b->bman = 3. * (cman * cman) / (4. * g);
      b->gb = g * b->bman;
      b->agb = b->aman - g * b->bman;
      b->agb13 = 13. * b->agb / 3.;
   }
}

int sl_basin_load(const char *dta_path, sl_basin *b, char err[256])
{
   return sl_basin_load_opts(dta_path, b, 0, err);
}

/* full_depth: the -fullDepth option (INITAL's MFULLDEP, crdrd2.f:441-444): the
 * depths are taken as read instead of capped at 600 ft. */
int sl_basin_load_opts(const char *dta_path, sl_basin *b, int full_depth, char err[256])
{
   frec r;
   memset(b, 0, sizeof *b);
   b->full_depth = full_depth;
   err[0] = 0;
   r.fp = fopen(dta_path, "r");
   r.eof = 0;
   if (!r.fp) {
      snprintf(err, 256, "cannot open basin file %s", dta_path);
      return -1;
   }
   if (read_header(&r, b, err) || read_body(&r, b, err)) {
      fclose(r.fp);
      return -1;
   }
   fclose(r.fp);
   if (depsfc(b, err)) return -1;
   btmstr(b);
   setdst(b);
   cnstnt(b);
   if (b->dltin[0] > 500. || b->dltin[0] < 5.) {
      snprintf(err, 256, "time step %g from the basin file is out of range", b->dltin[0]);
      return -1;
   }
   return 0;
}

void sl_basin_free(sl_basin *b)
{
   void *p[] = {b->zb, b->zbm, b->ylt, b->itree, b->kskp, b->imann, b->is, b->ie, b->ms,
                b->me, b->mf, b->sint, b->cost, b->sinl, b->cosl, b->sinl2, b->sint2,
                b->elpct, b->elpdt, b->elpcl, b->elpdl, b->elpcl2, b->elpdl2, b->ish,
                b->jsh, b->tidesh, b->ar2, b->ai2, b->br2, b->bi2, b->cr2, b->ci2,
                b->isqr, b->jsqr, b->iside, b->iang, b->ktree, b->hweir, b->zbmin, b->bank1,
                b->bank2, b->delcut, b->cos1d, b->sin1d, b->cutl, b->cutli, b->cutle, b->idry, b->jdry};
   size_t k;
   for (k = 0; k < sizeof p / sizeof p[0]; k++) free(p[k]);
   memset(b, 0, sizeof *b);
}

/* ------------------------------------------------------------------------ */
/* Cell latitude (memSetLLx, savellx.c:704-765, helpers savellx.c:40-227).   */
#define LL_RADPDG 0.01745329251994328
#define LL_ECCEN 0.082437112686627
#define LL_ARAD 6.378294E6
#define LL_METPSM 1609.344

void sl_cl2cnf(double clk, double *cnf, double *scale)
{
   double s = sin(clk * LL_RADPDG), t0 = LL_ECCEN * s;
   double t1 = (1. - s) * pow((1. + t0), LL_ECCEN);
   double t2 = (1. + s) * pow((1. - t0), LL_ECCEN);
   *cnf = atan2(t2 - t1, 2. * sqrt(t1 * t2)) / LL_RADPDG;
   *scale = pow((1 - t0 * t0), ((1 + LL_ECCEN) * .5)) * 2. / (t1 + t2);
}

void sl_cnf2cl(double cnf, double *clk)
{
   double s, t0, t1, t2, tl, grad;
   int i;
   *clk = cnf;
   for (i = 0; i < 3; i++) {
      s = sin(LL_RADPDG * (*clk));
      t0 = LL_ECCEN * s;
      t1 = (1. - s) * pow((1. + t0), LL_ECCEN);
      t2 = (1. + s) * pow((1. - t0), LL_ECCEN);
      tl = atan2(t2 - t1, 2 * sqrt(t1 * t2)) / LL_RADPDG;
      grad = 2. * (1. - LL_ECCEN * LL_ECCEN) / (t1 + t2) /
             pow((1. - t0 * t0), (1. - LL_ECCEN / 2.0));
      (*clk) = (*clk) + (cnf - tl) / grad;
   }
}

static double ll_a2lat(const char *src)
{
   char s[16], buf[24], *t, c;
   int deg, mn, sec;
   double ans;
   memcpy(s, src, 12);
   s[12] = 0;
   ans = atof(s);
   snprintf(buf, sizeof buf, "%+09.5f", ans);
   t = strchr(buf, '.');
   *t = 0;
   deg = atoi(buf);
   *t = '.';
   t++;
   c = t[2];
   t[2] = 0;
   mn = atoi(t);
   t[2] = c;
   t += 2;
   c = t[3];
   t[3] = 0;
   sec = atoi(t);
   t[3] = c;
   return (deg < 0) ? deg - mn / 60.0 - sec / (3600.0 * 10.0)
                    : deg + mn / 60.0 + sec / (3600.0 * 10.0);
}

static double ll_fatof(const char *src, unsigned w, int d)
{
   char s[32];
   size_t n = strlen(src);
   if (n > w) n = w;
   memcpy(s, src, n);
   s[n] = 0;
   return strchr(s, '.') ? atof(s) : atof(s) / pow(10, d);
}

int sl_basin_llx(sl_basin *b, const char *line, char err[256])
{
   char s[256];
   double glat, glon, plat, plon, xig, yjg, gsize, ximax, yjmax, elipty = 0, tmp;
   double tscald, stl, ctl, xppt, yppt, rg2p, thtg2p, abq, delrg, t1, t2;
   char type;
   int i, j;
   memset(s, ' ', sizeof s);
   s[sizeof s - 1] = 0;
   memcpy(s, line, strlen(line) < 200 ? strlen(line) : 200);
   type = s[3] == ' ' ? 'p' : s[3];
   glat = ll_a2lat(s + 22);
   glon = ll_a2lat(s + 32);
   plat = ll_a2lat(s + 44);
   plon = ll_a2lat(s + 54);
   xig = ll_fatof(s + 65, 5, 1);
   yjg = ll_fatof(s + 71, 5, 1);
   gsize = ll_fatof(s + 77, 6, 4);
   ximax = ll_fatof(s + 88, 5, 1);
   yjmax = ll_fatof(s + 98, 5, 1);
   if (type == 'h') elipty = ll_fatof(s + 123, 6, 5);
   /* tangent plane at G on the conformal sphere */
   sl_cl2cnf(plat, &plat, &tmp);
   sl_cl2cnf(glat, &glat, &tscald);
   stl = sin(LL_RADPDG * glat);
   ctl = cos(LL_RADPDG * glat);
   {
      /* sll2xy of P */
      double sl = sin(plat * LL_RADPDG), cl = cos(plat * LL_RADPDG);
      double dl = plon - glon, cdl = cos(dl * LL_RADPDG);
      double q1 = 1. + sl * stl + cl * ctl * cdl;
      double q2 = -1 * cl * sin(dl * LL_RADPDG);
      double q3 = sl * ctl - cl * stl * cdl;
      double q4 = ((LL_ARAD / LL_METPSM) * (2. / q1) / tscald);
      xppt = q4 * q2;
      yppt = q4 * q3;
   }
   t1 = xppt - 0.;
   t2 = yppt - 0.;
   rg2p = sqrt(t1 * t1 + t2 * t2);
   thtg2p = 180. + atan2(t1, t2) / LL_RADPDG;
   if (type == 'e') {
      abq = -1;
      delrg = LL_RADPDG * (90. - gsize) / (((int) (yjg + 0.001)) - 1.);
      rg2p = .5 * rg2p / cos(gsize * LL_RADPDG);
      thtg2p = thtg2p + 90.;
      xppt = 0.;
      yppt = 0.;
   } else if (type == 'h') {
      abq = (1. - elipty) / (1. + elipty);
      rg2p = rg2p / (1. + abq);
      if (s[4] == '+') delrg = 360. * LL_RADPDG / (yjmax - ll_fatof(s + 93, 5, 1));
      else delrg = gsize / rg2p;
   } else {
      abq = 0;
      t1 = gsize / (2. * rg2p);
      delrg = 2. * log(sqrt(1.0 + t1 * t1) + t1);
   }
   b->proj.type = type;
   b->proj.glat = glat;
   b->proj.glon = glon;
   b->proj.tscald = tscald;
   b->proj.stl = stl;
   b->proj.ctl = ctl;
   b->proj.xppt = xppt;
   b->proj.yppt = yppt;
   b->proj.rg2p = rg2p;
   b->proj.thtg2p = thtg2p;
   b->proj.abq = abq;
   b->proj.delrg = delrg;
   b->proj.xig = xig;
   b->proj.yjg = yjg;
   b->proj.ximax = ximax;
   b->proj.yjmax = yjmax;
   b->proj.have = 1;
   if ((int) ximax != b->imxb || (int) yjmax != b->jmxb) {
      snprintf(err, 256, "bnt line grid %g x %g does not match the dta grid %d x %d",
               ximax, yjmax, b->imxb, b->jmxb);
      return -1;
   }
   for (j = 0; j < b->jmxb; j++)
      for (i = 0; i < b->imxb; i++) {
         /* pq2xy then sxy2ll then cnf2cl (savellx.c:195-227, 146-183, 68-92) */
         double x, y, lat, lon;
         sl_proj_pq2xy(b, i + 1, j + 1, &x, &y);
         sl_proj_sxy2ll(b, x, y, &lat, &lon);
         sl_cnf2cl(lat, &lat);
         b->ylt[PIX(b, i + 1, j + 1)] = lat;
      }
   b->have_ylt = 1;
   return 0;
}

/* pq2xy, savellx.c:195-227: Z = exp(xa + i ya), Z2 = RG2P (Z + ABQ / Z) */
void sl_proj_pq2xy(const sl_basin *b, double p, double q, double *x, double *y)
{
   const double xa = (p - b->proj.xig) * b->proj.delrg, ya = (b->proj.yjg - q) * b->proj.delrg;
   const double ex = exp(xa), zr = ex * cos(ya), zi = ex * sin(ya);
   const double den = zr * zr + zi * zi, ir = zr / den, ii = -zi / den;
   const double z2r = b->proj.rg2p * (zr + b->proj.abq * ir), z2i = b->proj.rg2p * (zi + b->proj.abq * ii);
   const double x1 = z2i, y1 = z2r;
   const double cs = cos(b->proj.thtg2p * LL_RADPDG), ss = sin(b->proj.thtg2p * LL_RADPDG);
   *x = b->proj.xppt + x1 * cs + y1 * ss;
   *y = b->proj.yppt - x1 * ss + y1 * cs;
}

/* sxy2ll, savellx.c:146-183: plane -> conformal lat/lon */
void sl_proj_sxy2ll(const sl_basin *b, double x, double y, double *lat, double *lon)
{
   double xp = x * b->proj.tscald, yp = y * b->proj.tscald, rr = sqrt(xp * xp + yp * yp);
   double sn, cn, tt, tt2, sp, cp;
   if (rr > 0) {
      sn = xp / rr;
      cn = yp / rr;
   } else {
      sn = 0;
      cn = 1;
   }
   tt = 0.5 * rr * LL_METPSM / LL_ARAD;
   tt2 = tt * tt;
   sp = 2. * tt / (1. + tt2);
   cp = (1. - tt2) / (1. + tt2);
   xp = cp * b->proj.ctl - sp * b->proj.stl * cn;
   yp = -1 * sp * sn;
   rr = sqrt(xp * xp + yp * yp);
   *lon = (rr > 0) ? b->proj.glon + atan2(yp, xp) / LL_RADPDG : b->proj.glon;
   tt = cp * b->proj.stl + sp * b->proj.ctl * cn;
   *lat = atan2(tt, rr) / LL_RADPDG;
}

/* sll2xy, savellx.c:113-129: conformal lat/lon -> plane */
void sl_proj_sll2xy(const sl_basin *b, double lat, double lon, double *x, double *y)
{
   const double sl = sin(lat * LL_RADPDG), cl = cos(lat * LL_RADPDG);
   const double dl = lon - b->proj.glon, cdl = cos(dl * LL_RADPDG);
   const double q1 = 1. + sl * b->proj.stl + cl * b->proj.ctl * cdl;
   const double q2 = -1 * cl * sin(dl * LL_RADPDG);
   const double q3 = sl * b->proj.ctl - cl * b->proj.stl * cdl;
   const double q4 = ((LL_ARAD / LL_METPSM) * (2. / q1) / b->proj.tscald);
   *x = q4 * q2;
   *y = q4 * q3;
}
