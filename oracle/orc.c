/* orc.c -- CPU restatement of one SLOSH time step.  TEST INFRASTRUCTURE ONLY:
 * the checker for the HIP kernels, never linked into the product.
 *
 * Serial, in the reference's loop order, following (file:line under
 * sorc/slosh.fd):
 *   TMSTPC1  tmstpc1.f:55-92   CONTTY contty.f:85-221   SMPT2G smpt2g.f:23-175
 *            FLT2G flt2g.f:38-132   BDRYHT bdryht.f:72-86
 *   TMSTPC2  tmstpc2.f:54-118  HMXSV hmxsv.f:44-136   FILTER filter.f:121-211
 *            XFILTER xfilter.f:176-713
 *   TMSTPM   tmstpm.f:20        MOMNTM momntm.f:78-470  FRCPNT frcpnt.f:63-200
 *            MNTMBD mntmbd.f:58-117
 * TTIMER/STMVAL (the storm side) are supplied by the caller in orc_storm.
 * REAL is float or double; literals are cast to REAL as the reference's
 * default-real constants are.
 */
#include <math.h>
#include <string.h>
#include "orc.h"

typedef REAL R;
#define RC(x) ((R) (x))
/* square root rounded to REAL before it enters further arithmetic: sqrt of a
 * float computed in double and rounded once is the correctly rounded float */
#define SQRT(x) ((R) sqrt((double) (x)))
#define FABS(x) ((R) fabs((double) (x)))
#define F2(a, i, j) (a)[(size_t) ((j) - 1) * imxb + ((i) - 1)]
#define HB(i, j) F2(st->hb, i, j)
#define UB(i, j) F2(st->ub, i, j)
#define VB(i, j) F2(st->vb, i, j)
#define HMX(i, j) F2(st->hmx, i, j)
#define HSUB(i, j) F2(st->hsub, i, j)
#define ZB(i, j) F2(b->zb, i, j)
#define ZBM(i, j) F2(b->zbm, i, j)
#define YLT(i, j) F2(b->ylt, i, j)
#define ITREE(i, j) F2(b->itree, i, j)
#define KSKP(i, j) F2(b->kskp, i, j)
#define IMANN(i, j) F2(b->imann, i, j)
#define IS(j) b->is[(j) - 1]
#define IE(j) b->ie[(j) - 1]
#define MS(j) b->ms[(j) - 1]
#define ME(j) b->me[(j) - 1]
#define ELPCT(i) b->elpct[(i) - 1]
#define ELPDT(i) b->elpdt[(i) - 1]
#define ELPCL(i) b->elpcl[(i) - 1]
#define ELPDL(i) b->elpdl[(i) - 1]
#define ELPCL2(i) b->elpcl2[(i) - 1]
#define ELPDL2(i) b->elpdl2[(i) - 1]
#define SINT(j) b->sint[(j) - 1]
#define COST(j) b->cost[(j) - 1]
#define SINL(j) b->sinl[(j) - 1]
#define COSL(j) b->cosl[(j) - 1]
#define SINL2(j) b->sinl2[(j) - 1]
#define TAB(t, ii, n) b->t[(size_t) ((n) - 1) * 245 + ((ii) - 1)]
#define AMAX1(a, b) ((a) > (b) ? (a) : (b))
#define AMIN1(a, b) ((a) < (b) ? (a) : (b))

/* BLOCK DATA shift tables (flw1dm_block_c.f:37-38) */
static const int IP[5] = {0, 0, 1, 1, 0}, JP[5] = {0, 0, 0, 1, 1};
static const int IH[5] = {0, -1, 0, -1, 0}, JH[5] = {0, -1, -1, 0, 0};
static const int IIH[5] = {0, 0, 1, 0, -1}, JJH[5] = {0, -1, 0, 1, 0};
static const int IHH[5] = {0, -1, 1, 0, 0}, JHH[5] = {0, 0, 0, -1, 1};
/* 1-D flow: sign of the flow across each side (flw1dm_block_c.f:239 SIGA) */
static const REAL SIGA[5] = {0, -1, 1, -1, 1};
#define ISQR(l) b->isqr[(l) - 1]
#define JSQR(l) b->jsqr[(l) - 1]
#define ISIDE(l) b->iside[(l) - 1]
#define KTREE(l) b->ktree[(l) - 1]
#define HWEIR(l) b->hweir[(l) - 1]
#define ZBMIN(l) b->zbmin[(l) - 1]
#define BANK1(l) b->bank1[(l) - 1]
#define BANK2(l) b->bank2[(l) - 1]
#define DELCUT(l) b->delcut[(l) - 1]
#define COS1D(l) b->cos1d[(l) - 1]
#define SIN1D(l) b->sin1d[(l) - 1]
#define CUTL(l) b->cutl[(l) - 1]
#define CUTLI(l) b->cutli[(l) - 1]
#define CUTLE(l) b->cutle[(l) - 1]
#define FLWSQR(l) st->flwsqr[(l) - 1]
#define F1DACT(l) st->f1dact[(l) - 1]

#define M2G_ 30000
#define MCT_ 8000

/* ------------------------------------------------------------------ CONTTY */
void orc_contty(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d)
{
   const int imxb = b->imxb, jmxb = b->jmxb, imxb1 = imxb - 1, jmxb1 = jmxb - 1;
   int i, j, k, ib, ifn, jump, leap, iset = 0, ipass = 0;
   R dux, dvy, hpst, z, fct, hbj, chng, denm, anum, rfct;

   for (j = 1; j <= jmxb; j++)
      for (i = 1; i <= imxb1; i++) HSUB(i, j) = HB(i, j);
   for (j = 2; j <= jmxb1; j++)
      for (i = 2; i <= imxb1; i++)
         if (fabs(UB(i, j)) <= RC(1.E-10)) {
            UB(i, j) = 0;
            VB(i, j) = 0;
         }
   j = 1;
   ib = IS(j);
   ifn = IE(j);
   jump = 4;
   leap = 1;
L130:
   if (ifn >= ib) {
      for (i = ib; i <= ifn; i++) {
         if (KSKP(i, j) == '0') continue;
         if (leap == 2 && i == ipass) {
            HB(i, j) = -ZB(i, j);
            leap = 1;
            continue;
         }
         dux = -UB(i, j) + UB(i + 1, j) + UB(i + 1, j + 1) - UB(i, j + 1);
         if (ZB(i, j) <= RC(30.)) {
            if (dux == 0) continue;
         }
         hpst = HSUB(i, j);
         z = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
         fct = s->x1b5 / z;
         dvy = -VB(i, j) - VB(i + 1, j) + VB(i + 1, j + 1) + VB(i, j + 1);
         hbj = hpst - fct * (dux + dvy);
         chng = hbj + ZB(i, j);
         if (chng < 0) {
            if (-chng < RC(1E-3)) {
               HB(i, j) = -ZB(i, j);
               continue;
            }
         } else {
            HB(i, j) = hbj;
            continue;
         }
         if (HSUB(i, j) == -ZB(i, j)) continue;
         if (d) d->drying_events++;
         denm = hbj - hpst;
         denm = copysign(AMAX1(RC(1.E-5), fabs(denm)), denm);
         anum = -ZB(i, j) - hpst;
         rfct = anum / denm;
         if (fabs(rfct) < RC(1.E-5)) rfct = 0;
         for (k = 1; k <= 4; k++) {
            int ia = i + IP[k], ja = j + JP[k];
            UB(ia, ja) = rfct * UB(ia, ja);
            VB(ia, ja) = rfct * VB(ia, ja);
         }
         j = j - 1;
         jump = 1;
         iset = i;
         leap = 1;
         if (j == 0) goto L330;
         if (KSKP(i, j) == '0') goto L330;
         if (IS(j) > i) goto L330;
         ib = (i - 1 > IS(j)) ? i - 1 : IS(j);
         ifn = (i + 1 < IE(j)) ? i + 1 : IE(j);
         goto L130;
      }
   }
L330:
   jump = jump + 1;
   leap = 1;
   if (jump == 2) {
      j = j + 1;
      leap = 2;
      ib = (iset - 1 > IS(j)) ? iset - 1 : IS(j);
      ifn = IE(j);
      ipass = iset;
   } else {
      j = j + 1;
      if (j > jmxb1) goto L350;
      ib = IS(j);
      ifn = IE(j);
   }
   goto L130;
L350:
   for (i = 1; i <= imxb1; i++) HB(i, jmxb) = HB(i, 1);
}

/* ------------------------------------------------------------ SMPT2G/FLT2G */
/* per thread: several oracle instances may step at once, one per thread */
static _Thread_local short i2g[M2G_ + 1], j2g[M2G_ + 1];
static _Thread_local short mct[MCT_ + 1], i2gg[MCT_ + 1], j2gg[MCT_ + 1];

static void flt2g(const orc_basin *b, orc_state *st, int mm2g, int mcmx)
{
   const int imxb = b->imxb, jmxb = b->jmxb;
   int m, m1, mmm, nnn, i, j, k, ii, jj, kk, ia, ibb, ja, jb;
   R hct, hsm, hdif, gam0, gamf, gamfk, z, zz, zzz, hzz, hbtemp;
   if (mcmx != 0)
      for (m = 1; m <= mcmx; m++) HSUB(i2g[mm2g + m], j2g[mm2g + m]) = 999;
   hct = 1;
   m1 = 1;
   mmm = mm2g;
   nnn = 0;
L3000:
   nnn = nnn + 1;
   if (nnn > 2) return;
   for (m = m1; m <= mmm; m++) {
      i = i2g[m];
      j = j2g[m];
      if (HSUB(i, j) + ZB(i, j) <= hct) continue;
      hsm = 0;
      hdif = 0;
      gam0 = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
      for (k = 1; k <= 4; k++) {
         ii = i + IIH[k];
         jj = j + JJH[k];
         if (ii == 0 || jj == 0) {
            gamf = 0;
            goto L160;
         }
         if (b->ebsn == '$' || b->ebsn == '+') {
            gamfk = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
            z = gamfk / gam0;
            gamf = RC(2.) * z / (RC(1.) + z) - RC(1.);
         } else {
            gamf = b->gamfp[k - 1];
         }
         if (ii == imxb || jj == jmxb) goto L160;
         if (HSUB(ii, jj) == 999) goto L160;
         if (HSUB(ii, jj) + ZB(ii, jj) <= hct) goto L160;
         kk = k % 4 + 1;
         ia = i + IP[k];
         ibb = i + IP[kk];
         ja = j + JP[k];
         jb = j + JP[kk];
         z = ZBM(ia, ja);
         zz = ZBM(ibb, jb);
         zzz = AMIN1(z, zz) + hct;
         if (nnn == 2) {
            if (HSUB(i, j) < zzz && HSUB(ii, jj) < zzz) {
               hsm = hsm + HSUB(i, j);
               hzz = HSUB(i, j);
            } else {
               hsm = hsm + HSUB(ii, jj);
               hzz = HSUB(ii, jj);
            }
         } else {
            if (HSUB(i, j) < zzz || HSUB(ii, jj) < zzz) {
               hsm = hsm + HSUB(i, j);
               hzz = HSUB(i, j);
            } else {
               hsm = hsm + HSUB(ii, jj);
               hzz = HSUB(ii, jj);
            }
         }
         goto L210;
      L160:
         hsm = hsm + HSUB(i, j);
         hzz = HSUB(i, j);
      L210:
         hdif = hdif + hzz * gamf;
      }
      hbtemp = (RC(1.) / RC(3.)) * HSUB(i, j) + (hsm + hdif) / RC(6.);
      HB(i, j) = AMAX1(hbtemp, -ZB(i, j));
   }
   if (mcmx == 0) return;
   for (m = 1; m <= mm2g; m++) HSUB(i2g[m], j2g[m]) = 999;
   hct = 0;
   m1 = mm2g + 1;
   mmm = mm2g + mcmx;
   for (m = m1; m <= mmm; m++) HSUB(i2g[m], j2g[m]) = HB(i2g[m], j2g[m]);
   goto L3000;
}

int orc_smpt2g(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d)
{
   const int imxb = b->imxb, jmxb = b->jmxb, imxb1 = imxb - 1, jmxb1 = jmxb - 1;
   const int imxb2 = imxb - 2, jmxb2 = jmxb - 2;
   int m, mc, mm, mm2g, mcmx, i, j, k, ii, jj;
   (void) s;
   m = 0;
   for (j = 2; j <= jmxb2; j++)
      for (i = 2; i <= imxb2; i++)
         if (UB(i, j) != 0 && ZBM(i, j) > RC(-30.))
            if (UB(i - 1, j) == 0 || UB(i, j - 1) == 0 || UB(i, j + 1) == 0 || UB(i + 1, j) == 0) {
               if (m + 1 >= M2G_) {
                  if (d) d->smooth_overflow = 1;
                  return -1;
               }
               m = m + 1;
               i2g[m] = i;
               j2g[m] = j;
            }
   mm2g = m;
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++) HSUB(i, j) = 999;
   mc = 0;
   for (m = 1; m <= mm2g; m++) {
      ii = i2g[m];
      if (ii > 1 && ii < imxb) {
         jj = j2g[m];
         if (jj > 1 && jj < jmxb)
            for (k = 1; k <= 4; k++) {
               i = ii + IH[k];
               j = jj + JH[k];
               if (HB(i, j) + ZB(i, j) != 0) HSUB(i, j) = HB(i, j);
               else {
                  if (mc + 1 >= MCT_) {
                     if (d) d->smooth_overflow = 1;
                     return -1;
                  }
                  mc = mc + 1;
                  mct[mc] = m;
               }
            }
      }
   }
   mcmx = mc;
   if (mcmx != 0)
      for (m = 1; m <= mcmx; m++) {
         mm = mct[m];
         ii = i2g[mm];
         jj = j2g[mm];
         for (k = 1; k <= 4; k++) {
            i = ii + IH[k];
            j = jj + JH[k];
            if (HB(i, j) + ZB(i, j) != 0) HSUB(i, j) = 888;
         }
      }
   /* sides of 1-D flow not carrying water both ways at the last FLW1DM_C, with
    * either square at most 1 ft above the side's lower barrier: both squares
    * become special corner points (smpt2g.f:101-123) */
   if (b->nsqrw != 0) {
      int l;
      for (l = 1; l <= b->nsqrw; l++)
         if (F1DACT(l) != 'T') {
            i = ISQR(l);
            if (i > 1 && i < imxb1) {
               j = JSQR(l);
               if (j > 1 && j < jmxb1 && HSUB(i, j) != 999) {
                  k = ISIDE(l);
                  ii = i + IHH[k];
                  jj = j + JHH[k];
                  if (HSUB(ii, jj) != 999)
                     if (HSUB(i, j) <= (ZBMIN(l) + RC(1.)) || HSUB(ii, jj) <= (ZBMIN(l) + RC(1.))) {
                        HSUB(i, j) = 888;
                        HSUB(ii, jj) = 888;
                     }
               }
            }
         }
   }
   m = 0;
   mm = 0;
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++)
         if (HSUB(i, j) != 999) {
            if (HSUB(i, j) != 888) {
               if (m + mm >= M2G_) goto L140;
               m = m + 1;
               i2g[m] = i;
               j2g[m] = j;
            } else {
               if (mm >= MCT_) goto L140;
               mm = mm + 1;
               i2gg[mm] = i;
               j2gg[mm] = j;
            }
         }
L140:
   mm2g = m;
   mcmx = mm;
   if (d) {
      d->smooth_pts = mm2g;
      d->smooth_special = mcmx;
   }
   if (mm2g >= M2G_ || mcmx >= MCT_) {
      if (d) d->smooth_overflow = 1;
      return -1;
   }
   for (m = 1; m <= mcmx; m++) {
      i2g[m + mm2g] = i2gg[m];
      j2g[m + mm2g] = j2gg[m];
   }
   flt2g(b, st, mm2g, mcmx);
   return 0;
}

/* ------------------------------------------------------------------ BDRYHT */
void orc_bdryht(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb;
   int n, i, j, k;
   for (n = 0; n < b->nbcpt; n++) {
      R xr, yr, x, y, rsq, r1, r2, dr;
      i = b->ish[n];
      j = b->jsh[n];
      xr = ELPCL(i) * COSL(j);
      yr = ELPDL(i) * SINL(j);
      x = xr - s->c1 - s->ax;
      y = yr - s->c2 - s->ay;
      rsq = x * x + y * y;
      r1 = SQRT(rsq) / RC(5280.) + RC(1.);
      k = (int) r1;
      r2 = k;
      dr = r1 - r2;
      if (k > 790) k = 790;
      HB(i, j) = s->DELP[k] + dr * (s->DELP[k + 1] - s->DELP[k]) + s->seadtm + (st->tidesh ? st->tidesh[n] : RC(0.));
   }
}

/* ------------------------------------------------------------------- HMXSV */
static void seg_bounds(const orc_basin *b, int m, int *jl, int *jm, int *jn, int *il,
                       int *im, int *in)
{
   /* IBB(6,3) of hmxsv.f:49-63 / filter.f:101-115 */
   const int imxb1 = b->imxb - 1, jmxb1 = b->jmxb - 1, imxb2 = imxb1 - 1, jmxb2 = jmxb1 - 1;
   if (m == 1) {
      *jl = 2; *jm = jmxb2; *jn = 1; *il = 2; *im = imxb2; *in = 1;
   } else if (m == 2) {
      *jl = 1; *jm = jmxb1; *jn = jmxb2; *il = 1; *im = imxb1; *in = 1;
   } else {
      *jl = 2; *jm = jmxb2; *jn = 1; *il = 1; *im = imxb1; *in = imxb2;
   }
}

void orc_hmxsv(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb;
   int m, i, j, k, jl, jm, jn, il, im, in;
   (void) s;
   for (m = 1; m <= 3; m++) {
      seg_bounds(b, m, &jl, &jm, &jn, &il, &im, &in);
      for (j = jl; j <= jm; j += jn)
         for (i = il; i <= im; i += in) {
            R hx = HMX(i, j);
            if (hx < HB(i, j)) {
               if (ZB(i, j) > RC(40.) || b->ebsn1 == '&') {
                  HMX(i, j) = HB(i, j);
               } else {
                  R hst = HB(i, j) + ZB(i, j);
                  if (hst < RC(1.)) {
                     HMX(i, j) = HB(i, j);
                  } else {
                     R hsm = 0, hdif = 0, gam0, gamf, hzz, htemp;
                     gam0 = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
                     for (k = 1; k <= 4; k++) {
                        int ii = i + IIH[k], jj = j + JJH[k];
                        if (m != 1 && (ii == 0 || jj == 0)) {
                           gamf = 0;
                           hsm = hsm + HB(i, j);
                           hzz = HB(i, j);
                        } else {
                           R gamfk = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
                           R z = gamfk / gam0;
                           gamf = RC(2.) * z / (RC(1.) + z) - RC(1.);
                           if ((jj != jmxb && ii != imxb) && HB(ii, jj) + ZB(ii, jj) >= RC(1.)) {
                              int kk = k % 4 + 1;
                              int ia = i + IP[k], ibb = i + IP[kk], ja = j + JP[k], jb = j + JP[kk];
                              R zz1 = ZBM(ia, ja), zz2 = ZBM(ibb, jb);
                              R zzz = AMIN1(zz1, zz2) + RC(1.);
                              if (HB(i, j) < zzz || HB(ii, jj) < zzz) {
                                 hsm = hsm + HB(i, j);
                                 hzz = HB(i, j);
                              } else {
                                 hsm = hsm + HB(ii, jj);
                                 hzz = HB(ii, jj);
                              }
                           } else {
                              hsm = hsm + HB(i, j);
                              hzz = HB(i, j);
                           }
                        }
                        hdif = hdif + hzz * gamf;
                     }
                     htemp = (RC(4.) * HB(i, j) + hsm + hdif) / RC(8.);
                     HMX(i, j) = AMAX1(hx, htemp);
                  }
               }
            }
         }
   }
}

/* ------------------------------------------------------------------ FILTER */
void orc_filter(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb, imxb1 = imxb - 1, jmxb1 = jmxb - 1;
   const R hcrt = RC(0.5);
   int m, i, j, k, jl, jm, jn, il, im, in;
   (void) s;
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++) HSUB(i, j) = HB(i, j);
   for (m = 1; m <= 3; m++) {
      seg_bounds(b, m, &jl, &jm, &jn, &il, &im, &in);
      for (j = jl; j <= jm; j += jn)
         for (i = il; i <= im; i += in) {
            R hst, hsm, hdif, gam0, gamf, hzz, htemp;
            if (ZB(i, j) <= -b->hter) continue;
            hst = HSUB(i, j) + ZB(i, j);
            if (hst < hcrt) continue;
            hsm = 0;
            hdif = 0;
            gam0 = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
            for (k = 1; k <= 4; k++) {
               int ii = i + IIH[k], jj = j + JJH[k];
               if (m != 1) {
                  if (ii == 0 || jj == 0) {
                     gamf = 0;
                     hsm = hsm + HSUB(i, j);
                     hzz = HSUB(i, j);
                     hdif = hdif + hzz * gamf;
                     continue;
                  }
               }
               if (b->ebsn == '$' || b->ebsn == '+') {
                  R gamfk = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
                  R z = gamfk / gam0;
                  gamf = RC(2.) * z / (RC(1.) + z) - RC(1.);
               } else {
                  gamf = b->gamfp[k - 1];
               }
               if (ii == imxb || jj == jmxb) {
                  hsm = hsm + HSUB(i, j);
                  hzz = HSUB(i, j);
                  hdif = hdif + hzz * gamf;
                  continue;
               }
               if (HSUB(ii, jj) + ZB(ii, jj) < hcrt) {
                  hsm = hsm + HSUB(i, j);
                  hzz = HSUB(i, j);
                  hdif = hdif + hzz * gamf;
                  continue;
               }
               {
                  int kk = k % 4 + 1;
                  int ia = i + IP[k], ibb = i + IP[kk], ja = j + JP[k], jb = j + JP[kk];
                  R z1 = ZBM(ia, ja), zz = ZBM(ibb, jb), zzz = AMIN1(z1, zz) + hcrt;
                  if (HSUB(i, j) < zzz || HSUB(ii, jj) < zzz) {
                     hsm = hsm + HSUB(i, j);
                     hzz = HSUB(i, j);
                  } else {
                     hsm = hsm + HSUB(ii, jj);
                     hzz = HSUB(ii, jj);
                  }
                  hdif = hdif + hzz * gamf;
               }
            }
            htemp = RC(.5) * HSUB(i, j) + (hsm + hdif) / RC(8.);
            HB(i, j) = AMAX1(-ZB(i, j), htemp);
         }
   }
}

/* ----------------------------------------------------------------- XFILTER */
/* FLTER2, flter2.f:1-102: FILTER's smoother on a closed island basin (DOLLAR
 * '$ ', periodic in J): a neighbour row J-1 = 0 is row JMXB1, row JMXB is a
 * neighbour like any other, and HB(I,JMXB) = HB(I,1) after the pass.  The two
 * segments of the reference's loop (I = 2..IMXB2, then I = 1 and IMXB1) cover
 * every square once and read only HSUB, so one sweep in any order is the same. */
void orc_flter2(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb, imxb1 = imxb - 1, jmxb1 = jmxb - 1;
   const R hcrt = RC(0.5);
   int i, j, k;
   (void) s;
   for (j = 1; j <= jmxb; j++)
      for (i = 1; i <= imxb; i++) HSUB(i, j) = HB(i, j);
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++) {
         R hst, hsm, hdif, gam0, gamf, hzz, htemp;
         if (ZB(i, j) <= -b->hter) continue;
         hst = HSUB(i, j) + ZB(i, j);
         if (hst < hcrt) continue;
         hsm = 0;
         hdif = 0;
         gam0 = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
         for (k = 1; k <= 4; k++) {
            int ii = i + IIH[k], jj = j + JJH[k];
            if (jj == 0) jj = jmxb1;
            if (ii == 0) {
               gamf = 0;
               hsm = hsm + HSUB(i, j);
               hzz = HSUB(i, j);
               hdif = hdif + hzz * gamf;
               continue;
            }
            if (b->ebsn == '$' || b->ebsn == '+') {
               R gamfk = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
               R z = gamfk / gam0;
               gamf = RC(2.) * z / (RC(1.) + z) - RC(1.);
            } else {
               gamf = b->gamfp[k - 1];
            }
            if (ii == imxb || HSUB(ii, jj) + ZB(ii, jj) < hcrt) {
               hsm = hsm + HSUB(i, j);
               hzz = HSUB(i, j);
            } else {
               int kk = k % 4 + 1;
               int ia = i + IP[k], ibb = i + IP[kk], ja = j + JP[k], jb = j + JP[kk];
               R z1 = ZBM(ia, ja), zz = ZBM(ibb, jb), zzz = AMIN1(z1, zz) + hcrt;
               if (HSUB(i, j) < zzz || HSUB(ii, jj) < zzz) {
                  hsm = hsm + HSUB(i, j);
                  hzz = HSUB(i, j);
               } else {
                  hsm = hsm + HSUB(ii, jj);
                  hzz = HSUB(ii, jj);
               }
            }
            hdif = hdif + hzz * gamf;
         }
         htemp = RC(.5) * HSUB(i, j) + (hsm + hdif) / RC(8.);
         HB(i, j) = AMAX1(-ZB(i, j), htemp);
      }
   for (i = 1; i <= imxb; i++) HB(i, jmxb) = HB(i, 1);
}

void orc_xfilter(const orc_basin *b, const orc_storm *s, orc_state *st, int iversion)
{
   const int imxb = b->imxb, jmxb = b->jmxb, imxb1 = imxb - 1, jmxb1 = jmxb - 1;
   static const int IIX[5] = {0, -1, 1, 1, -1}, JJX[5] = {0, -1, -1, 1, 1};
   static const int IP1[5] = {0, -1, 1, 2, 0}, JP1[5] = {0, 0, -1, 1, 2};
   static const int IP2[5] = {0, 0, 2, 1, -1}, JP2[5] = {0, -1, 0, 2, 1};
   R hcrt = 0, exclude = 0;
   int m, i, j, k, jl, jm, jn, il, im, in;
   (void) s;
   switch (b->ismth) {
   case 2: hcrt = RC(0.5); exclude = hcrt; break;
   case 3: hcrt = RC(0.1); exclude = hcrt; break;
   case 4: hcrt = RC(0.2); exclude = hcrt; break;
   case 5: hcrt = RC(0.3); exclude = hcrt; break;
   case 6: hcrt = RC(0.4); exclude = hcrt; break;
   case 7: hcrt = RC(1.E-5); exclude = RC(1.E-5); break;
   case 8: hcrt = RC(0.2); exclude = RC(1.E-5); break;
   case 9: hcrt = RC(0.3); exclude = RC(1.E-5); break;
   case 10: hcrt = RC(0.4); exclude = RC(1.E-5); break;
   case 11: hcrt = RC(0.5); exclude = RC(1.E-5); break;
   }
   for (j = 1; j <= jmxb1; j++)
      for (i = 1; i <= imxb1; i++) HSUB(i, j) = HB(i, j);
   for (m = 1; m <= 3; m++) {
      seg_bounds(b, m, &jl, &jm, &jn, &il, &im, &in);
      for (j = jl; j <= jm; j += jn)
         for (i = il; i <= im; i += in) {
            int maskii[5], mbarri[5], mbouni[5], maskix[5], mbarrx[5], mbounx[5], km;
            const int maskc = 1;
            R hsm2 = 0, xsm2 = 0, hst, htemp;
            if (ZB(i, j) <= -b->hter) continue;
            hst = HSUB(i, j) + ZB(i, j);
            if (hst < exclude) continue;
            for (k = 1; k <= 4; k++) {
               int ii, jj, ix, jx, kk, km1, ia, ibb, ja, jb, ia2, ja2, ia1, ja1, iam1, jam1;
               maskii[k] = mbarri[k] = mbouni[k] = maskix[k] = mbarrx[k] = mbounx[k] = 0;
               ii = i + IIH[k];
               jj = j + JJH[k];
               ix = i + IIX[k];
               jx = j + JJX[k];
               if (ii == 0 || jj == 0 || ix == 0 || jx == 0 || i == 0 || j == 0) continue;
               if (ii == imxb || jj == jmxb || HSUB(ii, jj) + ZB(ii, jj) < hcrt) {
                  mbouni[k] = 1;
               } else {
                  int kk2 = k % 4 + 1;
                  R z = ZBM(i + IP[k], j + JP[k]), zz = ZBM(i + IP[kk2], j + JP[kk2]);
                  R zzz = AMIN1(z, zz) + hcrt;
                  if (HSUB(i, j) < zzz && HSUB(ii, jj) < zzz) mbarri[k] = 1;
                  else maskii[k] = 1;
               }
               kk = k % 4 + 1;
               km1 = (k + 2) % 4 + 1;
               ia = i + IP[k];
               ibb = i + IP[kk];
               ja = j + JP[k];
               jb = j + JP[kk];
               ia2 = i + IP2[k];
               ja2 = j + JP2[k];
               ia1 = i + IP1[k];
               ja1 = j + JP1[k];
               iam1 = i + IP[km1];
               jam1 = j + JP[km1];
               if (ix == imxb || jx == jmxb || HSUB(ix, jx) + ZB(ix, jx) < hcrt) {
                  mbounx[k] = 1;
               } else {
                  R z = ZBM(ia, ja), zz = ZBM(ibb, jb), zzz = AMIN1(z, zz) + hcrt;
                  R yy = ZBM(ia2, ja2), yyy = AMIN1(z, yy) + hcrt;
                  R xx = ZBM(ia1, ja1), xxx = AMIN1(z, xx) + hcrt;
                  R ww = ZBM(iam1, jam1), www = AMIN1(z, ww) + hcrt;
                  if (((HSUB(i, j) >= zzz && HSUB(ii, jj) >= yyy) ||
                       (HSUB(ix, jx) >= yyy && HSUB(ii, jj) >= zzz)) ||
                      ((HSUB(i, j) >= www && HSUB(ia1, ja1) >= xxx) ||
                       (HSUB(ix, jx) >= xxx && HSUB(ia1, ja1) >= www)))
                     maskix[k] = 1;
                  else
                     mbarrx[k] = 1;
               }
            }
            for (km = 1; km <= 4; km++) {
               int kkm = km % 4 + 1, km1m = (km + 2) % 4 + 1;
               int ix = i + IIX[km], jx = j + JJX[km], ii = i + IIH[km], jj = j + JJH[km];
               int ix2 = i + IIX[kkm], jx2 = j + JJX[kkm], ii2 = i + IIH[kkm], jj2 = j + JJH[kkm];
               int ix3 = i + IIX[km1m], jx3 = j + JJX[km1m], ii3 = i + IIH[km1m], jj3 = j + JJH[km1m];
               if (ii == 0 || jj == 0 || ix == 0 || jx == 0 || i == 0 || j == 0 || ix2 == 0 ||
                   jx2 == 0 || ix3 == 0 || jx3 == 0 || ii2 == 0 || jj2 == 0 || ii3 == 0 || jj3 == 0)
                  continue;
               hsm2 = hsm2 + (R) (maskc * maskii[km]) * HSUB(ii, jj) -
                      (R) (maskc * maskii[km]) * HSUB(i, j);
               if (iversion == 1) {
                  R w = (R) (maskix[km] * maskii[km] * maskc * maskix[kkm]);
                  xsm2 = xsm2 + RC(0.5) * (w * HSUB(ix, jx) - w * HSUB(i, j));
               } else {
                  xsm2 = xsm2 + RC(0.5) * (R) maskc * (R) maskix[km] * (R) maskii[km] *
                                (R) maskii[km1m] * (HSUB(ix, jx) - HSUB(i, j));
               }
            }
            htemp = HSUB(i, j) + RC(1.) / RC(8.) * (hsm2 - RC(1.) * (xsm2));
            HB(i, j) = AMAX1(-ZB(i, j), htemp);
         }
   }
}

/* ------------------------------------------------------------------ FRCPNT */
static int iik_of(const orc_basin *b, int i, int j)
{
   const int imxb = b->imxb;
   if (ZB(i, j) >= 0) return (b->iver == 201903) ? 4 : 1;
   if (b->iver == 201903) {
      if (IMANN(i, j) >= 100) {
         int iik = IMANN(i, j) - 6 + 1;
         return iik > 245 ? 245 : iik;
      }
      return 95;
   }
   return !strncmp(b->sta, "SOUTH FL", 8) ? 245 : 1;
}

static void frcpnt(const orc_basin *b, const orc_storm *s, R x, R y, R *fxb, R *fyb, int ncatg,
                   R *cshltr, int npls, int i, int j, R dph, int id)
{
   R xp = x - s->c1 - s->ax, yp = y - s->c2 - s->ay;
   R rsq = xp * xp + yp * yp, rs = SQRT(rsq), r1 = rs / RC(5280.) + RC(1.), r2, dr;
   R ccn, x1218, ck1, sk1, a, bb, fctt, fxx, fyy, aa, pxx, pyy;
   int k = (int) r1, iik = iik_of(b, i, j);
   r2 = k;
   dr = r1 - r2;
   if (k > 790) k = 790;
   ccn = s->x12[15] + rsq;
   x1218 = s->w1218[ncatg - 1];
   if (ncatg == 1) {
      ck1 = s->CW[k] + dr * (s->CW[k + 1] - s->CW[k]);
      sk1 = s->SW[k] + dr * (s->SW[k + 1] - s->SW[k]);
   } else {
      *cshltr = 1;
      ck1 = s->C[k] + dr * (s->C[k + 1] - s->C[k]);
      sk1 = s->S[k] + dr * (s->S[k + 1] - s->S[k]);
   }
   a = rs * s->x12[9] - x1218 * (yp * ck1 + xp * sk1);
   bb = rs * s->x12[10] + x1218 * (xp * ck1 - yp * sk1);
   if (ncatg == 1) {
      R rhol = fabs(xp * s->x12[20] + yp * s->x12[21]);
      R ccc = ccn * rhol / (s->x12[15] + rhol * rhol);
      a = a + ccc * s->x12[23];
      bb = bb + ccc * s->x12[24];
   }
   a = a * s->x12[4] / ccn;
   bb = bb * s->x12[4] / ccn;
   fctt = RC(3.E-6) * SQRT(a * a + bb * bb);
   fctt = fctt * *cshltr * s->bt;
   fxx = fctt * a;
   fyy = fctt * bb;
   *fxb = TAB(cr2, iik, id) * fxx - TAB(ci2, iik, id) * fyy;
   *fyb = TAB(cr2, iik, id) * fyy + TAB(ci2, iik, id) * fxx;
   if (npls != 0) {
      aa = -dph * (s->P[k] + dr * (s->P[k + 1] - s->P[k]));
      pxx = aa * xp;
      pyy = aa * yp;
      *fxb = *fxb + TAB(br2, iik, id) * pxx - TAB(bi2, iik, id) * pyy;
      *fyb = *fyb + TAB(br2, iik, id) * pyy + TAB(bi2, iik, id) * pxx;
   }
}

/* ------------------------------------------------------------------ MNTMBD */
static void mntmbd(const orc_basin *b, const orc_storm *s, orc_state *st, int jump, int ijump,
                   int isb, int jsb, R arj, R aii, int i, int j, R ubj, R vbj, R *fxb, R *fyb,
                   int ncatg, R *cshltr, int *npls, R dph, int id)
{
   const int imxb = b->imxb;
   R aiul = aii * (UB(isb, jsb) - ubj), aivb = aii * (VB(isb, jsb) - vbj);
   R xttx, xtty, fxd, fyd, xr, yr, fxbp, fybp;
   if (jump == 2) {
      xttx = 0;
      xtty = 0;
      UB(isb, jsb) = UB(i, j) + aiul + arj * (VB(isb, jsb) - vbj) + xttx;
      VB(isb, jsb) = VB(i, j) + aivb - arj * (UB(isb, jsb) - UB(i, j)) + xtty;
      return;
   }
   if (ijump != 2) {
      fxd = *fxb;
      fyd = *fyb;
      xr = ELPCT(isb) * COST(jsb);
      yr = ELPDT(isb) * SINT(jsb);
      *npls = 1;
      frcpnt(b, s, xr, yr, &fxbp, &fybp, ncatg, cshltr, *npls, i, j, dph, id);
      xr = ELPDT(isb) * COST(jsb);
      yr = ELPCT(isb) * SINT(jsb);
      *fxb = xr * fxbp + yr * fybp;
      *fyb = xr * fybp - yr * fxbp;
      xttx = s->delt * (*fxb - fxd);
      xtty = s->delt * (*fyb - fyd);
      UB(isb, jsb) = UB(i, j) + aiul + arj * (VB(isb, jsb) - vbj) + xttx;
      VB(isb, jsb) = VB(i, j) + aivb - arj * (UB(isb, jsb) - UB(i, j)) + xtty;
      return;
   }
   *npls = 0;
   *cshltr = 1;
   xr = ELPCT(isb) * COST(jsb);
   yr = ELPDT(isb) * SINT(jsb);
   frcpnt(b, s, xr, yr, &fxbp, &fybp, ncatg, cshltr, *npls, i, j, dph, id);
   xr = ELPDT(isb) * COST(jsb);
   yr = ELPCT(isb) * SINT(jsb);
   *fxb = xr * fxbp + yr * fybp;
   *fyb = xr * fybp - yr * fxbp;
   UB(isb, jsb) = aii * UB(isb, jsb) + arj * VB(isb, jsb) + s->delt * *fxb;
   VB(isb, jsb) = aii * VB(isb, jsb) - arj * UB(isb, jsb) + s->delt * *fyb;
}

/* ------------------------------------------------------------------ MOMNTM */
/* Returns -2 when a point it evaluates lies outside the Coriolis table: LLAT =
 * INT(YLT) must be 0..88 for CORL(LLAT+1) and CORL(LLAT+2) (momntm.f:294-295);
 * beyond it the reference reads outside COMMON /DUMB4I/ and has no defined
 * result. */
int orc_momntm(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb, jmxb1 = jmxb - 1;
   static const int ISBS[5] = {0, 0, -1, 0, 1}, JSBS[5] = {0, -1, 0, 1, 0};
   static const int ISHF[5][4] = {{0}, {0, 1, 2, 4}, {0, 1, 2, 3}, {0, 1, 3, 4}, {0, 2, 3, 4}};
   const R shgt = RC(50.);
   R slpai[601];
   int i, j, jend, iskip, icor = 0, ist, ifn;
   /* persistent across points like the reference's locals and COMMON /MF/ */
   R dph = 0, hpd[5] = {0, 0, 0, 0, 0}, arj = 0, aii = 0, fxb = 0, fyb = 0, cshltr = 0;
   int id = 0, ncatg = 1, npls = 1;

   for (i = 1; i <= 599; i++)
      slpai[i] = (b->iver == 201903) ? TAB(ai2, 4, i + 1) - TAB(ai2, 4, i)
                                     : TAB(ai2, 1, i + 1) - TAB(ai2, 1, i);
   slpai[600] = slpai[599];
   jend = jmxb1;
   if (b->dollar[0] == '$' && b->dollar[1] == ' ') jend = jmxb;
   for (j = 2; j <= jend; j++) {
      iskip = 2;
      if (!(b->dollar[0] == '$' && b->dollar[1] == ' ')) {
         if (j == 2) iskip = 1;
         if (j == jend) iskip = 3;
      }
      ist = MS(j);
      ifn = ME(j);
      if (ifn - ist < 0) continue;
      for (i = ist; i <= ifn; i++) {
         int iik, jump, ijump, kk;
         R dhx = 0, dhy = 0, aintrp, xttx, xtty, aiu, aiv, fct, ubj, vbj;
         iik = iik_of(b, i, j);
         jump = 1;
         if (ITREE(i, j) == '4' || ITREE(i, j) == '6') continue;
         ncatg = (ITREE(i, j) == '2' || ITREE(i, j) == '5') ? 2 : 1;
         if (ncatg != 1 && ITREE(i, j) == '2' && ZBM(i, j) < RC(-50.)) {
            R dphz, dphi;
            jump = 1;
            dhx = -HB(i - 1, j - 1) + HB(i, j - 1) - HB(i - 1, j) + HB(i, j);
            dhy = -HB(i - 1, j - 1) - HB(i, j - 1) + HB(i - 1, j) + HB(i, j);
            dph = RC(.25) * (HB(i - 1, j - 1) + HB(i, j - 1) + HB(i - 1, j) + HB(i, j) +
                             ZB(i - 1, j - 1) + ZB(i, j - 1) + ZB(i - 1, j) + ZB(i, j));
            dphz = AMIN1(RC(599.), dph);
            id = (int) AMIN1(RC(599.), dphz);
            if (id <= 0) id = 1;
            dphi = id;
            aintrp = TAB(ai2, iik, id) + slpai[id] * (dphz - dphi);
            goto L450;
         }
         {
            R hmn = AMIN1(AMIN1(AMIN1(HB(i, j), HB(i - 1, j)), HB(i, j - 1)), HB(i - 1, j - 1));
            if (hmn <= ZBM(i, j)) {
               R hmxx = AMAX1(AMAX1(AMAX1(HB(i, j), HB(i - 1, j)), HB(i, j - 1)), HB(i - 1, j - 1));
               if (hmxx <= ZBM(i, j)) {
                  UB(i, j) = 0;
                  VB(i, j) = 0;
                  goto L500;
               } else {
                  R hf1 = AMAX1(HB(i - 1, j - 1), ZBM(i, j)), hf2 = AMAX1(HB(i, j - 1), ZBM(i, j));
                  R hf3 = AMAX1(HB(i - 1, j), ZBM(i, j)), hf4 = AMAX1(HB(i, j), ZBM(i, j));
                  jump = 2;
                  dhx = -hf1 + hf2 - hf3 + hf4;
                  dhy = -hf1 - hf2 + hf3 + hf4;
               }
            } else {
               jump = 1;
               dhx = -HB(i - 1, j - 1) + HB(i, j - 1) - HB(i - 1, j) + HB(i, j);
               dhy = -HB(i - 1, j - 1) - HB(i, j - 1) + HB(i - 1, j) + HB(i, j);
            }
         }
         {
            R dphz;
            int k;
            hpd[1] = HB(i - 1, j - 1) + ZB(i - 1, j - 1);
            hpd[2] = HB(i, j - 1) + ZB(i, j - 1);
            hpd[3] = HB(i - 1, j) + ZB(i - 1, j);
            hpd[4] = HB(i, j) + ZB(i, j);
            dph = RC(.25) * (hpd[1] + hpd[2] + hpd[3] + hpd[4]);
            dphz = dph;
            id = (int) AMIN1(RC(599.), dphz);
            if (dph <= RC(5.)) {
               if (jump != 2) {
                  for (k = 1; k <= 4; k++) {
                     if (hpd[k] < RC(5.))
                        hpd[k] = hpd[k] * (RC(.4) + RC(.12) * hpd[k] - RC(.008) * hpd[k] * hpd[k]);
                     else
                        hpd[k] = hpd[k] - RC(1.);
                  }
                  dph = RC(.25) * (hpd[1] + hpd[2] + hpd[3] + hpd[4]);
                  dphz = dph;
                  id = (int) AMIN1(RC(599.), dphz);
               }
               if (id < 1) id = 1;
            }
            aintrp = 0;
            for (k = 1; k <= 4; k++) {
               R dz = hpd[k], dphi, aitr;
               int idd = (int) AMIN1(RC(599.), dz);
               if (idd < 1) idd = 1;
               dphi = idd;
               aitr = TAB(ai2, iik, idd) + slpai[idd] * (dz - dphi);
               aintrp = aintrp + aitr;
            }
            aintrp = RC(.25) * aintrp;
         }
      L450:
         aii = RC(1.) + s->x1b4 * aintrp;
         {
            int llat = (int) YLT(i, j);
            R zcor;
            if (llat < 0 || llat > 88) return -2;
            zcor = b->corl[llat] + (YLT(i, j) - llat) * (b->corl[llat + 1] - b->corl[llat]);
            arj = TAB(ar2, iik, id) * s->x1b4 * zcor;
         }
         aiu = aii * UB(i, j);
         aiv = aii * VB(i, j);
         fct = s->x1b6 * dph;
         if (jump != 2) {
            R xr, yr, fxbp, fybp;
            if (ncatg == 2) {
               cshltr = 1;
            } else {
               R hfx[5];
               hfx[1] = HB(i - 1, j - 1) - ZBM(i, j);
               hfx[2] = HB(i, j - 1) - ZBM(i, j);
               hfx[3] = HB(i - 1, j) - ZBM(i, j);
               hfx[4] = HB(i, j) - ZBM(i, j);
               if (ITREE(i, j) != '1') {
                  R c1 = AMIN1(hfx[1], RC(1.)), c2 = AMIN1(hfx[2], RC(1.));
                  R c3 = AMIN1(hfx[3], RC(1.)), c4 = AMIN1(hfx[4], RC(1.));
                  cshltr = RC(.25) * (c1 + c2 + c3 + c4);
               } else {
                  int kx;
                  cshltr = 0;
                  for (kx = 1; kx <= 4; kx++) {
                     R hfxx = 1;
                     if (hfx[kx] <= 0) hfxx = 0;
                     else if (hfx[kx] < shgt) {
                        R q = hfx[kx] / shgt;
                        hfxx = q * q;
                     }
                     cshltr = cshltr + hfxx;
                  }
                  cshltr = RC(.25) * cshltr;
               }
            }
            xr = ELPCT(i) * COST(j);
            yr = ELPDT(i) * SINT(j);
            npls = 1;
            frcpnt(b, s, xr, yr, &fxbp, &fybp, ncatg, &cshltr, npls, i, j, dph, id);
            xr = ELPDT(i) * COST(j);
            yr = ELPCT(i) * SINT(j);
            fxb = (xr * fxbp + yr * fybp);
            fyb = (xr * fybp - yr * fxbp);
            xttx = s->delt * fxb;
            xtty = s->delt * fyb;
         } else {
            xttx = 0;
            xtty = 0;
            fxb = 0;
            fyb = 0;
         }
         ubj = UB(i, j);
         vbj = VB(i, j);
         UB(i, j) = aiu - fct * (TAB(br2, iik, id) * dhx - TAB(bi2, iik, id) * dhy) +
                    arj * VB(i, j) + xttx;
         VB(i, j) = aiv - fct * (TAB(br2, iik, id) * dhy + TAB(bi2, iik, id) * dhx) -
                    arj * UB(i, j) + xtty;
         goto L500b;
      L500:
         ubj = UB(i, j);     /* not read on this path: the boundary is zeroed */
         vbj = VB(i, j);
      L500b:
         ijump = 1;
         if (iskip == 3) {
            if (i == ME(j)) icor = 4;
            if (i != MS(j)) goto L510;
            icor = 3;
            goto L600;
         } else if (iskip != 2) {
            if (i == ME(j)) icor = 2;
            if (i != MS(j)) goto L510;
            icor = 1;
            goto L600;
         }
         if (i != MS(j)) {
            if (i != ME(j)) continue;
            else iskip = 4;
         }
      L510:
         {
            int isb = i + ISBS[iskip], jsb = j + JSBS[iskip];
            if (UB(i, j) == 0) {
               UB(isb, jsb) = 0;
               VB(isb, jsb) = 0;
            } else if (ITREE(isb, jsb) != '6' && ITREE(isb, jsb) != '4') {
               if (ITREE(isb, jsb) == '8') ijump = 2;
               mntmbd(b, s, st, jump, ijump, isb, jsb, arj, aii, i, j, ubj, vbj, &fxb, &fyb,
                      ncatg, &cshltr, &npls, dph, id);
            }
         }
         continue;
      L600:
         for (kk = 1; kk <= 3; kk++) {
            int k = ISHF[icor][kk];
            int isb = i + IH[icor] + IP[k], jsb = j + JH[icor] + JP[k];
            if (UB(i, j) != 0) {
               if (ITREE(isb, jsb) == '6' || ITREE(isb, jsb) == '4') break;
               if (ITREE(isb, jsb) == '8') ijump = 2;
               mntmbd(b, s, st, jump, ijump, isb, jsb, arj, aii, i, j, ubj, vbj, &fxb, &fyb,
                      ncatg, &cshltr, &npls, dph, id);
            } else {
               UB(isb, jsb) = 0;
               VB(isb, jsb) = 0;
            }
         }
      }
   }
   if (b->dollar[0] == '$' && b->dollar[1] == ' ')
      for (i = 1; i <= imxb; i++) {
         UB(i, 1) = UB(i, jmxb);
         VB(i, 1) = VB(i, jmxb);
      }
   (void) hpd;
   return 0;
}

/* ------------------------------------------------------------------ CHANNL */
/* Fractional powers: libm pow in double (the reference's FP64 build calls it
 * for REAL**REAL); in float, the double result rounded. */
#define POWR(x, y) ((R) pow((double) (x), (double) (y)))

/* Channels between distant squares (channl.f:74-186): last step's flow Q
 * fills the exit squares and drains the entrance squares; then the new flow
 * from the head between the two ends: critical depth by Newton (at most 10
 * corrections), critical flow when the low end is below it, else Manning-type
 * subcritical flow; heads under 0.1 ft scale the flow down linearly. */
void orc_channl(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb;
   const R g = RC(32.2), cdely = RC(.1);
   int np, k;
   for (np = 0; np < b->npss; np++) {
      const R qwdt = st->chq[np] * s->delt;
      for (k = 0; k < 2; k++) {
         const int i = b->ipt0[np][k], j = b->jpt0[np][k], ii = b->iptl[np][k], jj = b->jptl[np][k];
         if (i > 0 && j > 0 && i < imxb && j < jmxb) {
            HB(i, j) = HB(i, j) - RC(.5) * qwdt / b->aresq0[np];
            HB(i, j) = AMAX1(HB(i, j), -ZB(i, j));
         }
         if (ii > 0 && jj > 0 && ii < imxb && jj < jmxb) {
            HB(ii, jj) = HB(ii, jj) + RC(.5) * qwdt / b->aresql[np];
            HB(ii, jj) = AMAX1(HB(ii, jj), -ZB(ii, jj));
         }
      }
   }
   for (np = 0; np < b->npss; np++) {
      const R clngth = b->chlnh[np] * RC(5280.), gl = g * clngth;
      R y0 = 0, yl = 0, sig = RC(1.), dely, hit, y, af, af3, ae, hcrit, qflw, q;
      int it = 0;
      for (k = 0; k < 2; k++) {
         const int i = b->ipt0[np][k], j = b->jpt0[np][k];
         if (i > 0 && j > 0 && i < imxb && j < jmxb) y0 = y0 + RC(.5) * HB(i, j);
      }
      y0 = y0 + b->chdph[np];
      for (k = 0; k < 2; k++) {
         const int ii = b->iptl[np][k], jj = b->jptl[np][k];
         if (ii > 0 && jj > 0 && ii < imxb && jj < jmxb) yl = yl + RC(.5) * HB(ii, jj);
      }
      yl = yl + b->chdph[np];
      if (y0 <= yl) {
         const R z = y0;
         y0 = yl;
         yl = z;
         sig = RC(-1.);
      }
      dely = y0 - yl;
      if (dely < cdely) yl = y0 - cdely;
      hit = yl;
      y = y0;
      af = gl + b->gb * POWR(y, RC(4.) / RC(3.));
      af3 = RC(3.) * af;
      ae = b->aman * POWR(y, RC(13.) / RC(3.));
      for (;;) {
         const R hito3 = POWR(hit, RC(1.) / RC(3.)), hit2 = hit * hit, hit3 = hit * hit2;
         const R hit10 = hit3 * hito3, hit13 = hit * hit10;
         const R cnum = af * hit3 + b->agb * hit13 - ae, cdnm = af3 * hit2 + b->agb13 * hit10;
         hcrit = hit - cnum / cdnm;
         if (FABS(hcrit - hit) >= RC(.001)) {
            it = it + 1;
            if (it <= 10) {
               hit = hcrit;
               continue;
            }
         }
         break;
      }
      if (yl <= hcrit) {
         qflw = SQRT(g * hcrit) * hcrit;
      } else {
         const R y0trd = POWR(y0, RC(1.) / RC(3.)), yltrd = POWR(yl, RC(1.) / RC(3.));
         const R y04 = y0 * y0trd, yl4 = yl * yltrd;
         R z = y0 * y0;
         const R y013 = z * z * y0trd;
         z = yl * yl;
         {
            const R yl13 = z * z * yltrd;
            const R qflwsq = b->aman * (y013 - yl13) / (clngth + b->bman * (y04 - yl4));
            qflw = SQRT(FABS(qflwsq));
         }
      }
      q = qflw * sig;
      q = q * b->chwth[np];
      if (dely < cdely) q = q * dely / cdely;
      st->chq[np] = q;
   }
}

/* ---------------------------------------------------------------- FLW1DM_C */
/* Continuity across the open sides of 1-D flow squares, one side at a time in
 * list order (flw1dm_block_c.f:229-430): the side's transport FLWSQR fills one
 * square and drains the other, shaped by banks (DELCUT) or cut widths (CUTL,
 * CUTLI, CUTLE); a square that would go below its bottom is left dry and the
 * transport is cut back.  HSUB of the two squares accumulates the water the
 * bank/cut shaping removed (read by FLW1DM_M as the friction head). */
void orc_flw1dm_c(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb;
   int l;
   for (l = 1; l <= b->nsqrwc; l++) {
      const int i = ISQR(l), j = JSQR(l);
      if (i > 0 && j > 0 && i < imxb && j < jmxb) {
         const int k = ISIDE(l), ii = i + IHH[k], jj = j + JHH[k];
         HSUB(i, j) = 0;
         if (ii > 0 && jj > 0 && ii < imxb && jj < jmxb) HSUB(ii, jj) = 0;
         F1DACT(l) = 'F';
      }
   }
   for (l = 1; l <= b->nsqrwc; l++) {
      int i, j, k, ii, jj, kcut, ll = 0, leap;
      R z, z1, chnghi, chnghe, chng, cstrt, ht, bk, cutw = 0, delb, hbj, hbjj, depli, deple;
      if (FLWSQR(l) == 0) continue;
      i = ISQR(l);
      j = JSQR(l);
      if (!(i > 0 && j > 0 && i < imxb && j < jmxb)) continue;
      k = ISIDE(l);
      ii = i + IHH[k];
      jj = j + JHH[k];
      z = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
      z1 = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
      chnghi = -(SIGA[k] * s->x1b5 * FLWSQR(l) / z);
      chnghe = -(chnghi * z / z1);
      if (l > b->nsqrw) {
         kcut = 2;
         ll = l - b->nsqrw;
         if (CUTLI(ll) >= RC(1.) && CUTLE(ll) >= RC(1.)) {
            chnghi = chnghi * CUTL(ll) / CUTLI(ll);
            chnghe = chnghe * CUTL(ll) / CUTLE(ll);
            goto L470;
         }
      } else {
         kcut = 1;
         if (DELCUT(l) >= RC(1.)) goto L470;
      }
      /* outflow from the interior square (leap 1) or from the exterior (leap 2) */
      if (chnghi > 0) {
         leap = 2;
         goto L194;
      }
      leap = 1;
   L192:
      chng = chnghi;
      cstrt = chnghi;
      ht = HB(i, j);
      bk = BANK1(l);
      if (kcut == 2) cutw = CUTLI(ll);
      if (leap != 2) goto L200;
      goto L300;
   L194:
      chng = chnghe;
      cstrt = chnghe;
      ht = HB(ii, jj);
      bk = BANK2(l);
      if (kcut == 2) cutw = CUTLE(ll);
      if (leap != 2) goto L300;
   L200: /* water leaving a square */
      if (ht <= bk) {
         if (kcut == 2) chng = chng * CUTL(ll) / cutw;
      } else {
         delb = ht - bk;
         if (kcut != 2) {
            if (-chng * DELCUT(l) <= delb) chng = chng * DELCUT(l);
            else chng = chng - delb + delb / DELCUT(l);
         } else {
            if (-chng * CUTL(ll) <= delb) chng = chng * CUTL(ll);
            else chng = (chng * CUTL(ll) + delb) / cutw - delb;
         }
      }
      if (leap != 2) {
         chnghi = chng;
         HSUB(i, j) = HSUB(i, j) + FABS(cstrt - chng);
         goto L194;
      }
      chnghe = chng;
      HSUB(ii, jj) = HSUB(ii, jj) + FABS(cstrt - chng);
      goto L192;
   L300: /* water entering a square */
      if (ht >= bk) {
         if (kcut != 2) chng = chng * DELCUT(l);
         else chng = chng * CUTL(ll);
      } else {
         delb = bk - ht;
         if (kcut != 2) {
            if (!(chng <= delb)) chng = delb + (chng - delb) * DELCUT(l);
         } else {
            if (chng * CUTL(ll) <= delb * cutw) chng = chng * CUTL(ll) / cutw;
            else chng = chng * CUTL(ll) - delb * cutw + delb;
         }
      }
      if (leap != 2) {
         chnghe = chng;
         HSUB(ii, jj) = HSUB(ii, jj) + FABS(cstrt - chng);
      } else {
         chnghi = chng;
         HSUB(i, j) = HSUB(i, j) + FABS(cstrt - chng);
      }
   L470:
      hbj = HB(i, j) + chnghi;
      hbjj = HB(ii, jj) + chnghe;
      depli = hbj + ZB(i, j);
      deple = hbjj + ZB(ii, jj);
      if (depli < 0) {
         const R cii = chnghi - depli, sube = -(cii * z / z1);
         HB(i, j) = -ZB(i, j);
         HB(ii, jj) = AMAX1(HB(ii, jj) - chnghe + sube, -ZB(ii, jj));
         FLWSQR(l) = FLWSQR(l) * cii / chnghi;
      } else if (deple < 0) {
         const R cee = chnghe - deple, subi = -(cee * z1 / z);
         HB(i, j) = AMAX1(HB(i, j) - chnghi + subi, -ZB(i, j));
         HB(ii, jj) = -ZB(ii, jj);
         FLWSQR(l) = FLWSQR(l) * cee / chnghe;
      } else {
         HB(i, j) = hbj;
         HB(ii, jj) = hbjj;
         F1DACT(l) = 'T';
      }
   }
}

/* ---------------------------------------------------------------- FLW1DM_M */
static R dim_(R a, R c)
{
   return a > c ? a - c : RC(0.);
}

/* Momentum across the open sides (flw1dm_block_m.f:153-397): the head between
 * the two squares, clamped between the weir height HWEIR and the lower barrier
 * ZBMIN, drives FLWSQR against bottom friction; below the barrier the lake
 * wind stress at the side's force point (FRCPNT, rotated into the side's flow
 * direction) adds on; the friction head HSUB left by FLW1DM_C damps the
 * result by a 20-step fixed point. */
void orc_flw1dm_m(const orc_basin *b, const orc_storm *s, orc_state *st)
{
   const int imxb = b->imxb, jmxb = b->jmxb;
   const R slpa = TAB(ai2, 1, 2) - TAB(ai2, 1, 1);
   const int ncatg = 1, npls = 1, sfl = !strncmp(b->sta, "SOUTH FL", 8);
   int l;
   for (l = 1; l <= b->nsqrwc; l++) {
      const int i = ISQR(l), j = JSQR(l);
      int k, ii, jj, iik, jump, id, kk;
      R hint, hext, zl, hovt, hint1, hext1, head, dph1d[3], dph, biobr, hcoff, flwt, frc12t;
      R aintrp, flwcof, flw, fwd;
      if (!(i > 0 && j > 0 && i < imxb && j < jmxb)) continue;
      iik = iik_of(b, i, j);
      hint = HB(i, j);
      k = ISIDE(l);
      ii = i + IHH[k];
      jj = j + JHH[k];
      if (ii <= 0 || jj <= 0 || ii >= imxb + 1 || jj >= jmxb + 1) {
         FLWSQR(l) = 0;
         HSUB(i, j) = 0;
         continue;
      }
      hext = HB(ii, jj);
      zl = ZBMIN(l);
      if (AMIN1(hint, hext) >= zl || (hint == -ZB(i, j) && hext == -ZB(ii, jj))) {
         FLWSQR(l) = 0;
         HSUB(i, j) = 0;
         HSUB(ii, jj) = 0;
         continue;
      }
      hovt = HWEIR(l);
      if (AMAX1(hint, hext) <= hovt) {
         FLWSQR(l) = 0;
         HSUB(i, j) = 0;
         HSUB(ii, jj) = 0;
         continue;
      }
      hint1 = HB(i, j);
      hext1 = HB(ii, jj);
      hint = AMIN1(AMAX1(hint1, hovt), zl);
      hext = AMIN1(AMAX1(hext1, hovt), zl);
      jump = 2;
      if (AMAX1(hint, hext) >= zl) jump = 1;
      head = hext - hint;
      head = SIGA[k] * head;
      dph1d[1] = hint + ZB(i, j);
      dph1d[2] = hext + ZB(ii, jj);
      dph = RC(.5) * (dph1d[1] + dph1d[2]);
      if (dph < RC(0.00000001)) {
         FLWSQR(l) = 0;
         HSUB(i, j) = 0;
         HSUB(ii, jj) = 0;
         continue;
      }
      id = (int) AMIN1(RC(599.), dph);
      if (id < 1) id = 1;
      biobr = TAB(bi2, iik, id) / TAB(br2, iik, id);
      if (b->nsqrw != b->nsqrs && l > b->nsqrs && l <= b->nsqrw) biobr = 0;
      hcoff = TAB(br2, iik, id) + TAB(bi2, iik, id) * biobr;
      flwt = -(s->x1b1 * hcoff * dph * head);
      if (jump != 1) {
         R cshltr, xr0, yr0, elpcz, elpdz, xr, yr, fxb, fyb, fxbp, fybp, frc1, frc2;
         int iii = i, jjj = j;
         if (KTREE(l) == 'T') {
            const R q = dph / RC(50.);
            cshltr = q * q;
         } else {
            const R cshint = RC(1.) - dim_(RC(1.), dim_(hint1, hovt));
            const R cshext = RC(1.) - dim_(RC(1.), dim_(hext1, hovt));
            const R csh2i = RC(1.) - dim_(RC(1.), dim_(zl, hint1));
            const R csh2e = RC(1.) - dim_(RC(1.), dim_(zl, hext1));
            cshltr = RC(.5) * AMIN1(cshint + cshext, csh2i + csh2e);
         }
         if (k != 3 && k != 4) {   /* force point on the side (polar grid (x,y)) */
            if (k == 2) iii = i + 1;
            xr0 = COSL(jjj);
            yr0 = SINL(jjj);
            elpcz = ELPCT(iii);
            elpdz = ELPDT(iii);
         } else {
            if (k == 4) jjj = j + 1;
            xr0 = COST(jjj);
            yr0 = SINT(jjj);
            elpcz = ELPCL(iii);
            elpdz = ELPDL(iii);
         }
         xr = xr0 * elpcz;
         yr = yr0 * elpdz;
         frcpnt(b, s, xr, yr, &fxb, &fyb, ncatg, &cshltr, npls, i, j, dph, id);
         xr = xr0 * elpdz;
         yr = yr0 * elpcz;
         fxbp = xr * fxb + yr * fyb;
         fybp = xr * fyb - yr * fxb;
         if (COS1D(l) != RC(1.)) {
            fxb = COS1D(l) * fxbp - SIN1D(l) * fybp;
            fyb = SIN1D(l) * fxbp + COS1D(l) * fybp;
         } else {
            fxb = fxbp;
            fyb = fybp;
         }
         frc1 = RC(.75) * fxb;
         frc2 = RC(.75) * fyb;
         if (k == 3 || k == 4) {
            const R t = frc2;
            frc2 = -frc1;
            frc1 = t;
         }
         frc12t = s->delt * (frc1 + biobr * frc2);
      } else {
         frc12t = 0;
      }
      aintrp = 0;
      for (kk = 1; kk <= 2; kk++) {
         const int idd = (int) AMIN1(RC(299.), dph1d[kk]);
         R aitr;
         if (idd <= 1) {
            if (b->iver < 201903 && sfl) aitr = TAB(ai2, 1, 1) + slpa * (dph1d[kk] - RC(1.));
            else aitr = TAB(ai2, iik, 1) + slpa * (dph1d[kk] - RC(1.));
         } else {
            const R slpai = TAB(ai2, iik, idd + 1) - TAB(ai2, iik, idd), dphi = (R) idd;
            aitr = TAB(ai2, iik, idd) + slpai * (dph1d[kk] - dphi);
         }
         aintrp = aintrp + aitr;
      }
      aintrp = RC(.5) * aintrp;
      flwcof = s->x1b4 * (aintrp - TAB(ar2, iik, id) * biobr) + RC(1.);
      flw = flwt + flwcof * FLWSQR(l) + frc12t;
      fwd = HSUB(ii, jj) - HSUB(i, j);
      if (fwd != 0) {
         R fd2 = -(SIGA[k] * fwd * s->x1b2 / (dph * dph)), z, z1, xy, flw1;
         int itr;
         fd2 = FABS(fd2);
         z = ELPDL2(i) + (ELPCL2(i) - ELPDL2(i)) * SINL2(j);
         z1 = ELPDL2(ii) + (ELPCL2(ii) - ELPDL2(ii)) * SINL2(jj);
         xy = RC(.5) * (RC(1.) / z + RC(1.) / z1);
         fd2 = fd2 * xy;
         flw1 = flw;
         for (itr = 1; itr <= 20; itr++) flw1 = flw / (RC(1.) + fd2 * FABS(flw1));
         FLWSQR(l) = flw1;
      } else {
         FLWSQR(l) = flw;
      }
   }
}

/* -------------------------------------------------------------------- step */
int orc_step(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d, REAL *hbs)
{
   int incsm = 20, ismpt2g = 1;
   char e1 = b->ebsn1;
   if (d) memset(d, 0, sizeof *d);
   if (e1 == '+' || e1 == '=') incsm = 1;
   if (e1 == '!' || e1 == '#') incsm = 5;
   if (e1 == '@' || e1 == '%') incsm = 10;
   if (e1 == '&') incsm = 20;
   if (e1 == '^' || e1 == '*') incsm = 30;
   if (e1 == '+' || e1 == '&' || e1 == '#' || e1 == '%' || e1 == '*') ismpt2g = 0;
   /* TMSTPC1 */
   orc_contty(b, s, st, d);
   if (ismpt2g != 0 && s->itime > 5 && incsm != 1)
      if (orc_smpt2g(b, s, st, d)) return -1;
   orc_bdryht(b, s, st);
   if (b->npss != 0) orc_channl(b, s, st);
   if (b->nsqrwc != 0 && b->nof1d != '+') orc_flw1dm_c(b, s, st);
   /* TMSTPC2 */
   if (b->nsqrwc != 0 && b->nof1d != '+') orc_flw1dm_m(b, s, st);
   orc_hmxsv(b, s, st);
   /* HB after HMXSV: what TMSTPC2 copies to the C driver's HHB (tmstpc2.f:70-85) */
   if (hbs) memcpy(hbs, st->hb, sizeof(REAL) * (size_t) b->imxb * b->jmxb);
   if ((s->itime + 1) % incsm == 0 || e1 == '+') {
      int isl = (b->dollar[0] == '$' && b->dollar[1] == ' ');
      if (d) d->filtered = 1;
      if (isl) orc_flter2(b, s, st);            /* closed island (tmstpc2.f:101-111) */
      else if (b->ismth == 1) orc_filter(b, s, st);
      else orc_xfilter(b, s, st, strncmp(b->sta, "CH2 (2008)", 10) == 0 ? 1 : 2);
   }
   /* TMSTPM */
   return orc_momntm(b, s, st);
}
