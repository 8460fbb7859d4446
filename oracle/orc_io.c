/* orc_io.c -- flat C entry points around the step restatement, for tests.
 * TEST INFRASTRUCTURE ONLY (part of the oracle).
 *
 * A context is opened from a tagged-record "init" file (the format written by
 * oracle/ref_driver.c from the reference's COMMON blocks, and by the product's
 * slg_export_init from its own host initialisation).  Values are read as double
 * and cast to REAL.  Each step takes the storm row that stood after the
 * previous step (ref_driver.c dump_storm layout) and the current radial tables.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

typedef REAL R;

typedef struct {
   orc_basin b;
   orc_state st;
   R *zb, *zbm, *ylt, *hb, *ub, *vb, *hmx, *hsub;
   char *itree, *kskp;
   int *imann, *is, *ie, *ms, *me, *ish, *jsh;
   R *elp[6], *trig[5], *tab[6];
   int *isqr, *jsqr, *iside;
   char *ktree, *f1dact;
   R *edge[10], *flwsqr;
   R S[801], C[801], P[801], DELP[801], SW[801], CW[801];
   R seadtm;
   int ic19;
   orc_diag diag;
} ctx_t;

typedef struct {
   char name[17];
   char type;
   long n;
   void *data;
} recd;

static int read_recs(const char *path, recd **out)
{
   FILE *fp = fopen(path, "rb");
   int n = 0, cap = 0;
   recd *r = NULL;
   char hdr[20];
   if (!fp) return -1;
   while (fread(hdr, 1, 20, fp) == 20) {
      long cnt;
      int esz;
      if (fread(&cnt, 8, 1, fp) != 1) break;
      if (n == cap) {
         cap = cap ? 2 * cap : 64;
         r = realloc(r, cap * sizeof *r);
      }
      memcpy(r[n].name, hdr, 16);
      r[n].name[16] = 0;
      r[n].type = hdr[16];
      r[n].n = cnt;
      esz = (r[n].type == 'd') ? 8 : (r[n].type == 'i') ? 4 : (r[n].type == 'h') ? 2 : 1;
      r[n].data = malloc(cnt * esz + 1);
      if (fread(r[n].data, esz, cnt, fp) != (size_t) cnt) break;
      n++;
   }
   fclose(fp);
   *out = r;
   return n;
}

static recd *find(recd *r, int n, const char *name)
{
   int k;
   for (k = 0; k < n; k++)
      if (!strcmp(r[k].name, name)) return &r[k];
   return NULL;
}

static R *toR(recd *x)
{
   R *o = malloc(sizeof(R) * (x->n ? x->n : 1));
   long k;
   for (k = 0; k < x->n; k++) o[k] = (R) ((double *) x->data)[k];
   return o;
}

static int *toI(recd *x)
{
   int *o = malloc(sizeof(int) * (x->n ? x->n : 1));
   memcpy(o, x->data, sizeof(int) * x->n);
   return o;
}

/* an optional record as REAL (zero-filled to at least n entries) */
static R *toR_opt(recd *x, long n)
{
   R *o = calloc((size_t) (n > 0 ? n : 1), sizeof(R));
   long k;
   if (x)
      for (k = 0; k < x->n && k < n; k++) o[k] = (R) ((double *) x->data)[k];
   return o;
}

static int *toI_opt(recd *x, long n)
{
   int *o = calloc((size_t) (n > 0 ? n : 1), sizeof(int));
   if (x) memcpy(o, x->data, sizeof(int) * (size_t) (x->n < n ? x->n : n));
   return o;
}

static void load_tables(ctx_t *c, recd *r, int n)
{
   const char *nm[6] = {"S", "C", "P", "DELP", "SW", "CW"};
   R *dst[6] = {c->S, c->C, c->P, c->DELP, c->SW, c->CW};
   int t, k;
   for (t = 0; t < 6; t++) {
      recd *x = find(r, n, nm[t]);
      if (!x) continue;
      for (k = 0; k < 800; k++) dst[t][k + 1] = (R) ((double *) x->data)[k];
   }
}

void *orc_open(const char *path)
{
   recd *r;
   int n = read_recs(path, &r), k;
   ctx_t *c;
   int *d3, *ints;
   const char *vnames[6] = {"ELPCT", "ELPDT", "ELPCL", "ELPDL", "ELPCL2", "ELPDL2"};
   const char *tnames[6] = {"AR2", "AI2", "BR2", "BI2", "CR2", "CI2"};
   const char *gnames[5] = {"SINT", "COST", "SINL", "COSL", "SINL2"};
   size_t np;
   if (n <= 0 || !find(r, n, "DUMB3")) return NULL;
   c = calloc(1, sizeof *c);
   d3 = (int *) find(r, n, "DUMB3")->data;
   c->b.imxb = d3[0];
   c->b.jmxb = d3[1];
   np = (size_t) d3[0] * d3[1];
   c->zb = toR(find(r, n, "ZB"));
   c->zbm = toR(find(r, n, "ZBM"));
   c->ylt = toR(find(r, n, "YLT"));
   c->hb = toR(find(r, n, "HB"));
   c->ub = toR(find(r, n, "UB"));
   c->vb = toR(find(r, n, "VB"));
   c->hmx = toR(find(r, n, "HMX"));
   c->hsub = calloc(np, sizeof(R));
   c->itree = malloc(np);
   memcpy(c->itree, find(r, n, "ITREE")->data, np);
   c->kskp = malloc(np);
   memcpy(c->kskp, find(r, n, "KSKP")->data, np);
   c->imann = toI(find(r, n, "IMANN"));
   c->is = toI(find(r, n, "IS"));
   c->ie = toI(find(r, n, "IE"));
   c->ms = toI(find(r, n, "MS"));
   c->me = toI(find(r, n, "ME"));
   c->ish = toI(find(r, n, "ISH"));
   c->jsh = toI(find(r, n, "JSH"));
   for (k = 0; k < 6; k++) c->elp[k] = toR(find(r, n, vnames[k]));
   for (k = 0; k < 5; k++) c->trig[k] = toR(find(r, n, gnames[k]));
   for (k = 0; k < 6; k++) c->tab[k] = toR(find(r, n, tnames[k]));
   ints = (int *) find(r, n, "INTS")->data;
   c->ic19 = ints[1];
   c->b.iver = ints[4];
   c->b.ismth = ints[6];
   {
      char *fl = (char *) find(r, n, "FLAGS")->data;
      c->b.dollar[0] = fl[0];
      c->b.dollar[1] = fl[1];
      c->b.ebsn = fl[2];
      c->b.ebsn1 = fl[3];
      c->b.nof1d = fl[5];
   }
   {
      /* sub-grid 1-D flow edges (absent from files written before they existed) */
      const char *en[10] = {"HWEIR", "ZBMIN", "BANK1", "BANK2", "DELCUT", "COS1D", "SIN1D",
                            "CUTL", "CUTLI", "CUTLE"};
      long nl = ints[18], nc = ints[20];
      recd *kt = find(r, n, "KTREE");
      c->b.nsqrs = ints[16];
      c->b.nsqrw = ints[17];
      c->b.nsqrwc = ints[18];
      c->b.ncut = ints[20];
      c->isqr = toI_opt(find(r, n, "EDGE_I"), nl);
      c->jsqr = toI_opt(find(r, n, "EDGE_J"), nl);
      c->iside = toI_opt(find(r, n, "EDGE_SIDE"), nl);
      c->ktree = calloc((size_t) nl + 1, 1);
      if (kt) memcpy(c->ktree, kt->data, (size_t) (kt->n < nl ? kt->n : nl));
      for (k = 0; k < 10; k++) c->edge[k] = toR_opt(find(r, n, en[k]), k < 7 ? nl : nc);
      c->flwsqr = calloc((size_t) nl + 1, sizeof(R));
      c->f1dact = malloc((size_t) nl + 1);
      memset(c->f1dact, '0', (size_t) nl + 1);
      c->b.isqr = c->isqr;
      c->b.jsqr = c->jsqr;
      c->b.iside = c->iside;
      c->b.ktree = c->ktree;
      c->b.hweir = c->edge[0];
      c->b.zbmin = c->edge[1];
      c->b.bank1 = c->edge[2];
      c->b.bank2 = c->edge[3];
      c->b.delcut = c->edge[4];
      c->b.cos1d = c->edge[5];
      c->b.sin1d = c->edge[6];
      c->b.cutl = c->edge[7];
      c->b.cutli = c->edge[8];
      c->b.cutle = c->edge[9];
      c->st.flwsqr = c->flwsqr;
      c->st.f1dact = c->f1dact;
   }
   {
      /* channels (absent from files written before they existed) */
      recd *pt = find(r, n, "CHNL_PTS"), *cd = find(r, n, "CHNLD"), *ca = find(r, n, "CHNARA");
      recd *cn = find(r, n, "CHNLNO");
      c->b.npss = ints[19];
      if (pt && cd && ca && cn) {
         const int *p = (const int *) pt->data;
         const double *d = (const double *) cd->data, *a = (const double *) ca->data, *q = (const double *) cn->data;
         int e;
         for (k = 0; k < 5; k++) {
            for (e = 0; e < 2; e++) {
               c->b.ipt0[k][e] = p[0 + k * 2 + e];
               c->b.jpt0[k][e] = p[10 + k * 2 + e];
               c->b.iptl[k][e] = p[20 + k * 2 + e];
               c->b.jptl[k][e] = p[30 + k * 2 + e];
            }
            c->b.chlnh[k] = (R) d[k];
            c->b.chwth[k] = (R) d[5 + k];
            c->b.chdph[k] = (R) d[10 + k];
            c->b.aresq0[k] = (R) a[k];
            c->b.aresql[k] = (R) a[5 + k];
         }
         c->b.aman = (R) q[0];
         c->b.bman = (R) q[1];
         c->b.gb = (R) q[2];
         c->b.agb = (R) q[3];
         c->b.agb13 = (R) q[4];
      } else {
         c->b.npss = 0;
      }
   }
   memcpy(c->b.sta, find(r, n, "STA")->data, 16);
   c->b.hter = (R) ((double *) find(r, n, "HTER")->data)[0];
   c->seadtm = (R) ((double *) find(r, n, "DATUM")->data)[0];
   {
      double *x = (double *) find(r, n, "CORL")->data, *g = (double *) find(r, n, "GAMAF")->data;
      for (k = 0; k < 90; k++) c->b.corl[k] = (R) x[k];
      for (k = 0; k < 4; k++) c->b.gamfp[k] = (R) g[2 + k];
   }
   c->b.nbcpt = ((int *) find(r, n, "NBCPT")->data)[0];
   c->b.zb = c->zb;
   c->b.zbm = c->zbm;
   c->b.ylt = c->ylt;
   c->b.itree = c->itree;
   c->b.kskp = c->kskp;
   c->b.imann = c->imann;
   c->b.is = c->is;
   c->b.ie = c->ie;
   c->b.ms = c->ms;
   c->b.me = c->me;
   c->b.ish = c->ish;
   c->b.jsh = c->jsh;
   c->b.elpct = c->elp[0];
   c->b.elpdt = c->elp[1];
   c->b.elpcl = c->elp[2];
   c->b.elpdl = c->elp[3];
   c->b.elpcl2 = c->elp[4];
   c->b.elpdl2 = c->elp[5];
   c->b.sint = c->trig[0];
   c->b.cost = c->trig[1];
   c->b.sinl = c->trig[2];
   c->b.cosl = c->trig[3];
   c->b.sinl2 = c->trig[4];
   c->b.ar2 = c->tab[0];
   c->b.ai2 = c->tab[1];
   c->b.br2 = c->tab[2];
   c->b.bi2 = c->tab[3];
   c->b.cr2 = c->tab[4];
   c->b.ci2 = c->tab[5];
   c->st.hb = c->hb;
   c->st.ub = c->ub;
   c->st.vb = c->vb;
   c->st.hmx = c->hmx;
   c->st.hsub = c->hsub;
   load_tables(c, r, n);
   for (k = 0; k < n; k++) free(r[k].data);
   free(r);
   return c;
}

int orc_info(void *vc, int *imxb, int *jmxb)
{
   ctx_t *c = vc;
   *imxb = c->b.imxb;
   *jmxb = c->b.jmxb;
   return 0;
}

/* tab: S, C, P, DELP, SW, CW, 800 each (Fortran index 1..800) */
int orc_set_tables(void *vc, const double *tab)
{
   ctx_t *c = vc;
   R *dst[6] = {c->S, c->C, c->P, c->DELP, c->SW, c->CW};
   int t, k;
   for (t = 0; t < 6; t++)
      for (k = 0; k < 800; k++) dst[t][k + 1] = (R) tab[t * 800 + k];
   return 0;
}

static void make_storm(ctx_t *c, const double *row, orc_storm *s);

/* The tide height at each BDRYHT square (COMMON /BCPTS/ TIDESH, which the C
 * driver's tide coupling sets between steps, slosh2.c:998-1025). */
int orc_set_tidesh(void *vc, const double *t, int n)
{
   ctx_t *c = vc;
   int k;
   if (n != c->b.nbcpt) return -1;
   if (!c->st.tidesh) c->st.tidesh = calloc((size_t) (n > 0 ? n : 1), sizeof(R));
   for (k = 0; k < n; k++) c->st.tidesh[k] = (R) t[k];
   return 0;
}

/* STATICHT (staticht.f:14-41): the static height at square (i,j) for the storm
 * values of row and the DELP column of the table set tab (S C P DELP SW CW, 800
 * each; NULL: the set in effect), in REAL. */
int orc_staticht(void *vc, const double *row, const double *tab, int i, int j, double *out)
{
   ctx_t *c = vc;
   const orc_basin *b = &c->b;
   orc_storm s;
   R xr, yr, x, y, rsq, r1, r2, dr, d0, d1;
   int k;
   make_storm(c, row, &s);
   xr = b->elpcl[i - 1] * b->cosl[j - 1];
   yr = b->elpdl[i - 1] * b->sinl[j - 1];
   x = xr - s.c1 - s.ax;
   y = yr - s.c2 - s.ay;
   rsq = x * x + y * y;
   r1 = (R) sqrt(rsq) / (R) 5280. + (R) 1.;
   k = (int) r1;
   r2 = (R) k;
   dr = r1 - r2;
   if (k > 790) k = 790;
   d0 = tab ? (R) tab[3 * 800 + k - 1] : c->DELP[k];
   d1 = tab ? (R) tab[3 * 800 + k] : c->DELP[k + 1];
   *out = (double) (d0 + dr * (d1 - d0) + s.seadtm);
   return 0;
}

static void make_storm(ctx_t *c, const double *row, orc_storm *s)
{
   int k;
   memset(s, 0, sizeof *s);
   s->itime = (int) row[0] + 1;                  /* TTIMER of this step */
   {
      double bt = 1.;
      int ic19 = (int) row[27];
      if (s->itime < ic19) {
         double time = s->itime, ang = time / ic19;
         bt = .5 * (1. - cos(ang * 3.14159265358979323846));
      }
      s->bt = (R) bt;
   }
   s->delt = (R) row[2];
   s->c1 = (R) row[3];
   s->c2 = (R) row[4];
   s->ax = (R) row[5];
   s->ay = (R) row[6];
   s->w1218[0] = (R) row[15];
   s->w1218[1] = (R) row[16];
   s->x1b1 = (R) row[31];
   s->x1b2 = (R) row[32];
   s->x1b4 = (R) row[24];
   s->x1b5 = (R) row[25];
   s->x1b6 = (R) row[26];
   s->seadtm = (R) row[28];
   for (k = 1; k <= 50; k++) s->x12[k] = (R) row[40 + k - 1];
   s->S = c->S;
   s->C = c->C;
   s->P = c->P;
   s->DELP = c->DELP;
   s->SW = c->SW;
   s->CW = c->CW;
}

/* one full step; diag[0..4] = drying events, smoothing pts, special, overflow,
 * filtered.  If hbs is not NULL it receives HB after HMXSV (before FILTER) as
 * double [jmxb][imxb]. */
int orc_advance_hbs(void *vc, const double *row, int *diag, double *hbs)
{
   ctx_t *c = vc;
   orc_storm s;
   int rc;
   size_t k, n = (size_t) c->b.imxb * c->b.jmxb;
   REAL *tmp = hbs ? malloc(sizeof(REAL) * n) : NULL;
   make_storm(c, row, &s);
   rc = orc_step(&c->b, &s, &c->st, &c->diag, tmp);
   if (hbs) {
      for (k = 0; k < n; k++) hbs[k] = (double) tmp[k];
      free(tmp);
   }
   if (diag) {
      diag[0] = c->diag.drying_events;
      diag[1] = c->diag.smooth_pts;
      diag[2] = c->diag.smooth_special;
      diag[3] = c->diag.smooth_overflow;
      diag[4] = c->diag.filtered;
   }
   return rc;
}

int orc_advance(void *vc, const double *row, int *diag)
{
   return orc_advance_hbs(vc, row, diag, NULL);
}

/* single routine, for step-injection tests */
int orc_call(void *vc, const char *routine, const double *row, int *diag)
{
   ctx_t *c = vc;
   orc_storm s;
   int rc = 0;
   make_storm(c, row, &s);
   memset(&c->diag, 0, sizeof c->diag);
   if (!strcmp(routine, "contty")) orc_contty(&c->b, &s, &c->st, &c->diag);
   else if (!strcmp(routine, "smpt2g")) rc = orc_smpt2g(&c->b, &s, &c->st, &c->diag);
   else if (!strcmp(routine, "bdryht")) orc_bdryht(&c->b, &s, &c->st);
   else if (!strcmp(routine, "hmxsv")) orc_hmxsv(&c->b, &s, &c->st);
   else if (!strcmp(routine, "filter")) orc_filter(&c->b, &s, &c->st);
   else if (!strcmp(routine, "xfilter1")) orc_xfilter(&c->b, &s, &c->st, 1);
   else if (!strcmp(routine, "xfilter2")) orc_xfilter(&c->b, &s, &c->st, 2);
   else if (!strcmp(routine, "momntm")) rc = orc_momntm(&c->b, &s, &c->st);
   else return -9;
   if (diag) {
      diag[0] = c->diag.drying_events;
      diag[1] = c->diag.smooth_pts;
      diag[2] = c->diag.smooth_special;
      diag[3] = c->diag.smooth_overflow;
   }
   return rc;
}

static R *field(ctx_t *c, const char *name)
{
   if (!strcmp(name, "HB")) return c->hb;
   if (!strcmp(name, "UB")) return c->ub;
   if (!strcmp(name, "VB")) return c->vb;
   if (!strcmp(name, "HMX")) return c->hmx;
   if (!strcmp(name, "HSUB")) return c->hsub;
   return NULL;
}

int orc_get(void *vc, const char *name, double *out)
{
   ctx_t *c = vc;
   R *f = field(c, name);
   size_t k, np = (size_t) c->b.imxb * c->b.jmxb;
   if (!strcmp(name, "CHQ")) {      /* one value per channel */
      for (k = 0; k < (size_t) c->b.npss; k++) out[k] = (double) c->st.chq[k];
      return c->b.npss;
   }
   if (!strcmp(name, "FLWSQR")) {   /* one value per 1-D flow side */
      for (k = 0; k < (size_t) c->b.nsqrwc; k++) out[k] = (double) c->flwsqr[k];
      return c->b.nsqrwc;
   }
   if (!f) return -1;
   for (k = 0; k < np; k++) out[k] = (double) f[k];
   return 0;
}

int orc_put(void *vc, const char *name, const double *in)
{
   ctx_t *c = vc;
   R *f = field(c, name);
   size_t k, np = (size_t) c->b.imxb * c->b.jmxb;
   if (!f) return -1;
   for (k = 0; k < np; k++) f[k] = (R) in[k];
   return 0;
}

void orc_close(void *vc)
{
   ctx_t *c = vc;
   int k;
   if (!c) return;
   free(c->st.tidesh);
   free(c->zb); free(c->zbm); free(c->ylt); free(c->hb); free(c->ub); free(c->vb);
   free(c->hmx); free(c->hsub); free(c->itree); free(c->kskp); free(c->imann);
   free(c->is); free(c->ie); free(c->ms); free(c->me); free(c->ish); free(c->jsh);
   for (k = 0; k < 6; k++) free(c->elp[k]);
   for (k = 0; k < 5; k++) free(c->trig[k]);
   for (k = 0; k < 6; k++) free(c->tab[k]);
   for (k = 0; k < 10; k++) free(c->edge[k]);
   free(c->isqr); free(c->jsqr); free(c->iside); free(c->ktree); free(c->f1dact); free(c->flwsqr);
   free(c);
}
