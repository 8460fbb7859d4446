/* orc.h -- CPU restatement of one SLOSH time step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU path is compared
 * against; nothing in the product links or calls it.  It is compiled twice:
 * with REAL=double it must reproduce the reference's default FP64 build bit
 * for bit (that is how it is pinned), with REAL=float it is the FP32 oracle
 * the HIP kernels must match.
 *
 * Layout: 2-D fields are packed [J-1][I-1] (I fastest, imxb x jmxb); 1-D
 * vectors by I-1 or J-1; bottom-stress tables [N-1][II-1] with N=1..600,
 * II=1..245 (the reference's AR2(245,600) column-major order).
 */
#ifndef ORC_H
#define ORC_H
#ifndef REAL
#define REAL double
#endif

typedef struct {
   int imxb, jmxb;
   char ebsn, dollar[2], ebsn1;
   int ismth, iver;
   REAL hter;
   char sta[16];
   const REAL *zb, *zbm, *ylt;
   const char *itree, *kskp;
   const int *imann;
   const int *is, *ie, *ms, *me;              /* by J-1 */
   const REAL *elpct, *elpdt, *elpcl, *elpdl, *elpcl2, *elpdl2;   /* by I-1 */
   const REAL *sint, *cost, *sinl, *cosl, *sinl2;                  /* by J-1 */
   const REAL *ar2, *ai2, *br2, *bi2, *cr2, *ci2;                  /* [600][245] */
   REAL corl[90], gamfp[4];
   int nbcpt;
   const int *ish, *jsh;
   /* sub-grid 1-D flow edges, L = 1..nsqrwc at index L-1 (1-D sides, then
    * weirs from nsqrs+1, then cuts from nsqrw+1; cut data at L-nsqrw-1) */
   int nsqrs, nsqrw, nsqrwc, ncut;
   char nof1d;
   const int *isqr, *jsqr, *iside;
   const char *ktree;
   const REAL *hweir, *zbmin, *bank1, *bank2, *delcut, *cos1d, *sin1d, *cutl, *cutli, *cutle;
   /* channels (at most 5): entrance/exit squares, length, width, depth, end
    * square areas, CNSTNT constants */
   int npss;
   int ipt0[5][2], jpt0[5][2], iptl[5][2], jptl[5][2];
   REAL chlnh[5], chwth[5], chdph[5], aresq0[5], aresql[5];
   REAL aman, bman, gb, agb, agb13;
} orc_basin;

/* Storm values the field routines read during one step: everything as it
 * stands after the previous step's STMVAL, plus this step's TTIMER ramp. */
typedef struct {
   int itime;                 /* after TTIMER */
   REAL bt;
   REAL delt, x1b1, x1b2, x1b4, x1b5, x1b6;
   REAL c1, c2, ax, ay;
   REAL x12[51];              /* 1-based */
   REAL w1218[2];
   REAL seadtm;
   const REAL *S, *C, *P, *DELP, *SW, *CW;     /* 1-based, >= 792 entries */
} orc_storm;

typedef struct {
   REAL *hb, *ub, *vb, *hmx;  /* imxb x jmxb */
   REAL *hsub;                /* scratch, imxb x jmxb, persists across steps */
   REAL *flwsqr;              /* transport across each 1-D flow side (nsqrwc) */
   char *f1dact;              /* 'T' where both squares of a side held water after FLW1DM_C */
   REAL chq[5];               /* each channel's flow (CHANNL's saved Q) */
   REAL *tidesh;              /* tide height at each BDRYHT square (BCPTS TIDESH, nbcpt), or NULL: 0 */
} orc_state;

typedef struct {
   int drying_events;         /* CONTTY squares that took the back-step branch */
   int smooth_pts, smooth_special, smooth_overflow;
   int filtered;
} orc_diag;

int orc_step(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d, REAL *hbs);

/* individual routines, for step-injection tests */
void orc_contty(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d);
int orc_smpt2g(const orc_basin *b, const orc_storm *s, orc_state *st, orc_diag *d);
void orc_bdryht(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_hmxsv(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_filter(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_xfilter(const orc_basin *b, const orc_storm *s, orc_state *st, int iversion);
int orc_momntm(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_flw1dm_c(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_channl(const orc_basin *b, const orc_storm *s, orc_state *st);
void orc_flw1dm_m(const orc_basin *b, const orc_storm *s, orc_state *st);
#endif
