/* sl_host.h -- host side of the SLOSH time-stepping core: basin and track
 * ingest, grid metrics, bottom-stress tables, initial water and the per-storm
 * state machine that advances the parametric storm each step.
 *
 * Everything here runs once per basin / once per storm / once per step on
 * scalars; it is computed in double so that it reproduces the reference's
 * default (FP64) build bit for bit.  The field arithmetic of the step (the hot
 * path) runs on the GPU in FP32 from the float-rounded results of this code.
 *
 * Index conventions: the reference is Fortran, (I,J) 1-based with I fastest.
 * Per-point arrays here are packed [J-1][I-1] (imxb x jmxb); per-I vectors are
 * indexed I-1 and per-J vectors J-1.
 */
#ifndef SL_HOST_H
#define SL_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SL_NT 999          /* track hours, inputn2.f:10 NT_ */
#define SL_NBCPTS 12000    /* static-height boundary squares, bdryht.f:59 */
#define SL_NTAB 800        /* radial tables at 1-mile spacing, force1.f */
#define SL_L1D 41000       /* 1-D flow points + weirs + cuts, parm.for:16 L_ */
#define SL_NCUT 3000       /* cuts, parm.for:16 NCT_ */

typedef struct sl_basin {
   int imxb, jmxb, imxb1, jmxb1, imxb2, jmxb2;
   char sta[17];
   char ebsn, dollar[3], ebsn1, ebsn2, isouth, nofld, nof1d, iterrain, ismooth;
   int iver, iprj, ismth;
   int full_depth;                    /* -fullDepth: depths not capped at 600 ft */
   double hter, fsouth;
   double azmth, alto, alno, degree, dela, rmouth, aaxis, baxis, abqab;
   double xmouth, ymouth, dstnt, asmpt, gsize;
   double phi, phi1, alt1, alt1c, aln1, cor;
   int irb, ire, ndb, nde, ndb1;
   double dltin[3];
   /* per point, [J-1][I-1] */
   double *zb, *zbm, *ylt;
   char *itree, *kskp;
   int *imann;
   /* per J */
   int *is, *ie, *ms, *me, *mf;
   double *sint, *cost, *sinl, *cosl, *sinl2, *sint2;
   /* per I */
   double *elpct, *elpdt, *elpcl, *elpdl, *elpcl2, *elpdl2;
   /* open-boundary squares (DEPSFC) */
   int nbcpt;
   int *ish, *jsh;
   double *tidesh;
   /* constants (CNSTNT) */
   double delt, x1b[10], corl[90], gama, gama1, gamfp[4];
   /* bottom-stress tables (BTMSTR), [N-1][II-1], N=1..600, II=1..245 */
   double *ar2, *ai2, *br2, *bi2, *cr2, *ci2;
   /* sub-grid 1-D flow edges (CRDRD2 / DEPSFC), L = 1..nsqrwc at index L-1:
    * 1..nsqrs 1-D flow sides, nsqrs+1..nsqrw raised weirs, nsqrw+1..nsqrwc
    * cuts (cut data cutl/cutli/cutle at index L-nsqrw-1). */
   int nsqrs, nsqrw, nsqrwc, npss, ncut, nodry;
   int *idry, *jdry;                  /* '1$' basins: squares dried at the start (COMMON /LDRY/) */
   double dtmlak_add;                 /* '2$' basins: DEPMSY's change to the lake datum */
   int *isqr, *jsqr, *iside, *iang;
   char *ktree;
   double *hweir, *zbmin, *bank1, *bank2, *delcut, *cos1d, *sin1d;
   double *cutl, *cutli, *cutle;
   /* channels (CRDRD2, crdrd2.f:286-299; at most 5): length (mi), width and
    * depth (ft), entrance / exit squares (two each), the entrance and exit
    * areas SETDST approximates (setdst.f:88-107) and CNSTNT's constants */
   double chlnh[5], chwth[5], chdph[5], aresq0[5], aresql[5];
   int ipt0[5][2], jpt0[5][2], iptl[5][2], jptl[5][2];
   double aman, bman, gb, agb, agb13;
   int ipn[10], jpn[10];
   int have_ylt;
   /* the grid's tangent-plane projection from the ?basins.dta line
    * (load_bsndta / memSetLLx, savellx.c:704-765, rexout.cd/convert.c:404-533):
    * grid type 'p' 'e' 'h'; G (conformal latitude, longitude as the line
    * holds it, deg), the map scale at G, sin / cos of G's latitude, P on the
    * plane, the grid origin and spacing */
   struct {
      int have;
      char type;
      double glat, glon, tscald, stl, ctl, xppt, yppt, rg2p, thtg2p, abq, delrg, xig, yjg, ximax, yjmax;
   } proj;
} sl_basin;

typedef struct sl_storm {
   /* hourly track after interpolation (1-based, index 0 unused) */
   int trklen;
   double X[SL_NT + 1], Y[SL_NT + 1], PT[SL_NT + 1], R[SL_NT + 1];
   double DIR[SL_NT + 1], SP[SL_NT + 1], XLAT[SL_NT + 1], YLONG[SL_NT + 1];
   double rlngth, seadtm, dtmlak;
   int ibgnt, itend, jhr;
   int date_yr, date_mo, date_dy, date_hr, date_mn;
   /* time stepping */
   double delt, dltb, dlthr, dlth2, bt, treal;
   int mhalt, ndlthr, ndlth2, ic12, ic19, itime, itmadv, nhrad, jumpr, iprt, key1;
   int iprtsv, iprtad;
   double x1b[10];
   /* storm state (STRMSB, WTEMP, FRST, DUMY44, DUMMY4) */
   double c1, c2, c21, c22, ax, ay, ptency, rtency;
   double ydelp, zdelp, pnn, yc24, zc24, zlato;
   double x12[51];                     /* 1-based */
   double w1218[2], wmax[2];
   double S[SL_NTAB + 1], C[SL_NTAB + 1], P[SL_NTAB + 1], DELP[SL_NTAB + 1];
   double SW[SL_NTAB + 1], CW[SL_NTAB + 1];
   int table_gen;                      /* bumps whenever FORCE1 refills the tables */
} sl_storm;

/* Basin: parse the dta file and derive everything the step needs.
 * Returns 0, or -1 with a message in err (e.g. an unsupported feature). */
int sl_basin_load(const char *dta_path, sl_basin *b, char err[256]);
int sl_basin_load_opts(const char *dta_path, sl_basin *b, int full_depth, char err[256]);
/* Cell latitudes for the Coriolis term from the ?basins.dta line
 * (restates memSetLLx, savellx.c:704-765).  0 on success. */
int sl_basin_llx(sl_basin *b, const char *bnt_line, char err[256]);
void sl_basin_free(sl_basin *b);

/* A track as INPUTN2 reads it (rows 1..n, hours counted from 1, longitudes
 * west positive), before the hourly interpolation. */
typedef struct sl_trk {
   int date_yr, date_mo, date_dy, date_hr, date_mn;
   int dhr, land, beg, end, adj, spline, n;
   int v1992;          /* the original (1992) track format: 100 hourly rows, no factors */
   double ocean, lake, windf, rmwf;
   int hrs[SL_NT + 2];
   double fla[SL_NT + 2], flo[SL_NT + 2], fdp[SL_NT + 2], frm[SL_NT + 2];
} sl_trk;
void sl_trk_init(sl_trk *t);
/* Add a row (longitude east positive, as a trk file holds it); 0, 1 when the
 * track is full (the row is not read), -1 with a message when out of bounds. */
int sl_trk_row(sl_trk *t, double hour, double lat, double lon, double dp, double rmw, char err[256]);
/* A trk file in either of INPUTN2's formats (FileVersion=2013, or the 1992
 * format of 100 hourly rows). */
int sl_trk_parse(const char *path, sl_trk *t, char err[256]);
int sl_trk_build(const sl_trk *t, sl_storm *s, char err[256]);

/* The conformal sphere and the basin's tangent plane (savellx.c:40-227):
 * Clarke <-> conformal latitude; lat/lon (conformal) <-> plane x, y (statute
 * miles, origin at G, y to local north); grid (p, q) -> plane. */
void sl_cl2cnf(double clk, double *cnf, double *scale);
void sl_cnf2cl(double cnf, double *clk);
void sl_proj_sll2xy(const sl_basin *b, double lat, double lon, double *x, double *y);
void sl_proj_sxy2ll(const sl_basin *b, double x, double y, double *lat, double *lon);
void sl_proj_pq2xy(const sl_basin *b, double p, double q, double *x, double *y);

/* MEOW seed points (conMeowGenCmd, rexout.cd/convert.c:1181-1288): the
 * landfall points of parallel tracks heading dir (deg from north) spaced
 * apart (mi) across the basin's shadow widened by r34 (mi), about the point
 * (lat, lon in the basin line's convention, deg).  xyflag 0 steps along the
 * tangent plane, 1 along great circles (Earth_Jump, convert.c:1127-1151).
 * num: at most num points (-1: as many as the shadow holds).  Points in the
 * reference's order into idx/lat/lon[cap]; returns their count, or -1. */
int sl_meow_seeds(const sl_basin *b, int xyflag, double lat, double lon, double dir, double apart, int num,
                  double r34, int cap, int *idx, double *slat, double *slon);
/* Earth_Jump (convert.c:1127-1151): the point dist statute miles from
 * (lat, lon) (deg) on bearing bear (deg from north). */
void sl_earth_jump(double lat, double lon, double dist, double bear, double *lat2, double *lon2);

/* The basin's harmonic constants for the gridded astronomical tide (-f_tide
 * V2 modes; tideutil.c): per square of the (IMXB-1) x (JMXB-1) grid, in the
 * file's order (I fastest), the 37 amplitudes (1e-6 ft) and phases (1e-6 deg)
 * as read, the NAVD88 - MTL offset and the square's lon / lat. */
#define SL_NUMT 37
typedef struct sl_tide {
   int ni, nj, ncell;
   char abbrev[6];
   int32_t *amp, *phs;           /* [ncell][SL_NUMT] */
   double *mtl, *lon, *lat;      /* [ncell] */
} sl_tide;
/* seconds from 1970-01-01 00:00 to 00:00 of the date (Clock_ScanDate) */
double sl_clock_scan_date(int year, int month, int day);
/* constituent k's speed in radians per hour */
double sl_tide_speed(int k);
/* bhc (HarmConstit v1) and adj (datum adjustment text; NULL or missing: 0)
 * for a basin of imxb x jmxb points */
int sl_tide_open(sl_tide *t, const char *bhc, const char *adj, int imxb, int jmxb, char err[256]);
/* one year's tables from ft03, [SL_NUMT][ncell] each: nodal factor x
 * amplitude, cos and sin of (equilibrium argument - phase) */
int sl_tide_year(const sl_tide *t, const char *ft03, int year, double *xamp, double *cosv, double *sinv,
                 char err[256]);
void sl_tide_free(sl_tide *t);

/* Track: parse the trk file and initialise the storm (INPUTN2, STERON, INTDLT,
 * INTVAL, SETCMP).  Needs the loaded basin.  sl_storm_init does the same from
 * a track already parsed or built in memory. */
int sl_storm_load(const char *trk_path, const sl_basin *b, sl_storm *s, char err[256]);
int sl_storm_init(const sl_trk *t, const sl_basin *b, sl_storm *s, char err[256]);
/* Initial water HB (INTLHT, intlht.f:53-185) into hb[jmxb*imxb]; also zeroes
 * nothing else -- UB=VB=0 and HMX=HB are the caller's. */
void sl_initial_water(const sl_basin *b, const sl_storm *s, double *hb);
/* How INTLHT fills square (i,j), i <= IMXB1, j <= JMXB1: SL_IW_STORM the
 * storm's static height (else a datum: the lake datum, or with SL_IW_SEA the
 * sea datum), SL_IW_DRY dried afterwards (-ZB). */
#define SL_IW_STORM 1
#define SL_IW_DRY 2
#define SL_IW_SEA 4
int sl_intlht_class(const sl_basin *b, int i, int j);

/* Per-step storm machinery.  Order inside one step, as TMSTPC1/TMSTPM run it:
 * sl_storm_ttimer() at the start (ITIME+1, ramp BT), then the field kernels,
 * then sl_storm_stmval() after the momentum update. */
void sl_storm_ttimer(sl_storm *s);
int sl_storm_stmval(sl_storm *s, const sl_basin *b);
/* storm centre lat/lon (XY2LLX, xy2llx.f) for the rex track record */
void sl_storm_latlon(const sl_storm *s, const sl_basin *b, double *lat, double *lon);

/* FORCE1 and INTWND, exposed for tests. */
int sl_force1(double pdrop, double rmax, double *vmax, double alto, int ncatg,
              double S[SL_NTAB + 1], double C[SL_NTAB + 1], double P[SL_NTAB + 1],
              double DELP[SL_NTAB + 1]);
double sl_intwnd(double pdrop, double rmax, double alto);

/* The datums of a surge-only run after InitWater_TideModeOverride
 * (slosh2.c:345-482, f_tide 0), as the env and rex trailers carry them. */
void sl_datum_override(float *ht1, float *ht2);
/* The same for a run with the gridded tide -f_tide f_tide (2, 21, 22, 24). */
void sl_datum_override_tide(float *ht1, float *ht2, int f_tide);

#ifdef __cplusplus
}
#endif
#endif
