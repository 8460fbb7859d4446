/* sl_dev.h -- device-side data of the SLOSH step (shared by the HIP kernels and
 * the host code that fills them), for R = float (the FP32 step) or double (the
 * FP64 step, the reference's default precision).
 *
 * Fields are [storm][J-1][pitch] with I fastest; pitch is imxb rounded up to 64
 * floats so every row starts on a 256-byte boundary.
 */
#ifndef SL_DEV_H
#define SL_DEV_H
#include <stdint.h>

#define SL_NTABV 801          /* radial tables, index 1..800 used (Fortran order) */
#define SL_NPAR 26
#define SL_EVCAP 4096         /* per-storm CONTTY event / dirty-square capacity */
#define SL_MAXSPLIT 4         /* parts of a batch on their own streams */
#define SL_SGROUP 4           /* default storms per group in the fused kernels' workgroup order */
#define T_AI2_OFF 147000      /* offset of AI2 in the [6][245][600] table block */
#define SL_M2G 30000          /* corner-smoothing list limits, smpt2g.f:17 / parm.for:26 */
#define SL_MCT 8000

/* per-point static flags */
#define F_ACT (1u << 0)       /* CONTTY square: KSKP == '1' */
#define F_BDRY (1u << 1)      /* BDRYHT static-height square */
#define F_MOM (1u << 2)       /* MOMNTM interior point (J 2..JEND, I MS..ME, ITREE not 4/6) */
#define F_TREE1 (1u << 3)     /* ITREE == '1' (trees: strong sheltering) */
#define F_OCEAN (1u << 4)     /* ITREE '2' or '5': ocean winds */
#define F_IT2 (1u << 5)       /* ITREE == '2' */
#define F_BSKIP (1u << 6)     /* ITREE '6' or '4' (as a boundary point: no computation) */
#define F_B8 (1u << 7)        /* ITREE == '8' (intermediate-depth boundary) */
#define F_ZR (1u << 19)       /* 2 <= I <= IMXB1 and 2 <= J <= JMXB1 (UB underflow flush) */
#define F_PFC (1u << 20)      /* SMPT2G candidate: 2 <= I <= IMXB-2, 2 <= J <= JMXB-2, ZBM > -30 */
#define F_IIK(f) (((f) >> 8) & 0xffu)
#define F_BOP(f) (((f) >> 16) & 7u)   /* 1..4: one boundary point, ISKIP; 5/6: corner 1/3 */

template <class R>
struct SlBasinDevT {
   int imxb, jmxb, pitch, jend;
   int npts;                  /* pitch * jmxb: elements of one field */
   int ebsn_ell;              /* grid type '$' or '+': metric GAMF in FILTER/FLT2G */
   int ebsn1_amp;             /* EBSN1 == '&' (HMXSV: no smoothing) */
   int island;                /* DOLLAR '$ ': closed island, periodic in J (FLTER2,
                                 MOMNTM to row JMXB, row 1 repeats row JMXB) */
   int ismth, xver;           /* FILTER (1) or XFILTER hcrt code; XFILTER version */
   R hter, hcrt, exclude;
   R gamfp[4];
   int nbcpt;
   const R *zb, *zbm, *zcor;
   const uint32_t *flags;
   const int *is, *ie, *ms, *me;                                      /* [jmxb] */
   const R *elpct, *elpdt, *elpcl, *elpdl, *elpcl2, *elpdl2;          /* [pitch] */
   const R *sint, *cost, *sinl, *cosl, *sinl2;                        /* [jmxb] */
   const R *tab8;             /* [245][600][8]: AI2 SLPAI AR2 BR2 BI2 CR2 CI2 0 per (iik, id) */
   const R *g4;               /* [4][pitch*jmxb]: metric weight GAMF of neighbour k (IIH/JJH order) */
   const int *bcidx;          /* [pitch*jmxb]: index n of a BDRYHT square in the ISH/JSH list, else -1 */
   /* sub-grid 1-D flow sides, L = 1..nsqrwc at index L-1 (1-D sides, weirs
    * from nsqrs+1, cuts from nsqrw+1); f1d_on: FLW1DM runs (sides and NOF1D
    * not '+'); sfl: STA starts with "SOUTH FL" */
   int nsqrs, nsqrw, nsqrwc, nsq, f1d_on, sfl;
   const int *e_int;          /* [nsqrwc][SL_EI]: see E1_* */
   const R *e_real;           /* [nsqrwc][8]: HWEIR ZBMIN BANK(,1) BANK(,2) DELCUT COS1D SIN1D 0 */
   const R *e_cut;            /* [ncut][4]: CUTL CUTLI CUTLE 0 */
   /* channels (CHANNL): entrance squares (i,j)x2, exit squares (i,j)x2 */
   int npss;
   int ch_pt[5][8];
   R ch_lnh[5], ch_wth[5], ch_dph[5], ch_a0[5], ch_al[5];
   R aman, bman, gb, agb, agb13;
};
#define SL_EI 10
enum { E1_I, E1_J, E1_K, E1_II, E1_JJ, E1_IIK, E1_T, E1_QI, E1_QE, E1_PAD };   /* E1_T: KTREE 'T' */
enum { E2_HWEIR, E2_ZBMIN, E2_BANK1, E2_BANK2, E2_DELCUT, E2_COS, E2_SIN };

/* per-step storm values (what the field routines read during that step);
 * P_ITIME, P_TVER and P_FLAGS hold small integers as values */
enum {
   P_BT, P_DELT, P_X1B4, P_X1B5, P_X1B6, P_C1, P_C2, P_AX, P_AY, P_X4, P_X9, P_X10,
   P_X15, P_X20, P_X21, P_X23, P_X24, P_W0, P_W1, P_SEA, P_ITIME, P_TVER, P_FLAGS, P_X1B1, P_X1B2,
   P_PAD
};
#define SF_SMPT (1 << 0)      /* SMPT2G/FLT2G this step */
#define SF_FILT (1 << 1)      /* FILTER/XFILTER this step */

/* Storm program: what the device needs to generate a run's initial water and
 * per-step values (the P_* above) itself.  Doubles, per storm:
 *   [0] nsteps  [1] IC19  [2] nev  [3] DTMLAK (the lake datum)
 *   [4 .. 4+SQ_N)  the storm state after SETCMP, in SQ_* order
 *   DELP(0 .. SL_NTABV-1): the static-height table of the first hour (INTLHT)
 *   BT(ITIME) for ITIME = 1 .. IC19-1 (TTIMER's ramp)
 *   nev events of SQ_EV doubles: n, then the state (SQ_* up to SQ_AX) left by
 *   the n-th STMVAL call wherever that call changed any of it */
enum {
   SQ_DELT, SQ_X1B0, SQ_X1B1, SQ_X1B3, SQ_X1B4, SQ_X1B5, SQ_C1, SQ_C2, SQ_X4, SQ_X9, SQ_X10, SQ_X11, SQ_X12,
   SQ_X15, SQ_X20, SQ_X21, SQ_X23, SQ_X24, SQ_W0, SQ_W1, SQ_SEA, SQ_TV, SQ_AX, SQ_AY, SQ_N
};
#define SQ_HDR 4
#define SQ_DELP (SQ_HDR + SQ_N)                   /* offset of the DELP table */
#define SQ_BT (SQ_DELP + SL_NTABV)                /* offset of BT(1) */
#define SQ_EV (SQ_AX + 1)

template <class R>
struct SlRunDevT {
   int nstorm;
   size_t fstride;            /* elements per storm field */
   R *hb0, *hb1, *ub, *vb, *hmx;   /* [nstorm][fstride] */
   R *hb2;                    /* [nstorm][fstride]: HB after TMSTPC1 on the fused path of
                                 basins with 1-D flow sides or channels (else null) */
   R *tidesh;                 /* [nstorm][nbcpt]: tide height BDRYHT adds at each boundary
                                 square (COMMON /BCPTS/ TIDESH), or null: 0 */
   R *hsubtop;                /* [nstorm][pitch]: HSUB row JMXB for XFILTER */
   uint8_t *pflag, *hclass;   /* [nstorm][fstride] (staged path) */
   uint8_t *uz;               /* [nstorm][fstride]: UB == 0 after CONTTY (fused path) */
   uint32_t *evkey;           /* [nstorm][SL_EVCAP] */
   int *evcnt;                /* [nstorm] */
   int *stat;                 /* [nstorm][8]: counters / error flags */
   const R *params;           /* [nstorm][maxsteps][SL_NPAR] */
   int maxsteps;
   const R *tables;           /* table pool [*][SL_NTABV][8]: C S CW SW P DELP 0 0 per radius */
   const int *tbase;          /* [nstorm] first table of each storm in the pool */
   const int *nsteps;         /* [nstorm] steps of each run (0: empty slot) */
   int *cur;                  /* [nstorm] steps completed; the kernels of a step work on
                                 step cur+1 of each storm and skip finished storms */
   R *flwsqr;                 /* [nstorm][nsqrwc]: transport across each 1-D flow side */
   uint8_t *f1dact;           /* [nstorm][nsqrwc]: 1 where FLW1DM_C left water on both sides */
   R *hsq;                    /* [nstorm][nsq]: HSUB of the 1-D flow squares (FLW1DM_C -> _M) */
   R *chq;                    /* [nstorm][5]: each channel's flow, kept from step to step */
   int evcap;                 /* drying squares listed per storm before the replay sweeps
                                 every square (1 .. SL_EVCAP; SL_EVCAP in production) */
   long vbudget;              /* visits one replay may make (0: 64 * squares + 4096) */
   int sgroup;                /* storms whose tiles follow each other in the fused
                                 kernels' workgroup order (1: storm by storm) */
};

/* The gridded tide between steps (-f_tide V2 modes; CalcTideGrid2,
 * slosh2.c:813-1027).  Runs that share a year and a model clock form a clock
 * group; a group's tide field is summed once per step and every run of the
 * group takes it from there.  Squares are the (IMXB-1) x (JMXB-1) grid of the
 * harmonic constants, I fastest. */
#define SL_NTIDE 37
struct SlTideDev {
   int ni, ncell;             /* squares per row (IMXB-1), squares */
   int nbcpt, ngrp;
   const double *amp, *cv, *sv;   /* [nyear][SL_NTIDE][ncell]: nodal factor x amplitude,
                                     cos and sin of (equilibrium argument - phase) */
   const double *mtl;         /* [ncell]: NAVD88 - MTL */
   const double *zb;          /* [ncell]: ZB in double */
   const uint8_t *cfl;        /* [ncell]: 1 the tide is added (SetTideUsableFlag),
                                 2 HB = STATICHT + tide (V2.1 / V2.2, ZB >= 290) */
   const int *bccell;         /* [nbcpt]: the square of each BDRYHT point */
   const int *list;           /* [nlist]: the squares with a flag (the only ones a step changes) */
   const int *call;           /* group g from goff[g] / (2 SL_NTIDE): per step 0 .. gmax[g] the
                                 number of the tide call made after it (0 before the first
                                 step), or -1: no call (the hsf basins call every 6 minutes) */
   int nlist;
   const double *casa;        /* group g from goff[g]: per step 0 .. gmax[g] the cos, then
                                 the sin, of each constituent's speed x hours */
   const long *goff;          /* [ngrp] */
   const int *gyear;          /* [ngrp]: index of the group's year in amp, cv, sv */
   const int *gmax;           /* [ngrp]: last step of the group's longest run */
   const int *sgrp;           /* [nstorm]: the group of each run */
   double *field;             /* [ngrp][2][ncell]: tide at the group's step n in slot n & 1 */
};

#endif
