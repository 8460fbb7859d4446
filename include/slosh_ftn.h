/* slosh_ftn.h -- the reference's Fortran-ABI entry points, exported by
 * libslosh_ftn.so over the HIP time step.
 *
 * Link the reference's C driver (sorc/slosh.fd: cstart.c, slosh2.c and the
 * other C sources of its Makefile:187-200) against libslosh_ftn.so and
 * libslosh_hip.so instead of the Fortran objects: the CLI, its list runs and
 * its env/rex writers then run unchanged with the model on the GPU.  ABI of the
 * reference default build (-DDOUBLE_FORTRAN -D_GCC_): lower-case names with a
 * trailing underscore, every argument by reference, grids double [1655][999]
 * (BAS_Y x BAS_X, halotype.h:60-61), I fastest.
 *
 * COMMON blocks exported with the layouts the C driver declares (#pragma
 * pack(2)): dumb5_ {ZB, ZBM, D_WAV} (slosh2.c:175-190), dumb7_ {UB, VB, HB}
 * (:209-224), dumb10_ {HMX} (:243-258), egth_ {DELS, DELT, G, COR} (:192-207),
 * frank_ {IRANK, ISTAR, ISTOP, JSTAR, JSTOP, IHALO} (:226-241), bcpts_
 * {TIDESH[12000], NBCPT, ISH[12000], JSH[12000]} (:285-303), llxfle_ {YLT, YLG}
 * (savellx.c:686-700).
 */
#ifndef SLOSH_FTN_H
#define SLOSH_FTN_H
#ifdef __cplusplus
extern "C" {
#endif

/* INITALCOMMON (initalcommon.f): zero the COMMON state (the mirrors start zeroed). */
int initalcommon_(void);

/* INITAL (inital.f:1-272; C prototype slosh2.c:111-135): read the basin (dta),
 * the track (trk) and initialise; returns MHALT, IMXB, JMXB, ZB, the run start
 * date, DELT and the model clock (0: no hot start).  Uses the cell latitudes
 * memSetLLx left in llxfle_. */
int inital_(int *mhalt, int *imxb, int *jmxb, double zb[1655][999], int *hour, int *min, int *day,
            int *month, int *year, char *trk, int *len_trk, char *dta, int *len_dta, char *env,
            int *len_env, char *ft40, int *len_ft40, double *delt, char *bsn_abrev, int *len_bsn,
            double *clock, short *wave, short *full_depth);

/* TMSTPC1 (tmstpc1.f:1-94; slosh2.c:24-37): first part of a step. */
int tmstpc1_(double hhb[1655][999], short *fpass, short *restart);

/* TMSTPC2 (tmstpc2.f:1-120; slosh2.c:39-52): second part; when *fpass, hhb
 * receives min(HB, HMX) after HMXSV. */
int tmstpc2_(double hhb[1655][999], short *fpass, short *restart);

/* TMSTPM (tmstpm.f:1-72; slosh2.c:54-70): momentum stage and storm advance;
 * returns ITIME, MHALT, DELT and the storm's lat/lon, DELP, size, speed and
 * direction. */
int tmstpm_(int *itime, int *mhalt, double *delt, double *storm_lat, double *storm_lon, double *delp,
            double *size, double *wspeed, double *wdirct, double *clock, short *restart);

/* CLNUP (clnup.f:1-24; slosh2.c:139-150): hhb receives the envelope HMX. */
int clnup_(double hhb[1655][999], int *f_save_env);

/* Tide / sub-grid entry points (staticht.f, issubgrid.f, tmstp2.f, intlhtsub.f,
 * intlhtadd.f): present for linking; this build stops with a message if the
 * driver reaches them (-f_tide, tide spin-up).  CLN_HS (cln_hs.f, called at the
 * end of every run, slosh2.c:2771) deletes hot-start files: a no-op here. */
int staticht_(int *i, int *j, double *ans);
int issubgrid_(int *i, int *j, int *ans);
int tmstp2_(void);
int intlhtsub_(void);
int intlhtadd_(void);
int cln_hs_(void);

#ifdef __cplusplus
}
#endif
#endif
