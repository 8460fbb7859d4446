/* ref_driver.c -- state-dumping driver over the reference SLOSH objects.
 *
 * Test infrastructure: links the reference's own Fortran objects (compiled from
 * /root/reference by oracle/Makefile) and replays the call sequence of the
 * reference's serial run (RunInit -> RunLoopStep* -> CleanUp, slosh2.c:1712,
 * 1321, 1674; PerformRun loop slosh2.c:2404) while dumping Fortran COMMON state
 * to tagged binary records that tests/ read with numpy.
 *
 *   refdrv64 <bnt-line-file> <dta> <trk> <outdir> <nsteps|-1> <dump-spec>
 *
 * dump-spec: comma list of step numbers to dump full state after ("0" = right
 * after INITAL), "all" for every step, "" for none.  Storm scalars are written
 * for every step to <outdir>/storm.bin; state to <outdir>/s%06d.bin; the static
 * init state to <outdir>/init.bin; the final envelope (after CLNUP) to final.bin.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define M_ 999
#define N_ 1655
typedef double R;                     /* FP64 reference build (-fdefault-real-8) */

/* ---- COMMON blocks (parm.for:30-92 and the per-routine declarations) ---- */
extern struct { R UB[N_][M_], VB[N_][M_], HB[N_][M_]; } dumb7_;
extern struct { R ZB[N_][M_], ZBM[N_][M_], D_WAV[N_][M_]; } dumb5_;
extern struct { R HMX[N_][M_]; } dumb10_;
extern struct { R HSUB[N_][M_]; } scrtch_;
extern struct { R YLT[N_][M_], YLG[N_][M_]; } llxfle_;
extern struct { char ITREE[N_][M_]; } itree_;
extern struct { int IMANN[N_][M_]; } imann_;
extern struct { char KSKP[N_][M_]; } kspcon_;
extern struct { R ELPCT[M_], ELPDT[M_], ELPCL[M_], ELPDL[M_]; } elpdst_;
extern struct { R ELPCL2[M_], ELPDL2[M_], SINL2[N_], SINT2[N_]; } elpds2_;
extern struct { R SINT[N_], COST[N_], SINL[N_], COSL[N_]; } plrpst_;
extern struct { int IS[N_], MS[N_], MF[N_]; } setmix_;
extern struct { int IE[N_], ME[N_]; } setmx1_;
extern struct { int IMXB, JMXB, IMXB1, JMXB1, IMXB2, JMXB2; } dumb3_;
extern struct { R X1B[10]; } dumb4_;
extern struct { R CORL[90]; } dumb4i_;
extern struct { R AR2[600][245], AI2[600][245], BR2[600][245], BI2[600][245],
                  CR2[600][245], CI2[600][245], SLPAI2[600][245]; } scnd2_;
extern struct { R DELS, DELT, G, COR; } egth_;
extern struct { int IC12, IC19; R BT; } frth_;
extern struct { int ITIME, MHALT; } ffth_;
extern struct { char DOLLAR[2]; char EBSN; } gprt1_;
extern struct { char EBSN1, EBSN2; } gprt2_;
extern struct { char STA[16]; } gprt_;
extern struct { int IVER, IPRJ; } dtaopt_;
extern struct { int ISMTH; } smth_;
extern struct { R HTER; } hterain_;
extern struct { R GAMA, GAMA1, GAMFP[4]; } gamaf_;
extern struct { R TIDESH[12000]; int NBCPT; int ISH[12000], JSH[12000]; } bcpts_;
extern struct { R SEADTM, DTMLAK; } datum_;
extern struct { R S[800], C[800], P[800], DELP[800]; } dummy4_;
extern struct { R SW[800], CW[800], W1218[2], WMAX[2]; } dumy44_;
extern struct { R X1[50], X12[50]; } frst_;
extern struct { R C1, C2, C21, C22, AX, AY, PTENCY, RTENCY; } strmsb_;
extern struct { R YDELP, ZDELP, PNN, YC24, ZC24; } wtemp_;
extern struct { R C7, C17, C19, C25; } thrd_;
extern struct { int ISTM, JHR, ITMADV, NHRAD, IBGNT, ITEND; } stime2_;
extern struct { int NDLTHR, NDLTH2; R DLTHR, DLTH2; } delth_;
extern struct { R DLTIN[3], DLTB; } dltdta_;
extern struct { R TREAL; } tmerel_;
extern struct { int KEY1; } keys_;
extern struct { int NSQRS, NSQRW, NSQRWC, NPSS, NCUT; } flwcpt_;
extern struct { R X[999], Y[999], PT[999], R_[999], DIR[999], SP[999]; } strmps_;
extern struct { R PHI, ALTO, ALNO, PHI1, ALT1, ALN1, ALT1C; } bsn_;
extern struct { R DEGREE, RMOUTH, AZMTH, DELA; } polar_;
extern struct { R AAXIS, BAXIS, ABQAB; } ellip_;
extern struct { R XMOUTH, YMOUTH; } orign_;
extern struct { R ZLATO; } zlato_;
extern struct { int IRANK, ISTAR, ISTOP, JSTAR, JSTOP, IHALO; } frank_;
extern struct { short MM2G, MCMX, I2G[30000], J2G[30000]; } sm2g_;
#define L_ 41000
#define NCT_ 3000
extern struct { R HWEIR[L_], ZBMIN[L_]; int ISQR[L_], JSQR[L_], ISIDE[L_]; } dumb18_;
extern struct { char KTREE[L_]; } ktree_;
extern struct { R FLWSQR[L_]; char F1DACT[L_]; } act1d_;
extern struct { R COS1D[L_], SIN1D[L_]; } cs1da_;
extern struct { R CUTL[NCT_], CUTLI[NCT_], CUTLE[NCT_]; } dumcut_;
extern struct { R DELCUT[L_], BANK[2][L_]; } bank2_;
extern struct { char NOFLD, NOF1D; } opts_;
extern struct { int IPT0[5][2], JPT0[5][2], IPTL[5][2], JPTL[5][2]; R ENTEXT[5]; } chnl_;
extern struct { R CHLNH[5], CHWTH[5], CHDPH[5]; } chnld_;
extern struct { R ARESQ0[5], ARESQL[5]; } chnara_;
extern struct { R AMAN, BMAN, GB, AGB, AGB13; } chnlno_;

extern void initalcommon_(void);
extern void inital_(int *mhalt, int *imxb, int *jmxb, R zb[N_][M_], int *hr,
                    int *mn, int *dy, int *mo, int *yr, char *trk, int *l1,
                    char *dta, int *l2, char *env, int *l3, char *ft40, int *l4,
                    R *delt, char *bsn, int *l5, R *clock, short *wave,
                    short *fulldep);
extern void tmstpc1_(R hb[N_][M_], short *fpass, short *restart);
extern void tmstpc2_(R hb[N_][M_], short *fpass, short *restart);
extern void tmstpm_(int *itime, int *mhalt, R *delt, R *lat, R *lon, R *delp,
                    R *size, R *wspd, R *wdir, R *clock, short *restart);
extern void clnup_(R hb[N_][M_], int *fsave);
extern int memSetLLx(char *buffer, int *Imxb, int *Jmxb);

static R hhb[N_][M_];
static R zbc[N_][M_];

/* ---- tagged record writer: name[16] type[1] pad[3] n(int64) data ---- */
static void rec(FILE *fp, const char *name, char type, long n, const void *data)
{
   char hdr[24];
   int esz = (type == 'd') ? 8 : (type == 'i') ? 4 : (type == 'h') ? 2 : 1;
   memset(hdr, 0, sizeof hdr);
   strncpy(hdr, name, 16);
   hdr[16] = type;
   fwrite(hdr, 1, 20, fp);
   fwrite(&n, 8, 1, fp);
   fwrite(data, esz, n, fp);
}

/* 2-D (N_,M_) Fortran array -> packed [jmxb][imxb] */
static void rec2(FILE *fp, const char *name, char type, const void *base)
{
   int imxb = dumb3_.IMXB, jmxb = dumb3_.JMXB, j;
   int esz = (type == 'd') ? 8 : (type == 'i') ? 4 : 1;
   char *buf = malloc((size_t) imxb * jmxb * esz);
   for (j = 0; j < jmxb; j++)
      memcpy(buf + (size_t) j * imxb * esz, (const char *) base + (size_t) j * M_ * esz,
             (size_t) imxb * esz);
   rec(fp, name, type, (long) imxb * jmxb, buf);
   free(buf);
}

static void dump_state(const char *path)
{
   FILE *fp = fopen(path, "wb");
   if (!fp) { perror(path); exit(2); }
   rec(fp, "ITIME", 'i', 1, &ffth_.ITIME);
   rec2(fp, "HB", 'd', dumb7_.HB);
   rec2(fp, "UB", 'd', dumb7_.UB);
   rec2(fp, "VB", 'd', dumb7_.VB);
   rec2(fp, "HMX", 'd', dumb10_.HMX);
   if (flwcpt_.NSQRWC > 0) rec(fp, "FLWSQR", 'd', flwcpt_.NSQRWC, act1d_.FLWSQR);
   fclose(fp);
}

static void dump_storm(FILE *fp)
{
   /* storm scalars as they stand after a step: what the next step consumes */
   R v[40];
   int n = 0, k;
   v[n++] = ffth_.ITIME;        v[n++] = frth_.BT;           v[n++] = egth_.DELT;
   v[n++] = strmsb_.C1;         v[n++] = strmsb_.C2;         v[n++] = strmsb_.AX;
   v[n++] = strmsb_.AY;         v[n++] = strmsb_.C21;        v[n++] = strmsb_.C22;
   v[n++] = wtemp_.YDELP;       v[n++] = wtemp_.ZDELP;       v[n++] = wtemp_.PNN;
   v[n++] = wtemp_.YC24;        v[n++] = wtemp_.ZC24;        v[n++] = tmerel_.TREAL;
   v[n++] = dumy44_.W1218[0];   v[n++] = dumy44_.W1218[1];   v[n++] = dumy44_.WMAX[0];
   v[n++] = dumy44_.WMAX[1];    v[n++] = stime2_.ITMADV;     v[n++] = stime2_.NHRAD;
   v[n++] = strmsb_.PTENCY;     v[n++] = strmsb_.RTENCY;     v[n++] = ffth_.MHALT;
   v[n++] = dumb4_.X1B[3];      v[n++] = dumb4_.X1B[4];      v[n++] = dumb4_.X1B[5];
   v[n++] = frth_.IC19;         v[n++] = datum_.SEADTM;      v[n++] = stime2_.JHR;
   v[n++] = zlato_.ZLATO;       v[n++] = dumb4_.X1B[0];      v[n++] = dumb4_.X1B[1];
   for (k = 0; k < 7; k++) v[n++] = 0;
   fwrite(v, 8, 40, fp);
   fwrite(frst_.X12, 8, 50, fp);
}

static void dump_tables(FILE *fp)
{
   rec(fp, "S", 'd', 800, dummy4_.S);
   rec(fp, "C", 'd', 800, dummy4_.C);
   rec(fp, "P", 'd', 800, dummy4_.P);
   rec(fp, "DELP", 'd', 800, dummy4_.DELP);
   rec(fp, "SW", 'd', 800, dumy44_.SW);
   rec(fp, "CW", 'd', 800, dumy44_.CW);
   rec(fp, "W1218", 'd', 2, dumy44_.W1218);
   rec(fp, "WMAX", 'd', 2, dumy44_.WMAX);
   rec(fp, "X12", 'd', 50, frst_.X12);
}

static void dump_init(const char *path)
{
   FILE *fp = fopen(path, "wb");
   int imxb = dumb3_.IMXB, jmxb = dumb3_.JMXB;
   int ints[64];
   R reals[64];
   int n;
   if (!fp) { perror(path); exit(2); }
   rec(fp, "DUMB3", 'i', 6, &dumb3_);
   rec2(fp, "ZB", 'd', dumb5_.ZB);
   rec2(fp, "ZBM", 'd', dumb5_.ZBM);
   rec2(fp, "ITREE", 'c', itree_.ITREE);
   rec2(fp, "IMANN", 'i', imann_.IMANN);
   rec2(fp, "KSKP", 'c', kspcon_.KSKP);
   rec2(fp, "YLT", 'd', llxfle_.YLT);
   rec2(fp, "YLG", 'd', llxfle_.YLG);
   rec2(fp, "HB", 'd', dumb7_.HB);
   rec2(fp, "UB", 'd', dumb7_.UB);
   rec2(fp, "VB", 'd', dumb7_.VB);
   rec2(fp, "HMX", 'd', dumb10_.HMX);
   rec(fp, "ELPCT", 'd', imxb, elpdst_.ELPCT);
   rec(fp, "ELPDT", 'd', imxb, elpdst_.ELPDT);
   rec(fp, "ELPCL", 'd', imxb, elpdst_.ELPCL);
   rec(fp, "ELPDL", 'd', imxb, elpdst_.ELPDL);
   rec(fp, "ELPCL2", 'd', imxb, elpds2_.ELPCL2);
   rec(fp, "ELPDL2", 'd', imxb, elpds2_.ELPDL2);
   rec(fp, "SINL2", 'd', jmxb, elpds2_.SINL2);
   rec(fp, "SINT2", 'd', jmxb, elpds2_.SINT2);
   rec(fp, "SINT", 'd', jmxb, plrpst_.SINT);
   rec(fp, "COST", 'd', jmxb, plrpst_.COST);
   rec(fp, "SINL", 'd', jmxb, plrpst_.SINL);
   rec(fp, "COSL", 'd', jmxb, plrpst_.COSL);
   rec(fp, "IS", 'i', jmxb, setmix_.IS);
   rec(fp, "MS", 'i', jmxb, setmix_.MS);
   rec(fp, "MF", 'i', jmxb, setmix_.MF);
   rec(fp, "IE", 'i', jmxb, setmx1_.IE);
   rec(fp, "ME", 'i', jmxb, setmx1_.ME);
   rec(fp, "X1B", 'd', 10, dumb4_.X1B);
   rec(fp, "CORL", 'd', 90, dumb4i_.CORL);
   rec(fp, "AR2", 'd', 600 * 245, scnd2_.AR2);
   rec(fp, "AI2", 'd', 600 * 245, scnd2_.AI2);
   rec(fp, "BR2", 'd', 600 * 245, scnd2_.BR2);
   rec(fp, "BI2", 'd', 600 * 245, scnd2_.BI2);
   rec(fp, "CR2", 'd', 600 * 245, scnd2_.CR2);
   rec(fp, "CI2", 'd', 600 * 245, scnd2_.CI2);
   rec(fp, "GAMAF", 'd', 6, &gamaf_);
   rec(fp, "EGTH", 'd', 4, &egth_);
   rec(fp, "THRD", 'd', 4, &thrd_);
   rec(fp, "BSN", 'd', 7, &bsn_);
   rec(fp, "POLAR", 'd', 4, &polar_);
   rec(fp, "ELLIP", 'd', 3, &ellip_);
   rec(fp, "ORIGN", 'd', 2, &orign_);
   rec(fp, "DATUM", 'd', 2, &datum_);
   rec(fp, "DLTDTA", 'd', 4, &dltdta_);
   rec(fp, "HTER", 'd', 1, &hterain_.HTER);
   rec(fp, "NBCPT", 'i', 1, &bcpts_.NBCPT);
   rec(fp, "ISH", 'i', bcpts_.NBCPT, bcpts_.ISH);
   rec(fp, "JSH", 'i', bcpts_.NBCPT, bcpts_.JSH);
   rec(fp, "TIDESH", 'd', bcpts_.NBCPT, bcpts_.TIDESH);
   n = 0;
   ints[n++] = frth_.IC12;  ints[n++] = frth_.IC19;  ints[n++] = ffth_.ITIME;
   ints[n++] = ffth_.MHALT; ints[n++] = dtaopt_.IVER; ints[n++] = dtaopt_.IPRJ;
   ints[n++] = smth_.ISMTH; ints[n++] = keys_.KEY1;
   ints[n++] = stime2_.ISTM; ints[n++] = stime2_.JHR; ints[n++] = stime2_.ITMADV;
   ints[n++] = stime2_.NHRAD; ints[n++] = stime2_.IBGNT; ints[n++] = stime2_.ITEND;
   ints[n++] = delth_.NDLTHR; ints[n++] = delth_.NDLTH2;
   ints[n++] = flwcpt_.NSQRS; ints[n++] = flwcpt_.NSQRW; ints[n++] = flwcpt_.NSQRWC;
   ints[n++] = flwcpt_.NPSS; ints[n++] = flwcpt_.NCUT;
   rec(fp, "INTS", 'i', n, ints);
   n = 0;
   reals[n++] = frth_.BT;   reals[n++] = delth_.DLTHR; reals[n++] = delth_.DLTH2;
   reals[n++] = tmerel_.TREAL; reals[n++] = zlato_.ZLATO;
   rec(fp, "REALS", 'd', n, reals);
   {
      char flags[8];
      flags[0] = gprt1_.DOLLAR[0]; flags[1] = gprt1_.DOLLAR[1];
      flags[2] = gprt1_.EBSN; flags[3] = gprt2_.EBSN1; flags[4] = gprt2_.EBSN2;
      flags[5] = opts_.NOF1D;
      flags[6] = flags[7] = ' ';
      rec(fp, "FLAGS", 'c', 8, flags);
   }
   {
      /* sub-grid 1-D flow edges (COMMON /DUMB18/ /KTREE/ /CS1DA/ /BANK2/ /DUMCUT/) */
      int nl = flwcpt_.NSQRWC, nc = flwcpt_.NCUT;
      rec(fp, "EDGE_I", 'i', nl, dumb18_.ISQR);
      rec(fp, "EDGE_J", 'i', nl, dumb18_.JSQR);
      rec(fp, "EDGE_SIDE", 'i', nl, dumb18_.ISIDE);
      rec(fp, "KTREE", 'c', nl, ktree_.KTREE);
      rec(fp, "HWEIR", 'd', nl, dumb18_.HWEIR);
      rec(fp, "ZBMIN", 'd', nl, dumb18_.ZBMIN);
      rec(fp, "BANK1", 'd', nl, bank2_.BANK[0]);
      rec(fp, "BANK2", 'd', nl, bank2_.BANK[1]);
      rec(fp, "DELCUT", 'd', nl, bank2_.DELCUT);
      rec(fp, "COS1D", 'd', nl, cs1da_.COS1D);
      rec(fp, "SIN1D", 'd', nl, cs1da_.SIN1D);
      rec(fp, "CUTL", 'd', nc, dumcut_.CUTL);
      rec(fp, "CUTLI", 'd', nc, dumcut_.CUTLI);
      rec(fp, "CUTLE", 'd', nc, dumcut_.CUTLE);
      rec(fp, "CHNL_PTS", 'i', 40, chnl_.IPT0);
      rec(fp, "CHNLD", 'd', 15, &chnld_);
      rec(fp, "CHNARA", 'd', 10, &chnara_);
      rec(fp, "CHNLNO", 'd', 5, &chnlno_);
   }
   rec(fp, "STA", 'c', 16, gprt_.STA);
   rec(fp, "STRMSB", 'd', 8, &strmsb_);
   rec(fp, "WTEMP", 'd', 5, &wtemp_);
   rec(fp, "TRK_X", 'd', 999, strmps_.X);
   rec(fp, "TRK_Y", 'd', 999, strmps_.Y);
   rec(fp, "TRK_PT", 'd', 999, strmps_.PT);
   rec(fp, "TRK_R", 'd', 999, strmps_.R_);
   rec(fp, "TRK_DIR", 'd', 999, strmps_.DIR);
   rec(fp, "TRK_SP", 'd', 999, strmps_.SP);
   dump_tables(fp);
   fclose(fp);
}

static int want(const char *spec, int step)
{
   char buf[4096], *tok;
   if (!spec || !*spec) return 0;
   if (strcmp(spec, "all") == 0) return 1;
   strncpy(buf, spec, sizeof buf - 1);
   buf[sizeof buf - 1] = 0;
   for (tok = strtok(buf, ","); tok; tok = strtok(NULL, ","))
      if (atoi(tok) == step) return 1;
   return 0;
}

int main(int argc, char **argv)
{
   char line[512], trk[257], dta[257], env[257], ft40[257], bsn[5], path[1024];
   FILE *fp, *fs, *ft;
   int imxb, jmxb, mhalt, hr, mn, dy, mo, yr, l1, l2, l3, l4, l5, nsteps, k, fsave;
   R delt, clock = 0, lat, lon, delp, size, wspd, wdir;
   short wave = 0, fulldep = 0, restart = 0, fpass;
   const char *outdir, *spec;
   static R last[4 * 800 + 2 * 800];

   if (argc < 7) {
      fprintf(stderr, "usage: %s bntline dta trk outdir nsteps dumpspec\n", argv[0]);
      return 1;
   }
   fp = fopen(argv[1], "r");
   if (!fp || !fgets(line, sizeof line, fp)) { perror(argv[1]); return 1; }
   fclose(fp);
   line[strcspn(line, "\n")] = 0;
   outdir = argv[4];
   nsteps = atoi(argv[5]);
   spec = argv[6];

   initalcommon_();                                  /* setup.c:115 */
   if (memSetLLx(line, &imxb, &jmxb) != 0) return 1; /* setup.c:29-81 */
   frank_.IRANK = 0;                                 /* RunInit, slosh2.c:1759-1764 */
   frank_.ISTAR = 1; frank_.ISTOP = imxb;
   frank_.JSTAR = 1; frank_.JSTOP = jmxb;
   frank_.IHALO = 7;
   memset(trk, ' ', sizeof trk); memset(dta, ' ', sizeof dta);
   memset(env, ' ', sizeof env); memset(ft40, ' ', sizeof ft40);
   strcpy(trk, argv[3]); strcpy(dta, argv[2]); strcpy(env, "none.env"); strcpy(ft40, "none");
   strcpy(bsn, "TST");
   l1 = strlen(trk); l2 = strlen(dta); l3 = strlen(env); l4 = strlen(ft40); l5 = strlen(bsn);
   inital_(&mhalt, &imxb, &jmxb, zbc, &hr, &mn, &dy, &mo, &yr, trk, &l1, dta, &l2,
           env, &l3, ft40, &l4, &delt, bsn, &l5, &clock, &wave, &fulldep);
   snprintf(path, sizeof path, "%s/init.bin", outdir);
   dump_init(path);
   if (want(spec, 0)) {
      snprintf(path, sizeof path, "%s/s%06d.bin", outdir, 0);
      dump_state(path);
   }
   if (nsteps < 0 || nsteps > mhalt) nsteps = mhalt;
   snprintf(path, sizeof path, "%s/storm.bin", outdir);
   fs = fopen(path, "wb");
   snprintf(path, sizeof path, "%s/tables.bin", outdir);
   ft = fopen(path, "wb");
   dump_storm(fs);
   memcpy(last, &dummy4_, sizeof(R) * 3200);
   memcpy(last + 3200, &dumy44_, sizeof(R) * 1600);
   for (k = 1; k <= nsteps; k++) {
      fpass = 0;
      tmstpc1_(hhb, &fpass, &restart);               /* slosh2.c:1400 */
      tmstpc2_(hhb, &fpass, &restart);               /* slosh2.c:1540 */
      tmstpm_(&ffth_.ITIME, &mhalt, &delt, &lat, &lon, &delp, &size, &wspd, &wdir,
              &clock, &restart);                     /* slosh2.c:1542 */
      dump_storm(fs);
      if (memcmp(last, &dummy4_, sizeof(R) * 3200) || memcmp(last + 3200, &dumy44_, sizeof(R) * 1600)) {
         int itime = ffth_.ITIME;
         rec(ft, "STEP", 'i', 1, &itime);
         dump_tables(ft);
         memcpy(last, &dummy4_, sizeof(R) * 3200);
         memcpy(last + 3200, &dumy44_, sizeof(R) * 1600);
      }
      if (want(spec, k)) {
         snprintf(path, sizeof path, "%s/s%06d.bin", outdir, k);
         dump_state(path);
      }
   }
   fclose(fs);
   fclose(ft);
   fsave = 0;
   clnup_(hhb, &fsave);                              /* slosh2.c:1701 */
   snprintf(path, sizeof path, "%s/final.bin", outdir);
   fp = fopen(path, "wb");
   rec(fp, "ITIME", 'i', 1, &ffth_.ITIME);
   rec2(fp, "HMX", 'd', dumb10_.HMX);
   rec2(fp, "HB", 'd', dumb7_.HB);
   rec2(fp, "UB", 'd', dumb7_.UB);
   rec2(fp, "VB", 'd', dumb7_.VB);
   fclose(fp);
   printf("refdrv: imxb=%d jmxb=%d mhalt=%d steps=%d delt=%g\n", dumb3_.IMXB,
          dumb3_.JMXB, mhalt, nsteps, (double) egth_.DELT);
   return 0;
}
