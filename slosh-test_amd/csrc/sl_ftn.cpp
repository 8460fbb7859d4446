// sl_ftn.cpp -- the reference's Fortran entry points and COMMON blocks, served
// by the HIP time step (libslosh_ftn.so).
//
// The reference's C driver (slosh2.c: RunInit / RunLoopStep / RunCleanup, and
// the CLI in cstart.c) calls the Fortran model through INITAL, TMSTPC1, TMSTPC2,
// TMSTPM and CLNUP and reads a few COMMON blocks directly (slosh2.c:24-303,
// savellx.c:686-700).  Linking that driver against this library instead of
// the Fortran objects runs the same CLI -- dta/bnt/trk in, env and rex files
// out, written by the C driver's own EnvSave / RexSaveStep -- with the time step
// on the GPU.  The ABI is the reference's default build (-DDOUBLE_FORTRAN):
// double arrays [BAS_Y][BAS_X] = [1655][999], #pragma pack(2) structs.
//
// Coherence with the C driver (what it reads and when):
//   DUMB5.ZB   after INITAL                      -> filled by inital_
//   DUMB7.HB   on the first step (mask init)     -> initial water, by inital_
//   hb (HHB)   after TMSTPC2 when FPASS          -> min(HB, HMX) after HMXSV
//   hb         after CLNUP                       -> HMX (the envelope)
//   DUMB7.HB   after every step (-f_stat reads it) -> HB, by tmstpc2_ (see below)
//   EGTH.DELT  on tide-only steps                -> kept current by tmstpm_
//   LLXFLE     written by memSetLLx before INITAL -> read by inital_ (Coriolis)
// Paths this build does not model (tides, hot start, waves, full depth) stop
// the run with a message, as the reference's Fortran STOPs on fatal input.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include "../../include/slosh_hip.h"
#include "../../include/slosh_ftn.h"

#define BAS_X 999
#define BAS_Y 1655
#define NBCPTS 12000

// COMMON blocks (parm.for; C views slosh2.c:175-303, savellx.c:686-700)
#pragma pack(push, 2)
extern "C" {
struct {
   double ZB[BAS_Y][BAS_X], ZBM[BAS_Y][BAS_X], D_WAV[BAS_Y][BAS_X];
} dumb5_;
struct {
   double UB[BAS_Y][BAS_X], VB[BAS_Y][BAS_X], HB[BAS_Y][BAS_X];
} dumb7_;
struct {
   double HMX[BAS_Y][BAS_X];
} dumb10_;
struct {
   double DELS, DELT, G, COR;
} egth_;
struct {
   int IRANK, ISTAR, ISTOP, JSTAR, JSTOP, IHALO;
} frank_;
struct {
   double TIDESH[NBCPTS];
   int NBCPT, ISH[NBCPTS], JSH[NBCPTS];
} bcpts_;
struct {
   double YLT[BAS_Y][BAS_X], YLG[BAS_Y][BAS_X];
} llxfle_;
}
#pragma pack(pop)

namespace {

struct Run {
   slg_basin *basin = nullptr;
   slg_batch *batch = nullptr;
   int imxb = 0, jmxb = 0, nsteps = 0, step = 0;
   bool hb_each_step = false;   // the driver reads COMMON HB after every step
   std::vector<double> buf;     // [jmxb][imxb] staging
   std::vector<double> hb_last; // COMMON HB as this library last wrote it
   std::vector<double> tide_last;   // TIDESH as last passed to the device
};
Run g_run;

[[noreturn]] void fatal(const char *what, const char *detail)
{
   fprintf(stderr, "slosh_ftn: %s%s%s\n", what, detail && *detail ? ": " : "", detail ? detail : "");
   fflush(stderr);
   exit(1);
}

std::string fstr(const char *s, int len)
{
   std::string r(s, s + (len > 0 ? len : 0));
   while (!r.empty() && (r.back() == ' ' || r.back() == '\0')) r.pop_back();
   return r;
}

void get(const char *field)
{
   char err[256];
   if (slg_batch_get(g_run.batch, 0, field, g_run.buf.data(), err, sizeof err)) fatal("field copy failed", err);
}

// COMMON HB from the staging buffer; what was written is remembered, so that a
// change the driver makes before the next step can be told apart
void mirror_hb()
{
   const int ni = g_run.imxb, nj = g_run.jmxb;
   for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) dumb7_.HB[j][i] = g_run.buf[(size_t) j * ni + i];
   g_run.hb_last = g_run.buf;
}

// Before a step: the C driver's tide coupling (slosh2.c:813-1027) changes
// COMMON HB (the tide added to every wet square) and TIDESH (the tide at the
// BDRYHT squares) between steps, as the reference's Fortran sees them in
// COMMON; both go to the device when they changed.  HB is compared only while
// the mirror is kept (without it COMMON HB is not the model's).
void take_common()
{
   char err[512];
   const int ni = g_run.imxb, nj = g_run.jmxb;
   if (g_run.hb_each_step) {
      bool changed = false;
      for (int j = 0; j < nj && !changed; j++)
         changed = memcmp(dumb7_.HB[j], &g_run.hb_last[(size_t) j * ni], sizeof(double) * ni) != 0;
      if (changed) {
         for (int j = 0; j < nj; j++)
            for (int i = 0; i < ni; i++) g_run.buf[(size_t) j * ni + i] = dumb7_.HB[j][i];
         if (slg_batch_put(g_run.batch, 0, "HB", g_run.buf.data(), err, sizeof err)) fatal("HB", err);
         g_run.hb_last = g_run.buf;
      }
   }
   const int n = bcpts_.NBCPT;
   if (memcmp(bcpts_.TIDESH, g_run.tide_last.data(), sizeof(double) * n)) {
      memcpy(g_run.tide_last.data(), bcpts_.TIDESH, sizeof(double) * n);
      if (slg_batch_set_boundary_tide(g_run.batch, 0, g_run.tide_last.data(), n, err, sizeof err))
         fatal("TIDESH", err);
   }
}

}   // namespace

extern "C" {

int initalcommon_(void)
{
   return 0;   // the mirrors start zeroed
}

// COMMON /DUMB7/ HB holds the step's result after every step, as the
// reference's Fortran leaves it.  The C driver reads it between steps in the
// -f_stat mode (MaxStatUpdateGrid, slosh2.c:1264-1316, called from
// RunLoopStep whatever the caller: the CLI, list runs, the Tcl binding), and
// the model is not told that mode is on, so the mirror is kept after every
// step.  SLG_HB_MIRROR=0 drops it for drivers known not to read HB between
// steps (a copy of one field per step saved).
static bool hb_mirror_on()
{
   const char *e = getenv("SLG_HB_MIRROR");
   return !(e && e[0] == '0' && e[1] == 0);
}

int inital_(int *mhalt, int *imxb, int *jmxb, double zb[BAS_Y][BAS_X], int *hr, int *mn, int *dy,
            int *mo, int *yr, char *trk, int *len_trk, char *dta, int *len_dta, char *env, int *len_env,
            char *ft40, int *len_ft40, double *delt, char *bsn, int *len_bsn, double *clock, short *wave,
            short *full_depth)
{
   (void) env;
   (void) len_env;
   (void) ft40;
   (void) len_ft40;
   (void) bsn;
   (void) len_bsn;
   char err[512];
   if (wave && *wave) fatal("the wave model is not part of this build", "");
   const std::string dta_path = fstr(dta, *len_dta), trk_path = fstr(trk, *len_trk);
   if (slg_basin_open_opts(dta_path.c_str(), nullptr, (full_depth && *full_depth) ? SLG_OPEN_FULL_DEPTH : 0,
                           &g_run.basin, err, sizeof err))
      fatal("basin", err);
   int ni, nj;
   slg_basin_dims(g_run.basin, &ni, &nj, nullptr);
   if (ni > BAS_X || nj > BAS_Y) fatal("basin larger than the compiled dimensions", dta_path.c_str());
   g_run.imxb = ni;
   g_run.jmxb = nj;
   g_run.buf.assign((size_t) ni * nj, 0.);
   // cell latitudes memSetLLx left in LLXFLE (savellx.c:752-762)
   for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) g_run.buf[(size_t) j * ni + i] = llxfle_.YLT[j][i];
   slg_basin_set_ylt(g_run.basin, g_run.buf.data());
   const char *trks[1] = {trk_path.c_str()};
   // The step runs in double, the reference default build's arithmetic (its env
   // and rex files then match that build's byte for byte); SLG_PRECISION=32
   // runs the FP32 step instead
   const char *pe = getenv("SLG_PRECISION");
   const int prec = (pe && atoi(pe) == 32) ? 32 : 64;
   // the per-step storm values stay on the host: STATICHT reads them there
   if (slg_batch_open_opts(g_run.basin, 1, trks, prec, SLG_SCHED_HOST, &g_run.batch, err, sizeof err))
      fatal("track", err);
   if (slg_batch_upload(g_run.batch, err, sizeof err)) fatal("device", err);
   g_run.nsteps = slg_batch_nsteps(g_run.batch, 0);
   g_run.step = 0;
   g_run.hb_each_step = hb_mirror_on();
   // outputs (inital.f:204-272)
   *mhalt = g_run.nsteps;
   *imxb = ni;
   *jmxb = nj;
   slg_basin_field(g_run.basin, "ZB", g_run.buf.data());
   for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) {
         zb[j][i] = g_run.buf[(size_t) j * ni + i];
         dumb5_.ZB[j][i] = zb[j][i];
      }
   slg_basin_field(g_run.basin, "ZBM", g_run.buf.data());
   for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) dumb5_.ZBM[j][i] = g_run.buf[(size_t) j * ni + i];
   // COMMON HB: the initial water as the step holds it (in the FP32 step,
   // rounded to float), which the driver's first tide call adds to
   get("HB");
   mirror_hb();
   // the BDRYHT squares (DEPSFC, depsfc.f:376-390) for the driver's tide
   // coupling, with no tide yet
   bcpts_.NBCPT = slg_basin_boundary(g_run.basin, bcpts_.ISH, bcpts_.JSH, NBCPTS);
   if (bcpts_.NBCPT > NBCPTS) fatal("basin", "more than 12000 boundary squares");
   memset(bcpts_.TIDESH, 0, sizeof bcpts_.TIDESH);
   g_run.tide_last.assign(NBCPTS, 0.);
   int d[5];
   slg_batch_start_date(g_run.batch, 0, d);
   *yr = d[0];
   *mo = d[1];
   *dy = d[2];
   *hr = d[3];
   *mn = d[4];
   double t[9];
   slg_batch_track(g_run.batch, 0, 0, t);
   *delt = t[2];
   *clock = 0.;   // DMCLOCK: no hot start
   egth_.DELS = 5280.;
   egth_.DELT = t[2];
   egth_.G = 32.2;
   return 0;
}

// TMSTPC1: the continuity stage runs with the rest of the step in tmstpc2_
// (nothing the C driver reads changes in between).
int tmstpc1_(double hb[BAS_Y][BAS_X], short *fpass, short *restart)
{
   (void) hb;
   (void) fpass;
   if (restart && *restart) fatal("restart files are not supported by this build", "");
   if (!g_run.batch) fatal("TMSTPC1 before INITAL", "");
   return 0;
}

// TMSTPC2: advance the GPU state through the whole step (CONTTY .. MOMNTM);
// when FPASS, hand back HHB = min(HB, HMX) of HB after HMXSV (tmstpc2.f:70-85).
int tmstpc2_(double hb[BAS_Y][BAS_X], short *fpass, short *restart)
{
   (void) restart;
   char err[512];
   const bool pass = fpass && *fpass;
   take_common();
   // the staged path keeps HB before FILTER, which HHB needs on filter steps
   slg_batch_set_path(g_run.batch, pass ? SLG_PATH_STAGED : SLG_PATH_FUSED, 16);
   if (slg_batch_step(g_run.batch, 1, err, sizeof err)) fatal("step", err);
   g_run.step++;
   if (pass) {
      const int ni = g_run.imxb, nj = g_run.jmxb;
      std::vector<double> hmx((size_t) ni * nj);
      if (slg_batch_get(g_run.batch, 0, "HMX", hmx.data(), err, sizeof err)) fatal("HMX", err);
      get("HBS");
      for (int j = 0; j < nj; j++)
         for (int i = 0; i < ni; i++) {
            const double h = g_run.buf[(size_t) j * ni + i], m = hmx[(size_t) j * ni + i];
            hb[j][i] = (m < h) ? m : h;
         }
   }
   if (g_run.hb_each_step) {   // HB at the end of the step, as COMMON holds it after TMSTPM
      get("HB");
      mirror_hb();
   }
   return 0;
}

// TMSTPM: the momentum stage ran in tmstpc2_; return the storm values
// (tmstpm.f:59-70) from the precomputed schedule.
int tmstpm_(int *itime, int *mhalt, double *delt, double *lat, double *lon, double *delp, double *size,
            double *wspd, double *wdir, double *clock, short *restart)
{
   (void) clock;
   (void) restart;
   double t[9];
   if (slg_batch_track(g_run.batch, 0, g_run.step, t)) fatal("storm schedule", "");
   *itime = (int) t[0];
   *mhalt = (int) t[1];
   *delt = t[2];
   *lat = t[3];
   *lon = t[4];
   *delp = t[5];
   *size = t[6];
   *wspd = t[7];
   *wdir = t[8];
   egth_.DELT = t[2];
   return 0;
}

// CLNUP: the envelope HMX to the driver (clnup.f:13-24); ARCHIV (iENV == 1) is
// the legacy Fortran envelope file, which the C driver never asks for.
int clnup_(double hb[BAS_Y][BAS_X], int *f_save_env)
{
   if (f_save_env && *f_save_env) fatal("the Fortran envelope archive (ARCHIV) is not supported", "");
   get("HMX");
   const int ni = g_run.imxb, nj = g_run.jmxb;
   for (int j = 0; j < nj; j++)
      for (int i = 0; i < ni; i++) {
         hb[j][i] = g_run.buf[(size_t) j * ni + i];
         dumb10_.HMX[j][i] = hb[j][i];
      }
   slg_batch_close(g_run.batch);
   slg_basin_close(g_run.basin);
   g_run.batch = nullptr;
   g_run.basin = nullptr;
   return 0;
}

// STATICHT (staticht.f:14-41), called by the tide coupling between steps: the
// static height at (i,j) with the storm values of the next step.
int staticht_(int *i, int *j, double *ans)
{
   if (slg_batch_static_height(g_run.batch, 0, *i, *j, ans)) fatal("STATICHT", "square outside the basin");
   return 0;
}
// ISSUBGRID (issubgrid.f:14-33): whether (i,j) is a narrowed 1-D side square.
int issubgrid_(int *i, int *j, int *ans)
{
   *ans = slg_basin_is_subgrid(g_run.basin, *i, *j);
   return 0;
}
// The tide spin-up (SpinUpTideGrid, slosh2.c:1029-1225: TMSTP2, INTLHTSUB,
// INTLHTADD) is not part of this build: its momentum step MOMNTM2 adds the
// wind-forcing terms XTTX/XTTY to every wet point but, with the forcing
// commented out, sets them only on partially flooded points (momntm2.f:
// 352-372), so it reads them uninitialised and has no defined result to match.
int tmstp2_(void)
{
   fatal("tide spin-up (TMSTP2) is not supported by this build", "MOMNTM2 reads XTTX/XTTY uninitialised");
}
int intlhtsub_(void)
{
   fatal("tide spin-up (INTLHTSUB) is not supported by this build", "");
}
int intlhtadd_(void)
{
   fatal("tide spin-up (INTLHTADD) is not supported by this build", "");
}
// CLN_HS (cln_hs.f) deletes the hot-start files; this build writes none.  The
// driver calls it at the end of every run (slosh2.c:2771).
int cln_hs_(void)
{
   return 0;
}

}   // extern "C"
