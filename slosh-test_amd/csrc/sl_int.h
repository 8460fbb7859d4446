// sl_int.h -- internal data of the C ABI (basins, storm schedules, batches),
// shared by sl_capi.cpp (the batch API) and sl_serve.cpp (the serving loop).
#ifndef SL_INT_H
#define SL_INT_H
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <vector>
#include <string>
#include <thread>
#include <algorithm>
#include "sl_host.h"
#include "sl_dev.h"
#include "../../include/slosh_hip.h"

extern "C" int sl_dev_launch_step_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, int smpt,
                                      int filt, int dyn, int parity, int path, int bth, hipStream_t st,
                                      hipEvent_t *ev);
extern "C" int sl_dev_launch_step_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, int smpt,
                                      int filt, int dyn, int parity, int path, int bth, hipStream_t st,
                                      hipEvent_t *ev);
static inline int dev_launch(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, int smpt, int filt, int dyn,
                      int parity, int path, int bth, hipStream_t st, hipEvent_t *ev)
{
   return sl_dev_launch_step_f32(B, R, smpt, filt, dyn, parity, path, bth, st, ev);
}
static inline int dev_launch(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, int smpt, int filt, int dyn,
                      int parity, int path, int bth, hipStream_t st, hipEvent_t *ev)
{
   return sl_dev_launch_step_f64(B, R, smpt, filt, dyn, parity, path, bth, st, ev);
}

#define PIX(b, i, j) ((size_t) ((j) - 1) * (size_t) (b)->imxb + (size_t) ((i) - 1))

static inline void seterr(char *err, int len, const char *fmt, const char *a = "", int v = 0)
{
   if (err && len > 0) snprintf(err, (size_t) len, fmt, a, v);
}

#define HIPOK(x)                                                                           \
   do {                                                                                    \
      hipError_t e_ = (x);                                                                 \
      if (e_ != hipSuccess) {                                                              \
         seterr(err, errlen, "HIP error %s (%d)", hipGetErrorString(e_), (int) e_);         \
         return SLG_EHIP;                                                                  \
      }                                                                                    \
   } while (0)

// Host images of the device static data in one precision, and their device copy.
template <class R>
struct BasinImg {
   std::vector<R> zb, zbm, zcor, vi[6], vj[5], tab, slpai, tab8, g4;
   std::vector<uint32_t> flags;
   std::vector<int> eint;          // 1-D flow sides [nsqrwc][SL_EI]
   std::vector<R> ereal, ecut;     // [nsqrwc][8], [ncut][4]
   int nsq = 0;                    // distinct squares the sides join
   bool built = false, on_dev = false;
   int lat_i = 0, lat_j = 0;       // first momentum point outside the Coriolis table
   double lat_v = 0;
   void *dmem = nullptr;
   SlBasinDevT<R> dev;
   const double *zbd = nullptr;   // device ZB in double (envelope dry test, as EnvSave)
};

struct slg_basin {
   sl_basin h;
   char bnt[512];
   int pitch, jend;
   int island;                         // DOLLAR '$ ': closed island, periodic in J
   std::vector<int> is, ie, ms, me;
   int incsm, ismpt2g;
   BasinImg<float> i32;
   BasinImg<double> i64;
};
static inline BasinImg<float> &img(slg_basin *B, float) { return B->i32; }
static inline BasinImg<double> &img(slg_basin *B, double) { return B->i64; }

struct storm_sched {
   sl_storm s0;                        // storm state after initialisation
   std::vector<double> rows;           // [nsteps+1][90] ref-format rows (kept for export)
   // device values of each step ([nsteps][SL_NPAR]) and the table sets
   // ([ntab][SL_NTABV][8], sl_dev.h), in the batch's precision
   std::vector<float> p32, t32;
   std::vector<double> p64, t64;
   std::vector<int> tab_step;          // step after which each table set took effect
   int nsteps;
   bool hostp = false;                 // p32 / p64 filled (else the device generates them)
   std::vector<double> prog;           // the storm program (sl_dev.h) the device generates them from
};

struct slg_batch {
   slg_basin *b;
   int prec;                           // 32 or 64
   int nstorm, maxsteps, step, parity;
   int path, tile_rows;                // launch path (SLG_PATH_*) and stage-B tile height
   int hbs_buf;                        // buffer holding HB before FILTER of the last step (-1: none)
   std::vector<int> last_par;          // per storm: buffer holding its HB after its last step
   int graph_steps = 0;                // > 0: replay the step loop as hipGraphs of this many steps
   hipGraph_t graph = nullptr;
   hipGraphExec_t gexec = nullptr;
   std::vector<storm_sched> sc;
   std::vector<double> hb0;            // initial HB per storm (double, imxb*jmxb)
   void *dmem;
   SlRunDevT<float> d32;
   SlRunDevT<double> d64;
   hipStream_t st;
   size_t bytes;
   // the batch in nsplit parts of storms on as many streams (fused path): one
   // part's latency-bound cascade replay overlaps the others' stencil kernels.
   // Part 0 runs on st, part p > 0 on stx[p - 1].
   int nsplit = 1;
   double *hbmax = nullptr;   // P-Surge windowed maxima, [nstorm][fstride], when on
   double *hbmax2 = nullptr;  // their second window (-envSave2), when on
   // rex frame buffers, kept from one slg_batch_rex_frames call to the next:
   // scratch, frames, packed frames (device), sizes and offsets, and pinned
   // host staging for the packed copy; for rex_chunk storms of rex_cap bytes
   void *rex_scr = nullptr, *rex_out = nullptr, *rex_pack = nullptr, *rex_nb = nullptr, *rex_off = nullptr;
   unsigned char *rex_host = nullptr;
   int rex_chunk = 0;
   long rex_cap = 0;
   // device buffer of slg_batch_rex_points: square indices, then the values
   void *pt_buf = nullptr;
   size_t pt_cap = 0;
   hipStream_t stx[SL_MAXSPLIT - 1] = {};
   hipEvent_t ev_join[SL_MAXSPLIT - 1] = {};
   hipEvent_t ev_phase[SL_MAXSPLIT][SLG_NKERN + 1] = {};   // marks of each part's first step
   // the gridded tide between steps (slg_batch_set_tide): mode 2, 21, 22 or
   // 24, or 0; its tables and fields on the device (tide_mem), the host's
   // copy of the harmonic constants until the upload
   int tide_mode = 0;
   void *tide_mem = nullptr;
   SlTideDev tdev = {};
};
static inline SlRunDevT<float> &run(slg_batch *X, float) { return X->d32; }
static inline SlRunDevT<double> &run(slg_batch *X, double) { return X->d64; }


extern "C" int sl_dev_envelopes_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const double *hx2,
                                    const double *zbd, int s0, int ns, short *env, short *mom, int acc,
                                    hipStream_t st);
extern "C" int sl_dev_envelopes_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hx2,
                                    const double *zbd, int s0, int ns, short *env, short *mom, int acc,
                                    hipStream_t st);
// env values (EnvSave) of HMX, or of hx2 (double [nstorm][fstride]) when given;
// with acc the MOM continues from what mom holds
static inline int dev_envelopes(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const double *zbd, int s0,
                                int ns, short *env, short *mom, hipStream_t st, const double *hx2 = nullptr,
                                int acc = 0)
{
   return sl_dev_envelopes_f32(B, R, hx2, zbd, s0, ns, env, mom, acc, st);
}
static inline int dev_envelopes(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *zbd, int s0,
                                int ns, short *env, short *mom, hipStream_t st, const double *hx2 = nullptr,
                                int acc = 0)
{
   return sl_dev_envelopes_f64(B, R, hx2, zbd, s0, ns, env, mom, acc, st);
}
extern "C" int sl_dev_maxstat_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hb,
                                  const double *zbd, double *hmax, double *hmax2, int init, hipStream_t st);
extern "C" int sl_dev_maxstat_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hb,
                                  const double *zbd, double *hmax, double *hmax2, int init, hipStream_t st);
static inline int dev_maxstat(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hb,
                              const double *zbd, double *hmax, double *hmax2, int init, hipStream_t st)
{
   return sl_dev_maxstat_f32(B, R, hb, zbd, hmax, hmax2, init, st);
}
static inline int dev_maxstat(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hb,
                              const double *zbd, double *hmax, double *hmax2, int init, hipStream_t st)
{
   return sl_dev_maxstat_f64(B, R, hb, zbd, hmax, hmax2, init, st);
}
extern "C" int sl_dev_rex_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hbs,
                              const double *hmax, const double *zbd, int s0, int ns, unsigned char *out, long cap,
                              int *nbytes, unsigned *scr, hipStream_t st);
extern "C" int sl_dev_rex_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hbs,
                              const double *hmax, const double *zbd, int s0, int ns, unsigned char *out, long cap,
                              int *nbytes, unsigned *scr, hipStream_t st);
extern "C" int sl_dev_rex_points_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hbs,
                                     const double *hmax, const double *zbd, int s0, int ns, const int *ks, int npts,
                                     int16_t *out, hipStream_t st);
extern "C" int sl_dev_rex_points_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hbs,
                                     const double *hmax, const double *zbd, int s0, int ns, const int *ks, int npts,
                                     int16_t *out, hipStream_t st);
static inline int dev_rex_points(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hbs,
                                 const double *hmax, const double *zbd, int s0, int ns, const int *ks, int npts,
                                 int16_t *out, hipStream_t st)
{
   return sl_dev_rex_points_f32(B, R, hbs, hmax, zbd, s0, ns, ks, npts, out, st);
}
static inline int dev_rex_points(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hbs,
                                 const double *hmax, const double *zbd, int s0, int ns, const int *ks, int npts,
                                 int16_t *out, hipStream_t st)
{
   return sl_dev_rex_points_f64(B, R, hbs, hmax, zbd, s0, ns, ks, npts, out, st);
}
extern "C" int sl_dev_rex_pack(const unsigned char *src, long cap, const int *nbytes, const long *off, int ns,
                               unsigned char *dst, hipStream_t st);
static inline int dev_rex(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const float *hbs,
                          const double *hmax, const double *zbd, int s0, int ns, unsigned char *out, long cap,
                          int *nbytes, unsigned *scr, hipStream_t st)
{
   return sl_dev_rex_f32(B, R, hbs, hmax, zbd, s0, ns, out, cap, nbytes, scr, st);
}
static inline int dev_rex(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const double *hbs,
                          const double *hmax, const double *zbd, int s0, int ns, unsigned char *out, long cap,
                          int *nbytes, unsigned *scr, hipStream_t st)
{
   return sl_dev_rex_f64(B, R, hbs, hmax, zbd, s0, ns, out, cap, nbytes, scr, st);
}

extern "C" int sl_dev_sched_f32(const double *prog, const long *off, int nstorm, float *params, int maxsteps,
                                int ismpt2g, int incsm, int filt_always, hipStream_t st);
extern "C" int sl_dev_sched_f64(const double *prog, const long *off, int nstorm, double *params, int maxsteps,
                                int ismpt2g, int incsm, int filt_always, hipStream_t st);
static inline int dev_sched(const double *prog, const long *off, int nstorm, float *params, int maxsteps,
                            int ismpt2g, int incsm, int filt_always, hipStream_t st)
{
   return sl_dev_sched_f32(prog, off, nstorm, params, maxsteps, ismpt2g, incsm, filt_always, st);
}
static inline int dev_sched(const double *prog, const long *off, int nstorm, double *params, int maxsteps,
                            int ismpt2g, int incsm, int filt_always, hipStream_t st)
{
   return sl_dev_sched_f64(prog, off, nstorm, params, maxsteps, ismpt2g, incsm, filt_always, st);
}

extern "C" int sl_dev_intlht_f32(const double *prog, const long *off, const double *bm, const unsigned char *ocn,
                                 int imxb, int jmxb, int imxb1, int jmxb1, int pitch, const SlRunDevT<float> *R,
                                 hipStream_t st);
extern "C" int sl_dev_intlht_f64(const double *prog, const long *off, const double *bm, const unsigned char *ocn,
                                 int imxb, int jmxb, int imxb1, int jmxb1, int pitch, const SlRunDevT<double> *R,
                                 hipStream_t st);
static inline int dev_intlht(const double *prog, const long *off, const double *bm, const unsigned char *ocn,
                             int imxb, int jmxb, int imxb1, int jmxb1, int pitch, const SlRunDevT<float> *R,
                             hipStream_t st)
{
   return sl_dev_intlht_f32(prog, off, bm, ocn, imxb, jmxb, imxb1, jmxb1, pitch, R, st);
}
static inline int dev_intlht(const double *prog, const long *off, const double *bm, const unsigned char *ocn,
                             int imxb, int jmxb, int imxb1, int jmxb1, int pitch, const SlRunDevT<double> *R,
                             hipStream_t st)
{
   return sl_dev_intlht_f64(prog, off, bm, ocn, imxb, jmxb, imxb1, jmxb1, pitch, R, st);
}
extern "C" int sl_dev_tide_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const SlTideDev *T, float *hb,
                               int n, int first, hipStream_t st);
extern "C" int sl_dev_tide_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const SlTideDev *T,
                               double *hb, int n, int first, hipStream_t st);
static inline int dev_tide(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, const SlTideDev *T, float *hb, int n,
                           int first, hipStream_t st)
{
   return sl_dev_tide_f32(B, R, T, hb, n, first, st);
}
static inline int dev_tide(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, const SlTideDev *T, double *hb,
                           int n, int first, hipStream_t st)
{
   return sl_dev_tide_f64(B, R, T, hb, n, first, st);
}

extern "C" int sl_dev_slot_init_f32(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, int slot, int tstart,
                                    int nsteps, hipStream_t st);
extern "C" int sl_dev_slot_init_f64(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, int slot, int tstart,
                                    int nsteps, hipStream_t st);
static inline int dev_slot_init(const SlBasinDevT<float> *B, const SlRunDevT<float> *R, int slot, int tstart,
                                int nsteps, hipStream_t st)
{
   return sl_dev_slot_init_f32(B, R, slot, tstart, nsteps, st);
}
static inline int dev_slot_init(const SlBasinDevT<double> *B, const SlRunDevT<double> *R, int slot, int tstart,
                                int nsteps, hipStream_t st)
{
   return sl_dev_slot_init_f64(B, R, slot, tstart, nsteps, st);
}

// a track given in memory, as sl_trk_parse leaves a file's
extern "C" int sl_track_from_mem(const slg_track *m, sl_trk *t, char e[256]);
// storm schedule of one run (host, double): rows kept when keep_rows; the
// per-step values in the batch's precision when hostp, else only the storm
// program the device generates them from.  trk: a file path, or with track
// set, the track itself (slg_batch_open_tracks).
int sl_make_schedule(const slg_basin *B, const char *trk, storm_sched &S, bool keep_rows, int prec, char *err,
                     int errlen, bool hostp = true, const slg_track *track = nullptr);
// the basin's static fields on the device in precision prec (uploads once)
int sl_basin_ensure(slg_basin *B, int prec, char *err, int errlen);
static inline const std::vector<float> &sched_params(const storm_sched &S, float) { return S.p32; }
static inline const std::vector<double> &sched_params(const storm_sched &S, double) { return S.p64; }
static inline const std::vector<float> &sched_tables(const storm_sched &S, float) { return S.t32; }
static inline const std::vector<double> &sched_tables(const storm_sched &S, double) { return S.t64; }

#endif
