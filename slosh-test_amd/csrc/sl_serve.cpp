// sl_serve.cpp -- serving storm runs on one GPU with continuous batching.
//
// Each registered basin gets a pool: device memory for a fixed number of run
// slots (fields, schedule, radial-table arena), sized from the free HBM at
// registration, and its own HIP stream.  Runs are submitted as they arrive;
// their storm schedule (the host-side TTIMER/STMVAL/FORCE1 sequence) and
// initial water are prepared on worker threads, then they wait in a FIFO and
// are admitted into a free slot of their basin's pool while the other slots
// keep stepping.  The kernels read each slot's own step counter (runs in one
// pool are at different steps) and skip empty or finished slots.  A run's
// fields depend only on its own track and basin, so its result is the same
// alone or in any mix.
//
// The host never waits for a whole stream.  Everything the device needs from
// the host goes through a pinned staging ring: an admission copies the run's
// initial water, schedule and tables into the ring, queues the copies and a
// slot-initialisation kernel on the pool's stream, and records an event that
// releases the ring space once the copies are done.  A run whose last step
// has been queued is retired in the same way: its envelope (EnvSave
// quantisation) is computed on the device and copied into the ring behind an
// event; its slot and tables are free at once (anything later on the stream
// runs after the envelope kernel), and the run is reported done when a later
// pump finds the event complete.  Each pool keeps at most two pumps of steps
// queued, which bounds how far the host runs ahead of the device.
//
// Launch shape: each pool tunes its tile on its first launch rounds, timing
// the 64 x 8, 32 x 16 and 16 x 32 tiles with events on real steps (all three
// compute identical fields) and keeping the fastest per slot-step.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include "sl_int.h"

namespace {

double now_s()
{
   return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Pinned host memory handed out in FIFO order; a range is reused once the
// device work that reads or writes it is complete (its event) and the host is
// done with it (consumed).
struct Ring {
   char *base = nullptr;
   size_t cap = 0, head = 0;
   struct Chunk {
      size_t off, len;
      hipEvent_t ev;
      bool consumed;
   };
   std::deque<Chunk> busy;
   std::vector<hipEvent_t> spare;

   hipEvent_t event()
   {
      if (!spare.empty()) {
         hipEvent_t e = spare.back();
         spare.pop_back();
         return e;
      }
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
      return e;
   }
   void reclaim()
   {
      while (!busy.empty() && busy.front().consumed && hipEventQuery(busy.front().ev) == hipSuccess) {
         spare.push_back(busy.front().ev);
         busy.pop_front();
      }
   }
   // a range of len bytes (multiple of 256), or -1 when the ring is too full
   long take(size_t len)
   {
      reclaim();
      if (busy.empty()) head = 0;
      const size_t tail = busy.empty() ? 0 : busy.front().off;
      if (busy.empty() || head > tail) {
         if (cap - head >= len) {
            head += len;
            return (long) (head - len);
         }
         if (!busy.empty() && len < tail) {   // wrap to the start
            head = len;
            return 0;
         }
         return -1;
      }
      if (tail - head > len) {   // wrapped: strictly below the oldest range
         head += len;
         return (long) (head - len);
      }
      return -1;
   }
};

struct Pool {
   slg_basin *basin = nullptr;
   int slots = 0, maxsteps = 0, tsets = 0;
   void *dmem = nullptr;
   size_t bytes = 0;
   SlRunDevT<float> d32;
   SlRunDevT<double> d64;
   hipStream_t st = nullptr;
   std::vector<int> run;     // run id in each slot, -1 free
   std::vector<int> left;    // steps of the slot's run not yet queued
   std::vector<std::pair<int, int>> free_ext;   // free table-set extents (start, count)
   short *denv = nullptr;    // device envelope of one slot (reused in stream order)
   size_t cells = 0;         // computational cells (IMXB1*JMXB1)
   // launch rounds in flight (back-pressure) and the tile tuner
   hipEvent_t ev_round[2] = {};
   int inflight = 0;
   int tile_rows = 32;
   int trial = 0;            // tiles tried so far (3: tuned)
   hipEvent_t ev_t0[3] = {}, ev_t1[3] = {};
   double trial_units[3] = {};
   bool trial_pending[3] = {};
};

struct Run {
   int basin = 0, slot = -1;
   std::atomic<int> state{-1};   // -1 being prepared, 0 queued, 1 running, 2 done, 3 failed
   // copied from the pool at submission: worker threads never touch the pools
   slg_basin *bas = nullptr;
   int maxsteps = 0, tsets = 0;
   std::string trk, err;
   storm_sched sched;
   std::vector<double> hb0;
   std::vector<short> env;
   int tstart = -1, ntab = 0;
   long ring_off = -1;          // the envelope's range in the ring while it is being copied back
   double t_submit = 0, t_admit = 0, t_done = 0;
};

int take_tables(Pool &P, int n)
{
   for (auto &e : P.free_ext)
      if (e.second >= n) {
         const int s = e.first;
         e.first += n;
         e.second -= n;
         return s;
      }
   return -1;
}

void give_tables(Pool &P, int start, int n)
{
   P.free_ext.push_back({start, n});
   // merge adjacent extents
   std::sort(P.free_ext.begin(), P.free_ext.end());
   std::vector<std::pair<int, int>> m;
   for (auto &e : P.free_ext) {
      if (e.second == 0) continue;
      if (!m.empty() && m.back().first + m.back().second == e.first) m.back().second += e.second;
      else m.push_back(e);
   }
   P.free_ext.swap(m);
}

size_t up256(size_t n)
{
   return (n + 255) & ~(size_t) 255;
}

}   // namespace

struct slg_server {
   int prec = 32;
   double frac = 0.5;
   std::vector<std::unique_ptr<Pool>> pools;
   std::vector<std::unique_ptr<Run>> runs;
   std::deque<int> queue;      // submitted, not yet admitted (submission order)
   int order = SLG_ORDER_FIFO; // admission order of the queued runs
   std::vector<int> finishing; // retired runs whose envelope is still being copied back
   Ring ring;
   long long cell_steps = 0;
   // schedule preparation workers
   std::mutex mu;
   std::condition_variable cv;
   std::deque<Run *> todo;
   bool stop = false;
   std::vector<std::thread> workers;
};

static void prepare(slg_server *S, Run *r)
{
   char e[512];
   e[0] = 0;
   int rc = sl_make_schedule(r->bas, r->trk.c_str(), r->sched, false, S->prec, e, sizeof e);
   if (!rc && (r->sched.nsteps > r->maxsteps || (int) r->sched.tab_step.size() > r->tsets)) {
      snprintf(e, sizeof e, "%s", "run longer (or with more table sets) than the pool's slots hold");
      rc = SLG_ELIMIT;
   }
   if (rc) {
      r->err = e;
      r->state = 3;
      return;
   }
   const size_t npt = (size_t) r->bas->h.imxb * r->bas->h.jmxb;
   r->hb0.resize(npt);
   sl_initial_water(&r->bas->h, &r->sched.s0, r->hb0.data());
   r->ntab = (int) r->sched.tab_step.size();
   r->state = 0;
}

static void worker(slg_server *S)
{
   for (;;) {
      Run *r;
      {
         std::unique_lock<std::mutex> lk(S->mu);
         S->cv.wait(lk, [S] { return S->stop || !S->todo.empty(); });
         if (S->stop && S->todo.empty()) return;
         r = S->todo.front();
         S->todo.pop_front();
      }
      prepare(S, r);
   }
}

// basins with 1-D flow sides or channels: per-slot side state and a third HB
static bool pool_sides(const slg_basin *B)
{
   return B->h.nsqrw > 0 || (B->h.nsqrwc > 0 && B->h.nof1d != '+') || B->h.npss > 0;
}

static size_t slot_bytes(const slg_basin *B, int maxsteps, int tsets_per_slot, int prec)
{
   const size_t rs = prec == 64 ? 8 : 4, np = (size_t) B->pitch * B->h.jmxb;
   const size_t nl = (size_t) std::max(B->h.nsqrwc, 1);
   // fields (a third HB with sides), masks, the HSUB top row, the event list,
   // counters, schedule and tables, FLWSQR + F1DACT, the side squares' HSUB
   // (at most two per side), the channel flows
   return np * rs * (pool_sides(B) ? 6 : 5) + np * 3 + (size_t) B->pitch * rs + SL_EVCAP * 4 + 64 +
          (size_t) maxsteps * SL_NPAR * rs + (size_t) tsets_per_slot * SL_NTABV * 8 * rs + nl * (rs + 1) +
          2 * nl * rs + 5 * rs + 4096;
}

template <class R>
static SlRunDevT<R> &pool_run(Pool &P)
{
   return *(SlRunDevT<R> *) (sizeof(R) == 8 ? (void *) &P.d64 : (void *) &P.d32);
}

template <class R>
static int pool_alloc(Pool &P, char *err, int errlen)
{
   const slg_basin *B = P.basin;
   const int n = P.slots;
   const size_t np = (size_t) B->pitch * B->h.jmxb, ns = np * n, rs = sizeof(R);
   size_t off[32], tot = 0;
   int k = 0;
   auto add = [&](size_t bytes) {
      off[k++] = tot;
      tot += up256(bytes);
   };
   add(ns * rs); add(ns * rs); add(ns * rs); add(ns * rs); add(ns * rs);   // hb0 hb1 ub vb hmx
   add((size_t) n * B->pitch * rs);                                       // hsubtop
   add(ns); add(ns); add(ns);                                             // pflag hclass uz
   add((size_t) n * SL_EVCAP * 4);                                        // evkey
   add((size_t) n * 4);                                                   // evcnt
   add((size_t) n * 8 * 4);                                               // stat
   add((size_t) n * P.maxsteps * SL_NPAR * rs);                           // params
   add((size_t) P.tsets * SL_NTABV * 8 * rs);                             // table arena
   add((size_t) n * 4);                                                   // tbase
   add((size_t) n * 4);                                                   // nsteps
   add((size_t) n * 4);                                                   // cur
   add((size_t) B->h.imxb * B->h.jmxb * 2);                               // one envelope
   const BasinImg<R> &I = img(const_cast<slg_basin *>(B), R());
   const size_t nl = (size_t) std::max(B->h.nsqrwc, 1), nq = (size_t) std::max(I.nsq, 1);
   add((size_t) n * nl * rs);                                             // flwsqr
   add((size_t) n * nl);                                                  // f1dact
   add((size_t) n * nq * rs);                                             // hsq
   add((size_t) n * 5 * rs);                                              // chq
   const bool sides = pool_sides(B);
   add(sides ? ns * rs : 0);                                              // hb2
   char *d = nullptr;
   HIPOK(hipMalloc((void **) &d, tot));
   HIPOK(hipMemset(d, 0, tot));   // nsteps = 0: every slot empty
   P.dmem = d;
   P.bytes = tot;
   SlRunDevT<R> &D = pool_run<R>(P);
   memset(&D, 0, sizeof D);
   D.nstorm = n;
   D.fstride = np;
   D.hb0 = (R *) (d + off[0]);
   D.hb1 = (R *) (d + off[1]);
   D.ub = (R *) (d + off[2]);
   D.vb = (R *) (d + off[3]);
   D.hmx = (R *) (d + off[4]);
   D.hsubtop = (R *) (d + off[5]);
   D.pflag = (uint8_t *) (d + off[6]);
   D.hclass = (uint8_t *) (d + off[7]);
   D.uz = (uint8_t *) (d + off[8]);
   D.evkey = (uint32_t *) (d + off[9]);
   D.evcnt = (int *) (d + off[10]);
   D.stat = (int *) (d + off[11]);
   D.params = (const R *) (d + off[12]);
   D.maxsteps = P.maxsteps;
   D.tables = (const R *) (d + off[13]);
   D.tbase = (const int *) (d + off[14]);
   D.nsteps = (const int *) (d + off[15]);
   D.cur = (int *) (d + off[16]);
   D.evcap = SL_EVCAP;
   D.vbudget = 0;
   D.sgroup = SL_SGROUP;
   P.denv = (short *) (d + off[17]);
   D.flwsqr = (R *) (d + off[18]);
   D.f1dact = (uint8_t *) (d + off[19]);
   D.hsq = (R *) (d + off[20]);
   D.chq = (R *) (d + off[21]);
   D.hb2 = sides ? (R *) (d + off[22]) : nullptr;
   HIPOK(hipStreamCreate(&P.st));
   for (auto &e : P.ev_round) HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
   for (int c = 0; c < 3; c++) {
      HIPOK(hipEventCreate(&P.ev_t0[c]));
      HIPOK(hipEventCreate(&P.ev_t1[c]));
   }
   return 0;
}

extern "C" int slg_server_open(int precision, double hbm_fraction, slg_server **out, char *err, int errlen)
{
   *out = nullptr;
   if ((precision != 32 && precision != 64) || !(hbm_fraction > 0 && hbm_fraction <= 1)) {
      seterr(err, errlen, "%s", "precision must be 32 or 64, hbm_fraction in (0, 1]");
      return SLG_EINPUT;
   }
   // pinned staging: admissions (initial water, schedule, tables) and
   // envelopes on their way back
   const size_t ring_bytes = (size_t) 512 << 20;
   char *base = nullptr;
   HIPOK(hipHostMalloc((void **) &base, ring_bytes, hipHostMallocDefault));
   slg_server *S = new slg_server();
   S->prec = precision;
   S->frac = hbm_fraction;
   S->ring.base = base;
   S->ring.cap = ring_bytes;
   const int nw = std::max(1, std::min<int>(8, (int) std::thread::hardware_concurrency()));
   for (int k = 0; k < nw; k++) S->workers.emplace_back(worker, S);
   *out = S;
   return 0;
}

// Register a basin: a pool of up to max_slots run slots, each able to hold a
// run of max_steps steps, with an arena of tables_per_slot radial-table sets
// per slot on average.  The slot count is cut to what the free HBM allows
// (hipMemGetInfo, times the server's fraction, less what other pools hold).
extern "C" int slg_server_add_basin(slg_server *S, const char *dta, const char *bnt_line, int max_slots,
                                    int max_steps, int tables_per_slot, int *basin_id, char *err, int errlen)
{
   if (max_slots < 1 || max_slots > SLG_MAX_STORMS || max_steps < 1 || tables_per_slot < 1) {
      seterr(err, errlen, "%s", "max_slots must be 1 .. SLG_MAX_STORMS; max_steps and tables_per_slot >= 1");
      return SLG_EINPUT;
   }
   slg_basin *B = nullptr;
   int rc = slg_basin_open(dta, bnt_line, &B, err, errlen);
   if (rc) return rc;
   rc = sl_basin_ensure(B, S->prec, err, errlen);
   if (rc) {
      slg_basin_close(B);
      return rc;
   }
   size_t fre = 0, total = 0;
   HIPOK(hipMemGetInfo(&fre, &total));
   const size_t per = slot_bytes(B, max_steps, tables_per_slot, S->prec);
   const double budget = (double) fre * S->frac;
   int slots = (int) std::min<double>(max_slots, budget / (double) per);
   if (slots < 1) {
      slg_basin_close(B);
      seterr(err, errlen, "%s", "not enough free HBM for one run slot");
      return SLG_ELIMIT;
   }
   std::unique_ptr<Pool> P(new Pool());
   P->basin = B;
   P->slots = slots;
   P->maxsteps = max_steps;
   P->tsets = slots * tables_per_slot;
   P->cells = (size_t) B->h.imxb1 * B->h.jmxb1;
   P->run.assign(slots, -1);
   P->left.assign(slots, 0);
   P->free_ext.push_back({0, P->tsets});
   rc = S->prec == 64 ? pool_alloc<double>(*P, err, errlen) : pool_alloc<float>(*P, err, errlen);
   if (rc) {
      slg_basin_close(B);
      return rc;
   }
   S->pools.push_back(std::move(P));
   *basin_id = (int) S->pools.size() - 1;
   return 0;
}

extern "C" int slg_server_slots(const slg_server *S, int basin_id)
{
   return (basin_id >= 0 && basin_id < (int) S->pools.size()) ? S->pools[basin_id]->slots : -1;
}

// Submit a run (trk file) on a registered basin; its schedule is prepared on a
// worker thread and the run waits for a slot.
extern "C" int slg_server_submit(slg_server *S, int basin_id, const char *trk, int *run_id, char *err,
                                 int errlen)
{
   if (basin_id < 0 || basin_id >= (int) S->pools.size()) {
      seterr(err, errlen, "%s", "unknown basin");
      return SLG_EINPUT;
   }
   const Pool &P = *S->pools[basin_id];
   std::unique_ptr<Run> r(new Run());
   r->basin = basin_id;
   r->bas = P.basin;
   r->maxsteps = P.maxsteps;
   r->tsets = P.tsets;
   r->trk = trk;
   r->t_submit = now_s();
   Run *rp = r.get();
   S->runs.push_back(std::move(r));
   *run_id = (int) S->runs.size() - 1;
   S->queue.push_back(*run_id);
   {
      std::lock_guard<std::mutex> lk(S->mu);
      S->todo.push_back(rp);
   }
   S->cv.notify_one();
   return 0;
}

// Queue a run's admission into a slot: its initial water, schedule and tables
// through the ring, then the slot-initialisation kernel.  Returns 1 when the
// ring has no room (the run waits for a later pump).
template <class R>
static int admit(slg_server *S, Pool &P, int slot, Run &r, char *err, int errlen)
{
   const slg_basin *B = P.basin;
   const sl_basin *b = &B->h;
   const int PI = B->pitch, NJ = b->jmxb;
   const size_t np = (size_t) PI * NJ;
   SlRunDevT<R> &D = pool_run<R>(P);
   const std::vector<R> &sp = sched_params(r.sched, R()), &stb = sched_tables(r.sched, R());
   const size_t lf = up256(np * sizeof(R)), lp = up256(sp.size() * sizeof(R)), lt = up256(stb.size() * sizeof(R));
   const long off = S->ring.take(lf + lp + lt);
   if (off < 0) return 1;
   char *h = S->ring.base + off;
   R *f = (R *) h;
   std::fill(f, f + np, R(0));
   for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= b->imxb; i++) f[(size_t) (j - 1) * PI + (i - 1)] = (R) r.hb0[PIX(b, i, j)];
   memcpy(h + lf, sp.data(), sp.size() * sizeof(R));
   memcpy(h + lf + lp, stb.data(), stb.size() * sizeof(R));
   const size_t so = (size_t) slot * np;
   HIPOK(hipMemcpyAsync(D.hb0 + so, h, np * sizeof(R), hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync((R *) D.params + (size_t) slot * P.maxsteps * SL_NPAR, h + lf, sp.size() * sizeof(R),
                        hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync((R *) D.tables + (size_t) r.tstart * SL_NTABV * 8, h + lf + lp, stb.size() * sizeof(R),
                        hipMemcpyHostToDevice, P.st));
   const int rc = dev_slot_init(&img(P.basin, R()).dev, &D, slot, r.tstart, r.sched.nsteps, P.st);
   if (rc) {
      seterr(err, errlen, "slot initialisation failed (%s%d)", "", -rc);
      return SLG_EHIP;
   }
   hipEvent_t ev = S->ring.event();
   if (!ev) {
      seterr(err, errlen, "%s", "event creation failed");
      return SLG_EHIP;
   }
   HIPOK(hipEventRecord(ev, P.st));
   S->ring.busy.push_back({(size_t) off, lf + lp + lt, ev, true});
   return 0;
}

// Queue a finished run's envelope and error flag into the ring; the slot and
// its tables are free from here on.  Returns 1 when the ring has no room.
template <class R>
static int retire(slg_server *S, Pool &P, int slot, char *err, int errlen)
{
   const int id = P.run[slot];
   Run &r = *S->runs[id];
   BasinImg<R> &I = img(P.basin, R());
   SlRunDevT<R> &D = pool_run<R>(P);
   const size_t n = (size_t) P.basin->h.imxb * P.basin->h.jmxb, le = up256(n * 2);
   const long off = S->ring.take(le + 256);
   if (off < 0) return 1;
   int rc = dev_envelopes(&I.dev, &D, I.zbd, slot, 1, P.denv, nullptr, P.st);
   if (rc) {
      seterr(err, errlen, "envelope kernel failed (%s%d)", "", -rc);
      return SLG_EHIP;
   }
   char *h = S->ring.base + off;
   HIPOK(hipMemcpyAsync(h, P.denv, n * 2, hipMemcpyDeviceToHost, P.st));
   HIPOK(hipMemcpyAsync(h + le, D.stat + (size_t) slot * 8, 32, hipMemcpyDeviceToHost, P.st));
   hipEvent_t ev = S->ring.event();
   if (!ev) {
      seterr(err, errlen, "%s", "event creation failed");
      return SLG_EHIP;
   }
   HIPOK(hipEventRecord(ev, P.st));
   S->ring.busy.push_back({(size_t) off, le + 256, ev, false});
   r.ring_off = off;
   give_tables(P, r.tstart, r.ntab);
   P.run[slot] = -1;
   S->finishing.push_back(id);
   return 0;
}

// Runs whose envelope has arrived in the ring are done (or failed, when the
// cascade replay flagged them).
static void finish_runs(slg_server *S)
{
   std::vector<int> still;
   for (int id : S->finishing) {
      Run &r = *S->runs[id];
      Ring::Chunk *c = nullptr;
      for (auto &k : S->ring.busy)
         if (!k.consumed && (long) k.off == r.ring_off) c = &k;
      if (!c || hipEventQuery(c->ev) != hipSuccess) {
         still.push_back(id);
         continue;
      }
      const slg_basin *B = r.bas;
      const size_t n = (size_t) B->h.imxb * B->h.jmxb;
      const char *h = S->ring.base + c->off;
      const int *st = (const int *) (h + up256(n * 2));
      r.t_done = now_s();
      if (st[7]) {
         r.err = "the drying-cascade replay ran out of visits; the run's fields are incomplete";
         r.state = 3;
      } else {
         r.env.assign((const short *) h, (const short *) h + n);
         r.state = 2;
      }
      c->consumed = true;
      r.ring_off = -1;
      // the schedule is no longer needed
      r.sched.p32 = std::vector<float>();
      r.sched.p64 = std::vector<double>();
      r.sched.t32 = std::vector<float>();
      r.sched.t64 = std::vector<double>();
      r.hb0 = std::vector<double>();
   }
   S->finishing.swap(still);
   S->ring.reclaim();
}

// The pool's next launch round: k steps over slots [0, hi).  The first three
// rounds time the three tiles; from then on the fastest per slot-step runs.
template <class R>
static int launch_round(Pool &P, int n, int hi, char *err, int errlen)
{
   SlRunDevT<R> d = pool_run<R>(P);
   d.nstorm = hi;
   const int tiles[3] = {32, 16, 8};
   int c = -1;
   if (P.trial < 3) {
      c = P.trial++;
      P.tile_rows = tiles[c];
      P.trial_units[c] = (double) n * hi;
      P.trial_pending[c] = true;
      HIPOK(hipEventRecord(P.ev_t0[c], P.st));
   }
   for (int t = 0; t < n; t++) {
      const int rc = dev_launch(&img(P.basin, R()).dev, &d, 0, 0, 1, 0, SLG_PATH_FUSED, P.tile_rows, P.st, nullptr);
      if (rc < 0) {
         seterr(err, errlen, "kernel launch failed (%s%d)", "", -rc);
         return SLG_EHIP;
      }
   }
   if (c >= 0) HIPOK(hipEventRecord(P.ev_t1[c], P.st));
   return 0;
}

// After the trial rounds have run: keep the tile with the least time per
// slot-step.
static void settle_tile(Pool &P)
{
   if (P.trial < 3) return;
   const int tiles[3] = {32, 16, 8};
   double best = 1e300;
   int bi = -1;
   for (int c = 0; c < 3; c++) {
      if (!P.trial_pending[c]) continue;
      if (hipEventQuery(P.ev_t1[c]) != hipSuccess) return;   // not measured yet
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, P.ev_t0[c], P.ev_t1[c]) != hipSuccess) continue;
      const double per = ms / std::max(1.0, P.trial_units[c]);
      if (per < best) {
         best = per;
         bi = c;
      }
   }
   if (bi >= 0) P.tile_rows = tiles[bi];
   for (auto &f : P.trial_pending) f = false;
}

// One round: collect finished envelopes, retire runs whose last step is
// queued, admit queued runs into free slots, and queue up to k further steps
// on every pool that has runs.  Returns the number of runs retired, or an
// error code.
extern "C" int slg_server_set_order(slg_server *S, int order)
{
   if (order != SLG_ORDER_FIFO && order != SLG_ORDER_HARDEST) return SLG_EINPUT;
   S->order = order;
   return 0;
}

extern "C" int slg_server_pump(slg_server *S, int k, char *err, int errlen)
{
   if (k < 0) {
      seterr(err, errlen, "%s", "k must be >= 0");
      return SLG_EINPUT;
   }
   int retired = 0;
   finish_runs(S);
   for (auto &PP : S->pools) {
      Pool &P = *PP;
      settle_tile(P);
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0 && P.left[slot] == 0) {
            const int rc = S->prec == 64 ? retire<double>(S, P, slot, err, errlen)
                                         : retire<float>(S, P, slot, err, errlen);
            if (rc < 0) return rc;
            if (rc == 0) retired++;
         }
   }
   // admission: in submission order, or hardest first: the prepared runs by
   // their work (steps x cells, the measure of the reference's job sort,
   // dev/b1.sortJobList.sh:20-73), largest first, ties in submission order.
   // A run that does not fit waits; later runs may go ahead of it.
   if (S->order == SLG_ORDER_HARDEST) {
      auto work = [&](int id) -> double {
         const Run &r = *S->runs[id];
         if (r.state != 0) return -1.;   // not prepared yet: after the prepared ones
         return (double) r.sched.nsteps * (double) (r.bas->h.imxb - 1) * (double) (r.bas->h.jmxb - 1);
      };
      std::vector<std::pair<double, int>> key;
      key.reserve(S->queue.size());
      for (int id : S->queue) key.push_back({work(id), id});
      std::stable_sort(key.begin(), key.end(), [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
         return a.first > b.first;
      });
      S->queue.clear();
      for (auto &k2 : key) S->queue.push_back(k2.second);
   }
   std::deque<int> wait;
   while (!S->queue.empty()) {
      const int id = S->queue.front();
      S->queue.pop_front();
      Run &r = *S->runs[id];
      if (r.state == 3) continue;   // failed in preparation
      if (r.state != 0) {           // still being prepared
         wait.push_back(id);
         continue;
      }
      Pool &P = *S->pools[r.basin];
      int slot = -1;
      for (int q = 0; q < P.slots && slot < 0; q++)
         if (P.run[q] < 0) slot = q;
      const int ts = slot >= 0 ? take_tables(P, r.ntab) : -1;
      if (slot < 0 || ts < 0) {
         wait.push_back(id);
         continue;
      }
      r.tstart = ts;
      const int rc = S->prec == 64 ? admit<double>(S, P, slot, r, err, errlen) : admit<float>(S, P, slot, r, err, errlen);
      if (rc < 0) return rc;
      if (rc == 1) {   // no staging room: give the tables back and wait
         give_tables(P, ts, r.ntab);
         wait.push_back(id);
         continue;
      }
      P.run[slot] = id;
      P.left[slot] = r.sched.nsteps;
      r.slot = slot;
      r.state = 1;
      r.t_admit = now_s();
   }
   S->queue.swap(wait);
   for (auto &PP : S->pools) {
      Pool &P = *PP;
      // launch over slots [0, hi): admission fills the lowest free slot, so the
      // grid shrinks with the load and empty slots beyond hi cost nothing
      int most = 0, hi = 0;
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0) {
            most = std::max(most, P.left[slot]);
            hi = slot + 1;
         }
      const int n = std::min(k, most);
      if (n == 0) continue;
      // at most two rounds queued per pool
      if (P.inflight == 2) {
         HIPOK(hipEventSynchronize(P.ev_round[0]));
         std::swap(P.ev_round[0], P.ev_round[1]);
         P.inflight = 1;
      }
      const int rc = S->prec == 64 ? launch_round<double>(P, n, hi, err, errlen)
                                   : launch_round<float>(P, n, hi, err, errlen);
      if (rc) return rc;
      HIPOK(hipEventRecord(P.ev_round[P.inflight], P.st));
      P.inflight++;
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0) {
            const int d = std::min(n, P.left[slot]);
            P.left[slot] -= d;
            S->cell_steps += (long long) d * (long long) P.cells;
         }
   }
   return retired;
}

// Run state (0 queued, 1 running, 2 done, 3 failed) and, when done, its
// envelope [jmxb][imxb] (EnvSave quantisation) and latency from submission (s).
extern "C" int slg_server_result(const slg_server *S, int run_id, short *env, double *latency_s)
{
   if (run_id < 0 || run_id >= (int) S->runs.size()) return SLG_EINPUT;
   const Run &r = *S->runs[run_id];
   const int st = r.state;
   if (st == 2) {
      if (env) memcpy(env, r.env.data(), r.env.size() * 2);
      if (latency_s) *latency_s = r.t_done - r.t_submit;
   }
   return st < 0 ? 0 : st;
}

// The message of a failed run (state 3).
extern "C" int slg_server_run_error(const slg_server *S, int run_id, char *err, int errlen)
{
   if (run_id < 0 || run_id >= (int) S->runs.size()) return SLG_EINPUT;
   const Run &r = *S->runs[run_id];
   seterr(err, errlen, "%s", r.state == 3 ? r.err.c_str() : "");
   return 0;
}

// Counters: [0] queued, [1] running, [2] done, [3] slots, [4] device bytes of
// the pools, [5] storm.cell.steps queued.
extern "C" int slg_server_stats(const slg_server *S, long long *out6)
{
   long long q = (long long) S->queue.size(), run = 0, done = 0, slots = 0, bytes = 0;
   for (auto &r : S->runs) {
      run += r->state == 1;
      done += r->state == 2;
   }
   for (auto &P : S->pools) {
      slots += P->slots;
      bytes += (long long) P->bytes;
   }
   out6[0] = q;
   out6[1] = run;
   out6[2] = done;
   out6[3] = slots;
   out6[4] = bytes;
   out6[5] = S->cell_steps;
   return 0;
}

// Tile rows each pool runs (8, 16 or 32; the tuner's choice once it has
// measured), out[npools]; returns the number of pools.
extern "C" int slg_server_tiles(const slg_server *S, int *out, int n)
{
   const int np = (int) S->pools.size();
   for (int p = 0; p < np && p < n; p++) out[p] = S->pools[p]->tile_rows;
   return np;
}

extern "C" void slg_server_close(slg_server *S)
{
   if (!S) return;
   {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
   }
   S->cv.notify_all();
   for (auto &w : S->workers) w.join();
   for (auto &PP : S->pools) {
      Pool &P = *PP;
      if (P.st) {
         (void) hipStreamSynchronize(P.st);
         (void) hipStreamDestroy(P.st);
      }
      for (auto &e : P.ev_round)
         if (e) (void) hipEventDestroy(e);
      for (int c = 0; c < 3; c++) {
         if (P.ev_t0[c]) (void) hipEventDestroy(P.ev_t0[c]);
         if (P.ev_t1[c]) (void) hipEventDestroy(P.ev_t1[c]);
      }
      if (P.dmem) (void) hipFree(P.dmem);
      slg_basin_close(P.basin);
   }
   for (auto &c : S->ring.busy) (void) hipEventDestroy(c.ev);
   for (auto &e : S->ring.spare) (void) hipEventDestroy(e);
   if (S->ring.base) (void) hipHostFree(S->ring.base);
   delete S;
}
