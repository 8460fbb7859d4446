// sl_serve.cpp -- serving storm runs on one GPU with continuous batching.
//
// Each registered basin gets a pool: device memory for a fixed number of run
// slots (fields, schedule, radial-table arena), sized from the free HBM at
// registration.  Runs are submitted as they arrive (their storm schedule and
// initial water are computed on the host at submission), wait in a FIFO, and
// are admitted into a free slot of their basin's pool while the other slots
// keep stepping.  Every pump advances all pools by up to k steps, each pool on
// its own HIP stream so the pools' kernels overlap; the kernels read each
// slot's own step counter (runs in one pool are at different steps) and skip
// empty or finished slots.  A finished run's envelope (EnvSave quantisation)
// is computed on the device and copied out; its slot and tables are freed.
// A run's storm schedule (the host-side TTIMER/STMVAL/FORCE1 sequence) and
// initial water are prepared on worker threads between submission and
// admission, so arrivals do not stall the stepping.
// A run's fields depend only on its own track and basin, so its result is the
// same alone or in any mix.
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

struct Pool {
   slg_basin *basin = nullptr;
   int slots = 0, maxsteps = 0, tsets = 0;
   void *dmem = nullptr;
   size_t bytes = 0;
   SlRunDevT<float> d32;
   SlRunDevT<double> d64;
   hipStream_t st = nullptr;
   std::vector<int> run;     // run id in each slot, -1 free
   std::vector<int> left;    // steps the run in the slot still has to go
   std::vector<std::pair<int, int>> free_ext;   // free table-set extents (start, count)
   short *denv = nullptr;    // device envelope of one slot
   size_t cells = 0;         // computational cells (IMXB1*JMXB1)
};

struct Run {
   int basin = 0, slot = -1;
   std::atomic<int> state{-1};   // -1 being prepared, 0 queued, 1 running, 2 done, 3 failed
   std::string trk, err;
   storm_sched sched;
   std::vector<double> hb0;
   std::vector<short> env;
   int tstart = -1, ntab = 0;
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

}   // namespace

struct slg_server {
   int prec = 32;
   double frac = 0.5;
   int tile_rows = 16;
   std::vector<Pool> pools;
   std::vector<std::unique_ptr<Run>> runs;
   std::deque<int> queue;   // submitted, not yet admitted (submission order)
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
   Pool &P = S->pools[r->basin];
   char e[512];
   e[0] = 0;
   int rc = sl_make_schedule(P.basin, r->trk.c_str(), r->sched, false, S->prec, e, sizeof e);
   if (!rc && (r->sched.nsteps > P.maxsteps || (int) r->sched.tab_step.size() > P.tsets)) {
      snprintf(e, sizeof e, "%s", "run longer (or with more table sets) than the pool's slots hold");
      rc = SLG_ELIMIT;
   }
   if (rc) {
      r->err = e;
      r->state = 3;
      return;
   }
   const size_t npt = (size_t) P.basin->h.imxb * P.basin->h.jmxb;
   r->hb0.resize(npt);
   sl_initial_water(&P.basin->h, &r->sched.s0, r->hb0.data());
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

static size_t slot_bytes(const slg_basin *B, int maxsteps, int tsets_per_slot, int prec)
{
   const size_t rs = prec == 64 ? 8 : 4, np = (size_t) B->pitch * B->h.jmxb;
   return np * rs * 5 + np * 3 + (size_t) B->pitch * rs + SL_EVCAP * 4 + 64 + (size_t) maxsteps * SL_NPAR * rs +
          (size_t) tsets_per_slot * SL_NTABV * 8 * rs + 4096;
}

template <class R>
static int pool_alloc(slg_server *S, Pool &P, char *err, int errlen)
{
   const slg_basin *B = P.basin;
   const int n = P.slots;
   const size_t np = (size_t) B->pitch * B->h.jmxb, ns = np * n, rs = sizeof(R);
   size_t off[20], tot = 0;
   int k = 0;
   auto add = [&](size_t bytes) {
      off[k++] = tot;
      tot += (bytes + 255) & ~(size_t) 255;
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
   char *d = nullptr;
   HIPOK(hipMalloc((void **) &d, tot));
   HIPOK(hipMemset(d, 0, tot));   // nsteps = 0: every slot empty
   P.dmem = d;
   P.bytes = tot;
   SlRunDevT<R> &D = *(SlRunDevT<R> *) (sizeof(R) == 8 ? (void *) &P.d64 : (void *) &P.d32);
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
   P.denv = (short *) (d + off[17]);
   HIPOK(hipStreamCreate(&P.st));
   (void) S;
   return 0;
}

extern "C" int slg_server_open(int precision, double hbm_fraction, slg_server **out, char *err, int errlen)
{
   *out = nullptr;
   if ((precision != 32 && precision != 64) || !(hbm_fraction > 0 && hbm_fraction <= 1)) {
      seterr(err, errlen, "%s", "precision must be 32 or 64, hbm_fraction in (0, 1]");
      return SLG_EINPUT;
   }
   slg_server *S = new slg_server();
   S->prec = precision;
   S->frac = hbm_fraction;
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
   if (B->h.nsqrw > 0 || (B->h.nsqrwc > 0 && B->h.nof1d != '+') || B->h.npss > 0) {
      slg_basin_close(B);
      seterr(err, errlen, "%s", "basins with 1-D flow sides or channels are stepped through the batch API, not served");
      return SLG_EUNSUP;
   }
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
   Pool P;
   P.basin = B;
   P.slots = slots;
   P.maxsteps = max_steps;
   P.tsets = slots * tables_per_slot;
   P.cells = (size_t) B->h.imxb1 * B->h.jmxb1;
   P.run.assign(slots, -1);
   P.left.assign(slots, 0);
   P.free_ext.push_back({0, P.tsets});
   rc = S->prec == 64 ? pool_alloc<double>(S, P, err, errlen) : pool_alloc<float>(S, P, err, errlen);
   if (rc) {
      slg_basin_close(B);
      return rc;
   }
   S->pools.push_back(P);
   *basin_id = (int) S->pools.size() - 1;
   return 0;
}

extern "C" int slg_server_slots(const slg_server *S, int basin_id)
{
   return (basin_id >= 0 && basin_id < (int) S->pools.size()) ? S->pools[basin_id].slots : -1;
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
   std::unique_ptr<Run> r(new Run());
   r->basin = basin_id;
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

template <class R>
static int admit(slg_server *S, Pool &P, int slot, int id, char *err, int errlen)
{
   Run &r = *S->runs[id];
   const slg_basin *B = P.basin;
   const sl_basin *b = &B->h;
   const int PI = B->pitch, NJ = b->jmxb;
   const size_t np = (size_t) PI * NJ;
   SlRunDevT<R> &D = *(SlRunDevT<R> *) (sizeof(R) == 8 ? (void *) &P.d64 : (void *) &P.d32);
   std::vector<R> f(np, R(0)), h(np, R(0));
   for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= b->imxb; i++) {
         const R v = (R) r.hb0[PIX(b, i, j)];
         f[(size_t) (j - 1) * PI + (i - 1)] = v;
         if (i < b->imxb && j < NJ) h[(size_t) (j - 1) * PI + (i - 1)] = v;   // HMX = HB on squares
      }
   const size_t so = (size_t) slot * np;
   HIPOK(hipMemcpyAsync(D.hb0 + so, f.data(), np * sizeof(R), hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync(D.hb1 + so, f.data(), np * sizeof(R), hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync(D.hmx + so, h.data(), np * sizeof(R), hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemsetAsync(D.ub + so, 0, np * sizeof(R), P.st));
   HIPOK(hipMemsetAsync(D.vb + so, 0, np * sizeof(R), P.st));
   HIPOK(hipMemsetAsync(D.hsubtop + (size_t) slot * PI, 0, PI * sizeof(R), P.st));
   HIPOK(hipMemsetAsync(D.evcnt + slot, 0, 4, P.st));
   HIPOK(hipMemsetAsync(D.stat + slot * 8, 0, 32, P.st));
   const std::vector<R> &sp = sched_params(r.sched, R()), &stb = sched_tables(r.sched, R());
   HIPOK(hipMemcpyAsync((R *) D.params + (size_t) slot * P.maxsteps * SL_NPAR, sp.data(), sp.size() * sizeof(R),
                        hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync((R *) D.tables + (size_t) r.tstart * SL_NTABV * 8, stb.data(), stb.size() * sizeof(R),
                        hipMemcpyHostToDevice, P.st));
   const int zero = 0;
   HIPOK(hipMemcpyAsync((int *) D.tbase + slot, &r.tstart, 4, hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync(D.cur + slot, &zero, 4, hipMemcpyHostToDevice, P.st));
   HIPOK(hipMemcpyAsync((int *) D.nsteps + slot, &r.sched.nsteps, 4, hipMemcpyHostToDevice, P.st));
   // the host buffers above are locals: finish the copies before they go
   HIPOK(hipStreamSynchronize(P.st));
   return 0;
}

template <class R>
static int retire(slg_server *S, Pool &P, int slot, char *err, int errlen)
{
   const int id = P.run[slot];
   Run &r = *S->runs[id];
   BasinImg<R> &I = img(P.basin, R());
   SlRunDevT<R> &D = *(SlRunDevT<R> *) (sizeof(R) == 8 ? (void *) &P.d64 : (void *) &P.d32);
   const size_t n = (size_t) P.basin->h.imxb * P.basin->h.jmxb;
   int rc = dev_envelopes(&I.dev, &D, I.zbd, slot, 1, P.denv, nullptr, P.st);
   if (rc) {
      seterr(err, errlen, "envelope kernel failed (%s%d)", "", -rc);
      return SLG_EHIP;
   }
   r.env.resize(n);
   HIPOK(hipMemcpyAsync(r.env.data(), P.denv, n * 2, hipMemcpyDeviceToHost, P.st));
   HIPOK(hipStreamSynchronize(P.st));
   r.state = 2;
   r.t_done = now_s();
   give_tables(P, r.tstart, r.ntab);
   P.run[slot] = -1;
   // the schedule is no longer needed
   r.sched.p32.clear();
   r.sched.p32.shrink_to_fit();
   r.sched.p64.clear();
   r.sched.p64.shrink_to_fit();
   r.sched.t32.clear();
   r.sched.t32.shrink_to_fit();
   r.sched.t64.clear();
   r.sched.t64.shrink_to_fit();
   r.hb0.clear();
   r.hb0.shrink_to_fit();
   return 0;
}

// One round: finish what the last round launched, retire finished runs, admit
// queued runs into free slots, and launch up to k steps on every pool that has
// runs (asynchronous).  Returns the number of runs retired, or an error code.
extern "C" int slg_server_pump(slg_server *S, int k, char *err, int errlen)
{
   int retired = 0;
   for (auto &P : S->pools) HIPOK(hipStreamSynchronize(P.st));
   for (auto &P : S->pools)
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0 && P.left[slot] == 0) {
            int rc = S->prec == 64 ? retire<double>(S, P, slot, err, errlen) : retire<float>(S, P, slot, err, errlen);
            if (rc) return rc;
            retired++;
         }
   // admission: FIFO per basin; a run that does not fit waits, later runs of
   // other basins may go ahead of it
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
      Pool &P = S->pools[r.basin];
      int slot = -1;
      for (int q = 0; q < P.slots && slot < 0; q++)
         if (P.run[q] < 0) slot = q;
      const int ts = slot >= 0 ? take_tables(P, r.ntab) : -1;
      if (slot < 0 || ts < 0) {
         wait.push_back(id);
         continue;
      }
      r.tstart = ts;
      int rc = S->prec == 64 ? admit<double>(S, P, slot, id, err, errlen) : admit<float>(S, P, slot, id, err, errlen);
      if (rc) return rc;
      P.run[slot] = id;
      P.left[slot] = r.sched.nsteps;
      r.slot = slot;
      r.state = 1;
      r.t_admit = now_s();
   }
   S->queue.swap(wait);
   for (auto &P : S->pools) {
      // launch over slots [0, hi): admission fills the lowest free slot, so the
      // grid shrinks with the load and empty slots beyond hi cost nothing
      int most = 0, hi = 0;
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0) {
            most = std::max(most, P.left[slot]);
            hi = slot + 1;
         }
      const int n = std::min(k, most);
      SlRunDevT<float> d32 = P.d32;
      SlRunDevT<double> d64 = P.d64;
      d32.nstorm = hi;
      d64.nstorm = hi;
      for (int t = 0; t < n; t++) {
         int rc = S->prec == 64 ? dev_launch(&P.basin->i64.dev, &d64, 0, 0, 1, 0, SLG_PATH_FUSED, S->tile_rows, P.st, nullptr)
                                : dev_launch(&P.basin->i32.dev, &d32, 0, 0, 1, 0, SLG_PATH_FUSED, S->tile_rows, P.st, nullptr);
         if (rc < 0) {
            seterr(err, errlen, "kernel launch failed (%s%d)", "", -rc);
            return SLG_EHIP;
         }
      }
      for (int slot = 0; slot < P.slots; slot++)
         if (P.run[slot] >= 0) {
            const int d = std::min(n, P.left[slot]);
            P.left[slot] -= d;
            S->cell_steps += (long long) d * (long long) P.cells;
         }
   }
   return retired;
}

// Run state (0 queued, 1 running, 2 done) and, when done, its envelope
// [jmxb][imxb] (EnvSave quantisation) and latency from submission (s).
extern "C" int slg_server_result(const slg_server *S, int run_id, short *env, double *latency_s)
{
   if (run_id < 0 || run_id >= (int) S->runs.size()) return SLG_EINPUT;
   const Run &r = *S->runs[run_id];
   if (r.state == 2) {
      if (env) memcpy(env, r.env.data(), r.env.size() * 2);
      if (latency_s) *latency_s = r.t_done - r.t_submit;
   }
   return r.state;
}

// Counters: [0] queued, [1] running, [2] done, [3] slots, [4] device bytes of
// the pools, [5] storm.cell.steps advanced.
extern "C" int slg_server_stats(const slg_server *S, long long *out6)
{
   long long q = (long long) S->queue.size(), run = 0, done = 0, slots = 0, bytes = 0;
   for (auto &r : S->runs) {
      run += r->state == 1;
      done += r->state == 2;
   }
   for (auto &P : S->pools) {
      slots += P.slots;
      bytes += (long long) P.bytes;
   }
   out6[0] = q;
   out6[1] = run;
   out6[2] = done;
   out6[3] = slots;
   out6[4] = bytes;
   out6[5] = S->cell_steps;
   return 0;
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
   for (auto &P : S->pools) {
      if (P.st) {
         (void) hipStreamSynchronize(P.st);
         (void) hipStreamDestroy(P.st);
      }
      if (P.dmem) (void) hipFree(P.dmem);
      slg_basin_close(P.basin);
   }
   delete S;
}
