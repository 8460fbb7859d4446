// sl_kernels.hip -- the SLOSH time step on CDNA4 (gfx950), FP32.
//
// Every stage of the step is written once, as a per-point device function that
// reads its neighbours through an accessor (global memory or an LDS tile), and
// restates a reference routine (file:line under sorc/slosh.fd):
//
//   contty_sq    CONTTY  contty.f:85-221 (Jacobi pass; drying squares listed)
//   k_replay     CONTTY  contty.f:146-213 serial back-step cascade, replayed
//                exactly from the first drying square, visiting only squares
//                whose inputs the cascade changed
//   pflag_pt     SMPT2G  smpt2g.f:23-45  (momentum points next to U=0 points)
//   hclass_pt    SMPT2G  smpt2g.f:51-120 (height points: copy / 888 / 999)
//   smooth_pt    FLT2G   flt2g.f:38-132 both passes + BDRYHT bdryht.f:72-86
//   hmxsv_pt     HMXSV   hmxsv.f:44-136
//   filter_pt    FILTER  filter.f:121-211, XFILTER xfilter.f:176-713
//   momntm_pt    MOMNTM  momntm.f:78-470 with FRCPNT frcpnt.f:63-200 and
//                MNTMBD mntmbd.f:58-117
//
// Two launch paths share those functions:
//
//   fused (default): k_conttyA (CONTTY over a 64x16 tile per workgroup, plus
//     the U==0 byte mask SMPT2G needs), k_replay, and k_stepB, which runs
//     SMPT2G/FLT2G/BDRYHT/HMXSV/FILTER/MOMNTM for one 64xTH tile of one storm
//     out of LDS: the tile and a halo wide enough for every stencil in the
//     chain are loaded once, each stage leaves its result in LDS for the next,
//     and only HB, HMX, UB and VB of the tile's own points go back to HBM.
//   staged: one kernel per stage over global memory (per-stage timing and an
//     independent check of the fused path).
//
// Workgroups are numbered so that each XCD (workgroup id mod 8) walks a
// contiguous run of tiles and storms: a tile's halo and the basin's static
// fields are then found in that XCD's own L2.
//
// Bit-exactness with the FP32 restatement of the reference depends on the
// operation order being the one written here: the file is compiled with
// -ffp-contract=off and correctly rounded fp32 division and square root.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sl_dev.h"

// One source, two builds: with SL_F64 every kernel computes in double (the
// reference's default build, and its results bit for bit), otherwise in float.
#ifdef SL_F64
typedef double real;
#define SL_NS sl_f64
#define SL_LAUNCH sl_dev_launch_step_f64
#else
typedef float real;
#define SL_NS sl_f32
#define SL_LAUNCH sl_dev_launch_step_f32
#endif

namespace SL_NS {
typedef SlBasinDevT<real> SlBasinDev;
typedef SlRunDevT<real> SlRunDev;

// bit equality (a store is skipped only when it would not change memory)
__device__ __forceinline__ bool same_bits(float a, float b) { return __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool same_bits(double a, double b)
{
   return __double_as_longlong(a) == __double_as_longlong(b);
}

#define TX 64
#define TY 4
// 32-bit element index of grid point (i,j): with a uniform base pointer the
// compiler addresses global memory as SGPR base + 32-bit VGPR offset.
#define IDX(B, i, j) ((uint32_t) (((j) - 1) * (B).pitch + ((i) - 1)))
// Bottom-stress record of slip class IIK and depth ID (one 32-byte line):
// AI2, SLPAI(ID), AR2, BR2, BI2, CR2, CI2 (btmstr.f:185-191, momntm.f:81-89).
enum { E_AI2, E_SLP, E_AR2, E_BR2, E_BI2, E_CR2, E_CI2 };
#define EREC(B, iik, id) ((B).tab8 + (uint32_t) ((((iik) - 1) * 600 + ((id) - 1)) * 8))
// Radial record K of a storm table set (one 32-byte line): C, S (ocean inflow),
// CW, SW (lake inflow), P (pressure / r), DELP (static height) (force1.f:171-181).
enum { R_C, R_S, R_CW, R_SW, R_P, R_DELP };

// BLOCK DATA shift tables (flw1dm_block_c.f:37-38), 1-based
__constant__ int cIP[5] = {0, 0, 1, 1, 0}, cJP[5] = {0, 0, 0, 1, 1};
__constant__ int cIH[5] = {0, -1, 0, -1, 0}, cJH[5] = {0, -1, -1, 0, 0};
__constant__ int cIIH[5] = {0, 0, 1, 0, -1}, cJJH[5] = {0, -1, 0, 1, 0};
// MOMNTM boundary offsets: one boundary point per ISKIP code, three per corner
__constant__ int cISBS[5] = {0, 0, -1, 0, 1}, cJSBS[5] = {0, -1, 0, 1, 0};
__constant__ int cISHF[4][3] = {{0, 0, 0}, {1, 2, 4}, {0, 0, 0}, {1, 3, 4}};

struct StepArgs {
   const real *hc;     // HB entering the stage (per storm base = hc + s*fstride)
   real *hn;           // HB leaving the stage
};

// The step storm s is on (1-based), or 0 when its run is complete.
__device__ __forceinline__ int storm_step(const SlRunDev &R, int s)
{
   const int c = R.cur[s];
   return c < R.nsteps[s] ? c + 1 : 0;
}

__device__ __forceinline__ const real *par(const SlRunDev &R, int s, int step)
{
   return R.params + ((size_t) s * R.maxsteps + (step - 1)) * SL_NPAR;
}

__device__ __forceinline__ const real *tabv(const SlRunDev &R, int s, const real *P)
{
   int v = (int) P[P_TVER];
   return R.tables + (size_t) (R.tbase[s] + v) * SL_NTABV * 8;
}

__device__ __forceinline__ bool zr(const SlBasinDev &B, int i, int j)
{
   return i >= 2 && i <= B.imxb - 1 && j >= 2 && j <= B.jmxb - 1;
}

// ------------------------------------------------------------------ accessors
// (i, j) are 1-based grid indices; every caller stays inside the grid.
struct GF {   // a field of one storm in global memory
   const real *p;
   int pitch;
   __device__ real operator()(int i, int j) const { return p[(uint32_t) ((j - 1) * pitch + (i - 1))]; }
};
struct LF {   // an LDS tile whose element 0 is grid point (i0, j0)
   const real *p;
   int i0, j0, w;
   __device__ real operator()(int i, int j) const { return p[(j - j0) * w + (i - i0)]; }
};
struct GB {
   const uint8_t *p;
   int pitch;
   __device__ uint8_t operator()(int i, int j) const { return p[(uint32_t) ((j - 1) * pitch + (i - 1))]; }
};
struct LB {
   const uint8_t *p;
   int i0, j0, w;
   __device__ uint8_t operator()(int i, int j) const { return p[(j - j0) * w + (i - i0)]; }
};
struct GUZ {  // "UB(i,j) == 0" straight from the transport field
   const real *ub;
   int pitch;
   __device__ bool operator()(int i, int j) const { return ub[(uint32_t) ((j - 1) * pitch + (i - 1))] == real(0.); }
};

// Static fields ZB (squares), ZBM (momentum points) and the per-point flag
// word: global or an LDS tile.
struct ZG {
   const real *zb, *zbm;
   const uint32_t *fl;
   int pitch;
   __device__ real b(int i, int j) const { return zb[(uint32_t) ((j - 1) * pitch + (i - 1))]; }
   __device__ real m(int i, int j) const { return zbm[(uint32_t) ((j - 1) * pitch + (i - 1))]; }
   __device__ uint32_t f(int i, int j) const { return fl[(uint32_t) ((j - 1) * pitch + (i - 1))]; }
};
struct ZL {
   const real *zb, *zbm;
   const uint32_t *fl;
   int i0, j0, w;
   __device__ real b(int i, int j) const { return zb[(j - j0) * w + (i - i0)]; }
   __device__ real m(int i, int j) const { return zbm[(j - j0) * w + (i - i0)]; }
   __device__ uint32_t f(int i, int j) const { return fl[(j - j0) * w + (i - i0)]; }
};

// Workgroup number -> (tile x, tile y, storm).  Hardware hands workgroup g to
// XCD g mod 8; the k-th workgroup of an XCD takes item x*q + min(x, r) + k of
// a bijection onto 0..G-1 that walks storms in groups of sg: every tile of a
// group, each tile for the group's sg storms in a row (so the tile's static
// basin data, read by each of them, is still in the XCD's L2), then the next
// group.  sg = 1 is the plain (storm, ty, tx) order.
__device__ __forceinline__ void tile_of(int g, int G, int ntx, int nty, int sg, int &tx, int &ty, int &s)
{
   const int x = g & 7, k = g >> 3, q = G >> 3, r = G & 7;
   const int lin = x * q + (x < r ? x : r) + k;
   const int T = ntx * nty, N = G / T;
   if (sg < 1) sg = 1;
   const int full = N / sg, F = full * T * sg;
   int tile, si, s0;
   if (lin < F) {
      const int grp = lin / (T * sg), w = lin - grp * T * sg;
      tile = w / sg;
      si = w - tile * sg;
      s0 = grp * sg;
   } else {   // the last, partial group of N % sg storms
      const int rem = N - full * sg, w = lin - F;
      tile = w / rem;
      si = w - tile * rem;
      s0 = full * sg;
   }
   s = s0 + si;
   tx = tile % ntx;
   ty = tile / ntx;
}

__device__ __forceinline__ int wave_sum(int v)
{
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
   return v;
}

// ---------------------------------------------------------------------- CONTTY
// Jacobi pass over one square.  HSUB is the entering HB (hc).  UB values with
// |UB| <= 1e-10 inside 2..IMXB1 x 2..JMXB1 count as 0 (and are flushed by their
// owner thread), as contty.f:90-96 does before the sweep.  Returns the new HB;
// squares that dry with a transport rescale are listed for the replay.
__device__ __forceinline__ real contty_sq(const SlBasinDev &B, const SlRunDev &R, int s, int step,
                                           int i, int j, const real *hc, real *ub, real *vb,
                                           uint8_t *uz)
{
   const uint32_t p = IDX(B, i, j);
   if (j == B.jmxb) {
      if (i <= B.imxb - 1) R.hsubtop[(size_t) s * B.pitch + (i - 1)] = hc[p];
      if (uz) uz[p] = ub[p] == real(0.);
      return hc[p];
   }
   if (i == B.imxb) {
      if (uz) uz[p] = ub[p] == real(0.);
      return hc[p];
   }
   const uint32_t fl = B.flags[p];
   real u0 = ub[p];
   if (fl & F_ZR) {
      if (fabs(u0) <= real(1.E-10)) {
         // stored only where the value is not +0 already (dry land, most of
         // them): the memory ends up the same, without the writes
         const real v0 = vb[p];
         if (!same_bits(u0, real(0.))) ub[p] = real(0.);
         if (!same_bits(v0, real(0.))) vb[p] = real(0.);
         u0 = real(0.);
      }
   }
   if (uz) uz[p] = u0 == real(0.);
   real out = hc[p];
   if (fl & F_ACT) {
      const uint32_t p2 = p + 1, p3 = p + B.pitch + 1, p4 = p + B.pitch;
      real u1 = ub[p], u2 = ub[p2], u3 = ub[p3], u4 = ub[p4];
      real v1 = vb[p], v2 = vb[p2], v3 = vb[p3], v4 = vb[p4];
      if (zr(B, i, j) && fabs(u1) <= real(1.E-10)) { u1 = real(0.); v1 = real(0.); }
      if (zr(B, i + 1, j) && fabs(u2) <= real(1.E-10)) { u2 = real(0.); v2 = real(0.); }
      if (zr(B, i + 1, j + 1) && fabs(u3) <= real(1.E-10)) { u3 = real(0.); v3 = real(0.); }
      if (zr(B, i, j + 1) && fabs(u4) <= real(1.E-10)) { u4 = real(0.); v4 = real(0.); }
      const real dux = -u1 + u2 + u3 - u4;
      const real zb = B.zb[p];
      if (!(zb <= real(30.) && dux == real(0.))) {
         const real *P = par(R, s, step);
         const real hpst = hc[p];
         const real z = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
         const real fct = P[P_X1B5] / z;
         const real dvy = -v1 - v2 + v3 + v4;
         const real hbj = hpst - fct * (dux + dvy);
         const real chng = hbj + zb;
         if (chng < real(0.)) {
            if (-chng < real(1E-3)) out = -zb;
            else if (hc[p] == -zb) out = hc[p];
            else {
               // the replay sets the final value; until then the square holds
               // what the serial sweep leaves in a square it has not assigned
               // (HSUB): a cascade may revisit it and store nothing
               out = hc[p];
               int k = atomicAdd(&R.evcnt[s], 1);
               if (k < R.evcap) R.evkey[(size_t) s * SL_EVCAP + k] = ((uint32_t) j << 16) | (uint32_t) i;
            }
         } else {
            out = hbj;
         }
      }
   }
   return out;
}

// Staged path: one thread per square.
__global__ void __launch_bounds__(TX * TY) k_contty(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (i > B.imxb || j > B.jmxb || !step) return;
   const size_t so = (size_t) s * R.fstride;
   A.hn[so + IDX(B, i, j)] = contty_sq(B, R, s, step, i, j, A.hc + so, R.ub + so, R.vb + so, nullptr);
}

// Fused path: a 64 x 16 tile per workgroup (each thread ATH / TYA rows), numbered
// XCD by XCD; also writes the U==0 mask of every point for k_stepB.
#define ATH 16
#ifndef TYA
#define TYA 8   /* rows of threads: each thread takes ATH / TYA rows */
#endif
__global__ void __launch_bounds__(TX * TYA) k_conttyA(SlBasinDev B, SlRunDev R, StepArgs A, int ntx,
                                                     int nty)
{
   int tx, ty, s;
   tile_of(blockIdx.x, gridDim.x, ntx, nty, 1, tx, ty, s);   // little static data: storm by storm
   const int i = tx * TX + threadIdx.x + 1, step = storm_step(R, s);
   if (i > B.imxb || !step) return;
   const size_t so = (size_t) s * R.fstride;
   const real *hc = A.hc + so;
   real *hn = A.hn + so, *ub = R.ub + so, *vb = R.vb + so;
   uint8_t *uz = R.uz + so;
   // The loads of the thread's rows are issued before any of its arithmetic:
   // rows do not depend on each other (a transport its owner flushes reads as
   // 0 to every square through the same |UB| <= 1e-10 test), so the values
   // are those contty_sq reads, without a dependent chain per row.
   constexpr int NR = ATH / TYA;
   struct Row {
      uint32_t fl;
      real u1, v1, u2, v2, u3, v3, u4, v4, hc, zb, sl;
   };
   Row L[NR];
   const bool in_i = i <= B.imxb - 1;
   const real ed = in_i ? B.elpdl2[i - 1] : real(0.), ec = in_i ? B.elpcl2[i - 1] : real(0.);
   for (int k = 0; k < NR; k++) {
      const int j = ty * ATH + threadIdx.y + k * TYA + 1;
      Row &l = L[k];
      l.fl = 0;
      l.u1 = l.v1 = l.u2 = l.v2 = l.u3 = l.v3 = l.u4 = l.v4 = l.hc = l.zb = l.sl = real(0.);
      if (j > B.jmxb) continue;
      const uint32_t p = IDX(B, i, j);
      l.u1 = ub[p];
      l.hc = hc[p];
      if (in_i && j <= B.jmxb - 1) {
         const uint32_t p3 = p + B.pitch;
         l.fl = B.flags[p];
         l.v1 = vb[p];
         l.u2 = ub[p + 1];
         l.v2 = vb[p + 1];
         l.u3 = ub[p3 + 1];
         l.v3 = vb[p3 + 1];
         l.u4 = ub[p3];
         l.v4 = vb[p3];
         l.zb = B.zb[p];
         l.sl = B.sinl2[j - 1];
      }
   }
   for (int k = 0; k < NR; k++) {
      const int j = ty * ATH + threadIdx.y + k * TYA + 1;
      if (j > B.jmxb) break;
      const Row &l = L[k];
      const uint32_t p = IDX(B, i, j);
      if (j == B.jmxb || !in_i) {   // the last row or column (contty_sq's first cases)
         if (j == B.jmxb && in_i) R.hsubtop[(size_t) s * B.pitch + (i - 1)] = l.hc;
         uz[p] = l.u1 == real(0.);
         hn[p] = l.hc;
         continue;
      }
      real u1 = l.u1, v1 = l.v1, u2 = l.u2, v2 = l.v2, u3 = l.u3, v3 = l.v3, u4 = l.u4, v4 = l.v4;
      if ((l.fl & F_ZR) && fabs(u1) <= real(1.E-10)) {
         if (!same_bits(u1, real(0.))) ub[p] = real(0.);
         if (!same_bits(v1, real(0.))) vb[p] = real(0.);
         u1 = real(0.);
         v1 = real(0.);
      }
      uz[p] = u1 == real(0.);
      real out = l.hc;
      if (l.fl & F_ACT) {
         if (zr(B, i + 1, j) && fabs(u2) <= real(1.E-10)) { u2 = real(0.); v2 = real(0.); }
         if (zr(B, i + 1, j + 1) && fabs(u3) <= real(1.E-10)) { u3 = real(0.); v3 = real(0.); }
         if (zr(B, i, j + 1) && fabs(u4) <= real(1.E-10)) { u4 = real(0.); v4 = real(0.); }
         const real dux = -u1 + u2 + u3 - u4;
         const real zb = l.zb;
         if (!(zb <= real(30.) && dux == real(0.))) {
            const real *P = par(R, s, step);
            const real hpst = l.hc;
            const real z = ed + (ec - ed) * l.sl;
            const real fct = P[P_X1B5] / z;
            const real dvy = -v1 - v2 + v3 + v4;
            const real hbj = hpst - fct * (dux + dvy);
            const real chng = hbj + zb;
            if (chng < real(0.)) {
               if (-chng < real(1E-3)) out = -zb;
               else if (l.hc == -zb) out = l.hc;
               else {
                  out = l.hc;   // HSUB until the replay sets it (see contty_sq)
                  int q = atomicAdd(&R.evcnt[s], 1);
                  if (q < R.evcap) R.evkey[(size_t) s * SL_EVCAP + q] = ((uint32_t) j << 16) | (uint32_t) i;
               }
            } else {
               out = hbj;
            }
         }
      }
      hn[p] = out;
   }
}

// ------------------------------------------------------------- CONTTY replay
// Exact serial semantics of contty.f:108-213 from the first drying square on.
// Squares whose corner transports were never rescaled ("clean") and that did
// not dry in the Jacobi pass give, on every serial visit, the value the Jacobi
// pass already wrote; so the replay runs the reference control flow verbatim
// but only visits drying squares, squares whose corners a rescale touched
// ("dirty", kept sorted in evkey) and the square being re-dried (IPASS).  A
// dirty square the serial sweep has not reached yet is reset to HSUB, which is
// what the reference holds for an unvisited square.  If the candidate list
// overflows, the replay falls back to visiting every square (still exact).
// The walk runs on one wavefront per storm with every lane following the same
// control flow.  Along a row the next up to 64 candidates are visited at once,
// one per lane, without side effects; the visits before the first one that
// rescales are committed (their HB stores), and the sweep continues from that
// rescale exactly as the serial code would.  The sorted candidate list lives
// in LDS; new candidates are merged into it by the whole wave.
struct Replay {
   const SlBasinDev *B;
   const real *hc;
   real *hn, *ub, *vb;
   uint8_t *uz;
   uint32_t *key;             // sorted candidate keys (j << 16 | i), LDS
   int n, full, cap;          // keys held, full-sweep mode, key capacity
   uint32_t frontier;         // largest key the sweep has reached
   real x1b5;
};

__device__ static int lbound(const uint32_t *key, int n, uint32_t k)
{
   int lo = 0, hi = n;
   while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (key[mid] < k) lo = mid + 1;
      else hi = mid;
   }
   return lo;
}

// lbound by the whole wave (every lane gets the result): the keys below k are
// counted among 64 evenly spaced samples, then inside the one gap that holds
// the boundary (n <= 4096 keeps a gap under 64 keys).
__device__ static int lbound_w(const uint32_t *key, int n, uint32_t k, int lane)
{
   if (n <= 64) return __popcll(__ballot(lane < n && key[lane] < k));
   const int stp = (n + 63) >> 6, a = lane * stp;
   const int c = __popcll(__ballot(a < n && key[a] < k));
   if (c == 0) return 0;
   const int base = (c - 1) * stp, b = base + 1 + lane;
   return base + 1 + __popcll(__ballot(lane < stp - 1 && b < n && key[b] < k));
}

__device__ __forceinline__ uint32_t rd_lane(uint32_t v, int l)
{
   return (uint32_t) __shfl((int) v, l);
}

// The candidate list overflowed: from here on the sweep visits every square,
// and every active square it has not reached yet holds HSUB again.
__device__ static void go_full(Replay &X, int lane)
{
   const SlBasinDev &B = *X.B;
   X.full = 1;
   for (int j = 1; j <= B.jmxb - 1; j++)
      for (int i0 = 1; i0 <= B.imxb - 1; i0 += 64) {
         const int i = i0 + lane;
         if (i > B.imxb - 1) continue;
         const uint32_t k = ((uint32_t) j << 16) | (uint32_t) i;
         const uint32_t p = IDX(B, i, j);
         if (k > X.frontier && (B.flags[p] & F_ACT)) X.hn[p] = X.hc[p];
      }
   __syncthreads();
}

// The nine squares around (i,j) become candidates for the rest of the sweep
// (their transports changed).  A square not yet reached by the sweep gets
// HB = HSUB back, as the serial sweep would still find it.  Lane q < 9 takes
// neighbour q; the new keys are merged into the sorted list by the wave.
// Returns whether the square below, (i, j-1), is active (the back-step test).
__device__ static bool dirty9(Replay &X, int i, int j, int lane)
{
   const SlBasinDev &B = *X.B;
   bool isnew = false, act = false;
   uint32_t k = 0;
   int pos = 0;
   if (lane < 9) {
      const int ii = i - 1 + lane % 3, jj = j - 1 + lane / 3;
      if (ii >= 1 && jj >= 1 && ii <= B.imxb - 1 && jj <= B.jmxb - 1) {
         const uint32_t p = IDX(B, ii, jj);
         const uint32_t fl = B.flags[p];
         const real hv = X.hc[p];
         act = (fl & F_ACT) != 0;
         if (act && !X.full) {
            k = ((uint32_t) jj << 16) | (uint32_t) ii;
            pos = lbound(X.key, X.n, k);
            isnew = !(pos < X.n && X.key[pos] == k);
            if (isnew && k > X.frontier) X.hn[p] = hv;
         }
      }
   }
   const bool below = (__ballot(act) >> 1) & 1;
   const uint64_t nm = __ballot(isnew);
   const int m = __popcll(nm);
   if (!m) return below;
   if (X.n + m > X.cap) {
      go_full(X, lane);
      return below;
   }
   // the new keys, in lane (= row-major = key) order, in every lane
   uint32_t nk[9];
   int cnt = 0;
   for (int q = 0; q < 9; q++) {
      const uint32_t v = rd_lane(k, q);
      if ((nm >> q) & 1) nk[cnt++] = v;
   }
   // every old key moves up by the number of new keys below it; chunks of 64
   // from the top, so no key is overwritten before it has been read
   for (int base = ((X.n - 1) >> 6) << 6; base >= 0; base -= 64) {
      const int a = base + lane;
      uint32_t v = 0;
      int sh = 0;
      if (a < X.n) {
         v = X.key[a];
         for (int r = 0; r < cnt; r++) sh += nk[r] < v;
      }
      __syncthreads();
      if (a < X.n && sh) X.key[a + sh] = v;
      __syncthreads();
   }
   if (isnew) X.key[pos + __popcll(nm & ((1ull << lane) - 1))] = k;
   X.n += m;
   __syncthreads();
   return below;
}

// What one visit of square (i,j) (the DO 320 body) does, without doing it:
// 0 inactive, 1 active and nothing stored, 2 store HB = val, 3 dries with a
// transport rescale by val.  All operands are loaded before the first test.
struct Visit {
   int kind;
   real val;
   real u[4], v[4];
};

__device__ static Visit visit_eval(const Replay &X, int i, int j, int leap, int ipass)
{
   const SlBasinDev &B = *X.B;
   Visit o;
   o.kind = 0;
   o.val = real(0.);
   const uint32_t p = IDX(B, i, j);
   const uint32_t p2 = p + 1, p3 = p + B.pitch + 1, p4 = p + B.pitch;
   const uint32_t fl = B.flags[p];
   const real zb = B.zb[p], hpst = X.hc[p];
   o.u[0] = X.ub[p];
   o.u[1] = X.ub[p2];
   o.u[2] = X.ub[p3];
   o.u[3] = X.ub[p4];
   o.v[0] = X.vb[p];
   o.v[1] = X.vb[p2];
   o.v[2] = X.vb[p3];
   o.v[3] = X.vb[p4];
   const real ed = B.elpdl2[i - 1], ec = B.elpcl2[i - 1], sl = B.sinl2[j - 1];
   if (!(fl & F_ACT)) return o;
   o.kind = 1;
   if (leap == 2 && i == ipass) {
      o.kind = 2;
      o.val = -zb;
      return o;
   }
   const real dux = -o.u[0] + o.u[1] + o.u[2] - o.u[3];
   if (zb <= real(30.) && dux == real(0.)) return o;
   const real z = ed + (ec - ed) * sl;
   const real fct = X.x1b5 / z;
   const real dvy = -o.v[0] - o.v[1] + o.v[2] + o.v[3];
   const real hbj = hpst - fct * (dux + dvy);
   const real chng = hbj + zb;
   if (chng < real(0.)) {
      if (-chng < real(1E-3)) {
         o.kind = 2;
         o.val = -zb;
         return o;
      }
   } else {
      o.kind = 2;
      o.val = hbj;
      return o;
   }
   if (hpst == -zb) return o;
   real denm = hbj - hpst;
   denm = copysign(fmax(real(1.E-5), fabs(denm)), denm);
   const real anum = -zb - hpst;
   real rfct = anum / denm;
   if (fabs(rfct) < real(1.E-5)) rfct = real(0.);
   o.kind = 3;
   o.val = rfct;
   return o;
}

// One wavefront per storm.  The wave sorts the storm's drying squares in LDS
// (bitonic, row-major key order), then walks the serial cascade.
__global__ __launch_bounds__(64) void k_replay(SlBasinDev B, SlRunDev R, StepArgs A, int use_uz)
{
   __shared__ uint32_t sk[SL_EVCAP];
   const int s = blockIdx.x, lane = threadIdx.x;
   if (s >= R.nstorm) return;
   const int step = storm_step(R, s);
   int n = R.evcnt[s];
   if (n == 0 || !step) return;
   if (n <= R.evcap) {
      const uint32_t *gk = R.evkey + (size_t) s * SL_EVCAP;
      int m = 1;
      while (m < n) m <<= 1;
      for (int e = lane; e < m; e += 64) sk[e] = e < n ? gk[e] : 0xffffffffu;
      __syncthreads();
      for (int k = 2; k <= m; k <<= 1)
         for (int h = k >> 1; h > 0; h >>= 1) {
            for (int e = lane; e < m; e += 64) {
               const int x = e ^ h;
               if (x > e) {
                  const uint32_t a = sk[e], b = sk[x];
                  if ((a > b) == ((e & k) == 0)) {
                     sk[e] = b;
                     sk[x] = a;
                  }
               }
            }
            __syncthreads();
         }
   }
   const size_t so = (size_t) s * R.fstride;
   Replay X;
   X.B = &B;
   X.hc = A.hc + so;
   X.hn = A.hn + so;
   X.ub = R.ub + so;
   X.vb = R.vb + so;
   X.uz = use_uz ? R.uz + so : nullptr;
   X.key = sk;
   X.full = 0;
   X.cap = R.evcap;
   X.x1b5 = par(R, s, step)[P_X1B5];
   if (n > R.evcap) {
      n = 0;
      X.full = 1;
   }
   X.n = n;
   const int jmxb1 = B.jmxb - 1;
   int j, ib, ifn, jump = 4, leap = 1, iset = 0, ipass = 0;
   if (X.full) {
      j = 1;
      X.frontier = 0;
      go_full(X, lane);
   } else {
      j = (int) (X.key[0] >> 16);
      X.frontier = (uint32_t) j << 16;
   }
   ib = B.is[j - 1];
   ifn = B.ie[j - 1];
   // The reference sweep ends after a bounded number of visits; the budget
   // only guards the device against a malformed bound table.  A run that
   // exhausts it is flagged (stat[7]) and its results are refused by the host.
   long budget = R.vbudget > 0 ? R.vbudget : 64L * B.imxb * B.jmxb + 4096;
L130:
   if (--budget < 0) {
      if (lane == 0) atomicOr(&R.stat[s * 8 + 7], 1);
      goto L350;
   }
   if (ifn >= ib) {
      const uint32_t row = (uint32_t) j << 16;
      int pos = X.full ? 0 : lbound_w(X.key, X.n, row | (uint32_t) ib, lane);
      const int pend = X.full ? 0 : lbound_w(X.key, X.n, (row | (uint32_t) ifn) + 1u, lane);
      int icur = ib;
      for (;;) {
         const int cnt = X.full ? ifn - icur + 1 : pend - pos;
         if (cnt <= 0) break;
         const int nb = cnt < 64 ? cnt : 64;
         const bool have = lane < nb;
         const int i = have ? (X.full ? icur + lane : (int) (X.key[pos + lane] & 0xffffu)) : 0;
         Visit o;
         o.kind = 0;
         if (have) o = visit_eval(X, i, j, leap, ipass);
         const uint64_t resc = __ballot(have && o.kind == 3);
         const int first = resc ? __ffsll((unsigned long long) resc) - 1 : nb;
         if (have && lane < first && o.kind == 2) X.hn[IDX(B, i, j)] = o.val;
         const uint64_t act = __ballot(have && o.kind != 0 && lane <= first);
         if (act) {
            const int last = 63 - __clzll((long long) act);
            const uint32_t kl = row | (uint32_t) __shfl(i, last);
            if (kl > X.frontier) X.frontier = kl;
         }
         if (leap == 2 && (__ballot(have && lane < first && o.kind == 2 && i == ipass) != 0)) leap = 1;
         if (first == nb) {
            pos += nb;
            icur += nb;
            continue;
         }
         // the first rescale: its corner transports, then the sweep backs up a row
         const int ir = __shfl(i, first);
         if (lane == first) {
            const uint32_t p = IDX(B, i, j);
            const uint32_t c[4] = {p, p + 1, p + B.pitch + 1, p + B.pitch};
            for (int q = 0; q < 4; q++) {
               const real un = o.val * o.u[q];
               X.ub[c[q]] = un;
               X.vb[c[q]] = o.val * o.v[q];
               if (X.uz) X.uz[c[q]] = un == real(0.);
            }
         }
         __syncthreads();
         const bool below = dirty9(X, ir, j, lane);
         j = j - 1;
         jump = 1;
         iset = ir;
         leap = 1;
         if (j == 0) goto L330;
         if (!below) goto L330;
         if (B.is[j - 1] > ir) goto L330;
         ib = (ir - 1 > B.is[j - 1]) ? ir - 1 : B.is[j - 1];
         ifn = (ir + 1 < B.ie[j - 1]) ? ir + 1 : B.ie[j - 1];
         goto L130;
      }
   }
L330:
   jump = jump + 1;
   leap = 1;
   if (jump == 2) {
      j = j + 1;
      leap = 2;
      ib = (iset - 1 > B.is[j - 1]) ? iset - 1 : B.is[j - 1];
      ifn = B.ie[j - 1];
      ipass = iset;
   } else {
      j = j + 1;
      if (!X.full) {   // rows without candidates change nothing: skip ahead
         int pos = lbound_w(X.key, X.n, (uint32_t) j << 16, lane);
         if (pos >= X.n) goto L350;
         int jn = (int) (X.key[pos] >> 16);
         if (jn > j) j = jn;
      }
      if (j > jmxb1) goto L350;
      if (((uint32_t) j << 16) > X.frontier) X.frontier = (uint32_t) j << 16;
      ib = B.is[j - 1];
      ifn = B.ie[j - 1];
   }
   goto L130;
L350:
   if (lane == 0) {
      R.evcnt[s] = 0;
      atomicAdd(&R.stat[s * 8 + 4], 1);
   }
}

// ------------------------------------------------------------------ SMPT2G
// Momentum point flag: bit 0 = smoothing candidate, bit 1 = next to a dry
// height point; *ndry = number of dry height points around it.
template <class UZ, class HS, class Z>
__device__ __forceinline__ uint8_t pflag_pt(const SlBasinDev &B, int i, int j, const UZ &uz,
                                            const HS &hs, const Z &z, int *ndry)
{
   uint8_t f = 0;
   *ndry = 0;
   if (z.f(i, j) & F_PFC) {
      if (!uz(i, j)) {
         if (uz(i - 1, j) || uz(i, j - 1) || uz(i, j + 1) || uz(i + 1, j)) {
            int nd = 0;
            f = 1;
            for (int k = 1; k <= 4; k++) {
               const int a = i + cIH[k], b = j + cJH[k];
               if (hs(a, b) + z.b(a, b) == real(0.)) nd++;
            }
            if (nd) f |= 2;
            *ndry = nd;
         }
      }
   }
   return f;
}

// Height point class: 0 none, 1 copy (HCT 1 ft), 2 next to dry (HCT 0 ft).
template <class HS, class PF, class Z>
__device__ __forceinline__ uint8_t hclass_pt(const SlBasinDev &B, int i, int j, const HS &hs, const PF &pf,
                                             const Z &z)
{
   if (i <= B.imxb - 1 && j <= B.jmxb - 1 && hs(i, j) + z.b(i, j) != real(0.)) {
      // momentum points having (i,j) as a corner: (i+1,j+1) (i,j+1) (i+1,j) (i,j)
      const uint8_t a = pf(i + 1, j + 1) | pf(i, j + 1) | pf(i + 1, j) | pf(i, j);
      return (a & 2) ? 2 : (a & 1) ? 1 : 0;
   }
   return 0;
}

// Metric weight of neighbour k of square p: on elliptic/hyperbolic grids
// 2z/(1+z) - 1 with z = GAMFK/GAM0 (precomputed per square, g4), on polar grids
// the constants GAMFP (flt2g.f:60-75, filter.f:140-160).
__device__ __forceinline__ real gamf_of(const SlBasinDev &B, int k, uint32_t p)
{
   return B.ebsn_ell ? B.g4[(uint32_t) (k - 1) * B.npts + p] : B.gamfp[k - 1];
}

// FLT2G (both passes, when smpt) + BDRYHT for square (i,j), j <= JMXB-1.
template <class HS, class CL, class Z>
__device__ __forceinline__ real smooth_pt(const SlBasinDev &B, const real *P, const real *T, bool smpt,
                                           int i, int j, const HS &hs, const CL &cl, const Z &z,
                                           const real *tide)
{
   const uint32_t p = IDX(B, i, j);
   real out = hs(i, j);
   if (i > B.imxb - 1) return out;
   if (smpt) {
      const uint8_t c = cl(i, j);
      if (c) {
         const real hct = (c == 1) ? real(1.) : real(0.);
         const real hsub = out, zb = z.b(i, j);
         if (!(hsub + zb <= hct)) {
            real hsm = real(0.), hdif = real(0.), hzz, gamf;
            for (int k = 1; k <= 4; k++) {
               const int ii = i + cIIH[k], jj = j + cJJH[k];
               bool self = true;
               if (ii == 0 || jj == 0) {
                  gamf = real(0.);
               } else {
                  gamf = gamf_of(B, k, p);
                  if (!(ii == B.imxb || jj == B.jmxb)) {
                     if (cl(ii, jj) == c) {                     // HSUB(n) != 999
                        const real hq = hs(ii, jj);
                        if (!(hq + z.b(ii, jj) <= hct)) {
                           const int kk = k % 4 + 1;
                           const real z1 = z.m(i + cIP[k], j + cJP[k]);
                           const real z2 = z.m(i + cIP[kk], j + cJP[kk]);
                           const real zzz = fmin(z1, z2) + hct;
                           if (c == 1) self = (hsub < zzz || hq < zzz);
                           else self = (hsub < zzz && hq < zzz);
                           if (!self) {
                              hsm = hsm + hq;
                              hzz = hq;
                           }
                        }
                     }
                  }
               }
               if (self) {
                  hsm = hsm + hsub;
                  hzz = hsub;
               }
               hdif = hdif + hzz * gamf;
            }
            const real hbtemp = (real(1.) / real(3.)) * hsub + (hsm + hdif) / real(6.);
            out = fmax(hbtemp, -zb);
         }
      }
   }
   if (z.f(i, j) & F_BDRY) {
      const real xr = B.elpcl[i - 1] * B.cosl[j - 1], yr = B.elpdl[i - 1] * B.sinl[j - 1];
      const real x = xr - P[P_C1] - P[P_AX], y = yr - P[P_C2] - P[P_AY];
      const real rsq = x * x + y * y, r1 = sqrt(rsq) / real(5280.) + real(1.);
      int k = (int) r1;
      const real r2 = (real) k, dr = r1 - r2;
      if (k > 790) k = 790;
      const real *Rk = T + (uint32_t) k * 8, *Rk1 = Rk + 8;
      out = Rk[R_DELP] + dr * (Rk1[R_DELP] - Rk[R_DELP]) + P[P_SEA] + (tide ? tide[B.bcidx[p]] : real(0.));
   }
   return out;
}

// HMXSV at square (i,j) (i <= IMXB-1, j <= JMXB-1): the new envelope value.
template <class HB, class Z>
__device__ __forceinline__ real hmxsv_pt(const SlBasinDev &B, int i, int j, const HB &hb, real hx,
                                          const Z &z)
{
   const real h = hb(i, j);
   if (!(hx < h)) return hx;
   const real zb = z.b(i, j);
   if (zb > real(40.) || B.ebsn1_amp) return h;
   if (h + zb < real(1.)) return h;
   real hsm = real(0.), hdif = real(0.), hzz, gamf;
   const uint32_t p = IDX(B, i, j);
   for (int k = 1; k <= 4; k++) {
      const int ii = i + cIIH[k], jj = j + cJJH[k];
      if (ii == 0 || jj == 0) {
         gamf = real(0.);
         hsm = hsm + h;
         hzz = h;
      } else {
         gamf = B.g4[(uint32_t) (k - 1) * B.npts + p];   // hmxsv.f:84-92
         if (jj != B.jmxb && ii != B.imxb && hb(ii, jj) + z.b(ii, jj) >= real(1.)) {
            const real hq = hb(ii, jj);
            const int kk = k % 4 + 1;
            const real z1 = z.m(i + cIP[k], j + cJP[k]);
            const real z2 = z.m(i + cIP[kk], j + cJP[kk]);
            const real zzz = fmin(z1, z2) + real(1.);
            if (h < zzz || hq < zzz) {
               hsm = hsm + h;
               hzz = h;
            } else {
               hsm = hsm + hq;
               hzz = hq;
            }
         } else {
            hsm = hsm + h;
            hzz = h;
         }
      }
      hdif = hdif + hzz * gamf;
   }
   const real htemp = (real(4.) * h + hsm + hdif) / real(8.);
   return fmax(hx, htemp);
}

// ------------------------------------------------------------ FILTER/XFILTER
// Square (i,j) with i <= IMXB-1, j <= JMXB-1; hs is the HSUB snapshot, top
// the HSUB row JMXB CONTTY saved (XFILTER reads it; IMXB column reads 0).
template <class HS, class Z>
__device__ __forceinline__ real filter_pt(const SlBasinDev &B, int i, int j, const HS &hs, const real *top,
                                           const Z &z)
{
   const real zb = z.b(i, j), hsub = hs(i, j);
   real out = hsub;
   if (zb <= -B.hter) return out;
   if (B.ismth == 1) {
      const real hcrt = real(0.5);
      if (!(hsub + zb < hcrt)) {
         real hsm = real(0.), hdif = real(0.), hzz, gamf;
         const uint32_t p = IDX(B, i, j);
         for (int k = 1; k <= 4; k++) {
            const int ii = i + cIIH[k], jj = j + cJJH[k];
            if (ii == 0 || jj == 0) {
               gamf = real(0.);
               hsm = hsm + hsub;
               hzz = hsub;
            } else {
               gamf = gamf_of(B, k, p);
               if (ii == B.imxb || jj == B.jmxb || hs(ii, jj) + z.b(ii, jj) < hcrt) {
                  hsm = hsm + hsub;
                  hzz = hsub;
               } else {
                  const real hq = hs(ii, jj);
                  const int kk = k % 4 + 1;
                  const real z1 = z.m(i + cIP[k], j + cJP[k]);
                  const real z2 = z.m(i + cIP[kk], j + cJP[kk]);
                  const real zzz = fmin(z1, z2) + hcrt;
                  if (hsub < zzz || hq < zzz) {
                     hsm = hsm + hsub;
                     hzz = hsub;
                  } else {
                     hsm = hsm + hq;
                     hzz = hq;
                  }
               }
            }
            hdif = hdif + hzz * gamf;
         }
         const real htemp = real(0.5) * hsub + (hsm + hdif) / real(8.);
         out = fmax(-zb, htemp);
      }
   } else if (!(hsub + zb < B.exclude)) {
      const real hcrt = B.hcrt;
#define HX(a, b) ((a) == B.imxb ? real(0.) : ((b) == B.jmxb ? top[(a) - 1] : hs(a, b)))
#define ZX(a, b) (z.m(a, b))
      const int IIX[5] = {0, -1, 1, 1, -1}, JJX[5] = {0, -1, -1, 1, 1};
      const int IP1[5] = {0, -1, 1, 2, 0}, JP1[5] = {0, 0, -1, 1, 2};
      const int IP2[5] = {0, 0, 2, 1, -1}, JP2[5] = {0, -1, 0, 2, 1};
      int maskii[5] = {0, 0, 0, 0, 0}, maskix[5] = {0, 0, 0, 0, 0};
      for (int k = 1; k <= 4; k++) {
         const int ii = i + cIIH[k], jj = j + cJJH[k], ix = i + IIX[k], jx = j + JJX[k];
         if (ii == 0 || jj == 0 || ix == 0 || jx == 0) continue;
         if (!(ii == B.imxb || jj == B.jmxb || HX(ii, jj) + z.b(ii, jj) < hcrt)) {
            const int kk2 = k % 4 + 1;
            const real zzz = fmin(ZX(i + cIP[k], j + cJP[k]), ZX(i + cIP[kk2], j + cJP[kk2])) + hcrt;
            if (!(hsub < zzz && HX(ii, jj) < zzz)) maskii[k] = 1;
         }
         if (!(ix == B.imxb || jx == B.jmxb || HX(ix, jx) + z.b(ix, jx) < hcrt)) {
            const int kk = k % 4 + 1, km1 = (k + 2) % 4 + 1;
            const real z0 = ZX(i + cIP[k], j + cJP[k]), zz = ZX(i + cIP[kk], j + cJP[kk]);
            const real zzz = fmin(z0, zz) + hcrt;
            const real yyy = fmin(z0, ZX(i + IP2[k], j + JP2[k])) + hcrt;
            const real xxx = fmin(z0, ZX(i + IP1[k], j + JP1[k])) + hcrt;
            const real www = fmin(z0, ZX(i + cIP[km1], j + cJP[km1])) + hcrt;
            const real hii = HX(ii, jj), hix = HX(ix, jx), ha1 = HX(i + IP1[k], j + JP1[k]);
            if (((hsub >= zzz && hii >= yyy) || (hix >= yyy && hii >= zzz)) ||
                ((hsub >= www && ha1 >= xxx) || (hix >= xxx && ha1 >= www)))
               maskix[k] = 1;
         }
      }
      real hsm2 = real(0.), xsm2 = real(0.);
      for (int km = 1; km <= 4; km++) {
         const int kkm = km % 4 + 1, km1m = (km + 2) % 4 + 1;
         const int ix = i + IIX[km], jx = j + JJX[km], ii = i + cIIH[km], jj = j + cJJH[km];
         if (ii == 0 || jj == 0 || ix == 0 || jx == 0 || i + IIX[kkm] == 0 || j + JJX[kkm] == 0 ||
             i + IIX[km1m] == 0 || j + JJX[km1m] == 0 || i + cIIH[kkm] == 0 ||
             j + cJJH[kkm] == 0 || i + cIIH[km1m] == 0 || j + cJJH[km1m] == 0)
            continue;
         hsm2 = hsm2 + (real) maskii[km] * HX(ii, jj) - (real) maskii[km] * hsub;
         if (B.xver == 1) {
            const real w = (real) (maskix[km] * maskii[km] * maskix[kkm]);
            xsm2 = xsm2 + real(0.5) * (w * HX(ix, jx) - w * hsub);
         } else {
            xsm2 = xsm2 + real(0.5) * real(1.) * (real) maskix[km] * (real) maskii[km] *
                          (real) maskii[km1m] * (HX(ix, jx) - hsub);
         }
      }
      const real htemp = hsub + real(1.) / real(8.) * (hsm2 - real(1.) * (xsm2));
      out = fmax(-zb, htemp);
#undef HX
#undef ZX
   }
   return out;
}

// FLTER2 (flter2.f:17-95) at square (i,j), i <= IMXB-1, j <= JMXB-1, of a
// closed island basin: FILTER's weights with the angular direction periodic
// (a neighbour in row 0 is in row JMXB1; row JMXB is an ordinary neighbour).
// hs is the HSUB snapshot, rows 1..JMXB.
template <class HS, class Z>
__device__ __forceinline__ real flter2_pt(const SlBasinDev &B, int i, int j, const HS &hs, const Z &z)
{
   const real zb = z.b(i, j), hsub = hs(i, j), hcrt = real(0.5);
   if (zb <= -B.hter) return hsub;
   if (hsub + zb < hcrt) return hsub;
   real hsm = real(0.), hdif = real(0.), hzz, gamf;
   const real gam0 = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
   for (int k = 1; k <= 4; k++) {
      const int ii = i + cIIH[k];
      int jj = j + cJJH[k];
      if (jj == 0) jj = B.jmxb - 1;
      if (ii == 0) {
         gamf = real(0.);
         hsm = hsm + hsub;
         hzz = hsub;
      } else {
         if (B.ebsn_ell) {
            const real gamfk = B.elpdl2[ii - 1] + (B.elpcl2[ii - 1] - B.elpdl2[ii - 1]) * B.sinl2[jj - 1];
            const real zz = gamfk / gam0;
            gamf = real(2.) * zz / (real(1.) + zz) - real(1.);
         } else {
            gamf = B.gamfp[k - 1];
         }
         if (ii == B.imxb || hs(ii, jj) + z.b(ii, jj) < hcrt) {
            hsm = hsm + hsub;
            hzz = hsub;
         } else {
            const int kk = k % 4 + 1;
            const real z1 = z.m(i + cIP[k], j + cJP[k]), z2 = z.m(i + cIP[kk], j + cJP[kk]);
            const real zzz = fmin(z1, z2) + hcrt, hq = hs(ii, jj);
            if (hsub < zzz || hq < zzz) {
               hsm = hsm + hsub;
               hzz = hsub;
            } else {
               hsm = hsm + hq;
               hzz = hq;
            }
         }
      }
      hdif = hdif + hzz * gamf;
   }
   const real htemp = real(.5) * hsub + (hsm + hdif) / real(8.);
   return fmax(-zb, htemp);
}

// ------------------------------------------------------------------ MOMNTM
struct Forcing {
   real dph;
   int id, iik;
};

__device__ static void frcpnt(const SlBasinDev &B, const real *P, const real *T, real x, real y,
                              real &fxb, real &fyb, int ncatg, real &cshltr, int npls,
                              const Forcing &F)
{
   const real xp = x - P[P_C1] - P[P_AX], yp = y - P[P_C2] - P[P_AY];
   const real rsq = xp * xp + yp * yp, rs = sqrt(rsq), r1 = rs / real(5280.) + real(1.);
   int k = (int) r1;
   const real r2 = (real) k, dr = r1 - r2;
   if (k > 790) k = 790;
   const real ccn = P[P_X15] + rsq;
   const real x1218 = (ncatg == 1) ? P[P_W0] : P[P_W1];
   real ck1, sk1;
   const real *Rk = T + (uint32_t) k * 8, *Rk1 = Rk + 8;
   if (ncatg == 1) {
      ck1 = Rk[R_CW] + dr * (Rk1[R_CW] - Rk[R_CW]);
      sk1 = Rk[R_SW] + dr * (Rk1[R_SW] - Rk[R_SW]);
   } else {
      cshltr = real(1.);
      ck1 = Rk[R_C] + dr * (Rk1[R_C] - Rk[R_C]);
      sk1 = Rk[R_S] + dr * (Rk1[R_S] - Rk[R_S]);
   }
   real a = rs * P[P_X9] - x1218 * (yp * ck1 + xp * sk1);
   real b = rs * P[P_X10] + x1218 * (xp * ck1 - yp * sk1);
   if (ncatg == 1) {
      const real rhol = fabs(xp * P[P_X20] + yp * P[P_X21]);
      const real ccc = ccn * rhol / (P[P_X15] + rhol * rhol);
      a = a + ccc * P[P_X23];
      b = b + ccc * P[P_X24];
   }
   a = a * P[P_X4] / ccn;
   b = b * P[P_X4] / ccn;
   real fctt = real(3.E-6) * sqrt(a * a + b * b);
   fctt = fctt * cshltr * P[P_BT];
   const real fxx = fctt * a, fyy = fctt * b;
   const real *E = EREC(B, F.iik, F.id);
   const real cr = E[E_CR2], ci = E[E_CI2];
   fxb = cr * fxx - ci * fyy;
   fyb = cr * fyy + ci * fxx;
   if (npls != 0) {
      const real aa = -F.dph * (Rk[R_P] + dr * (Rk1[R_P] - Rk[R_P]));
      const real pxx = aa * xp, pyy = aa * yp;
      const real br = E[E_BR2], bi = E[E_BI2];
      fxb = fxb + br * pxx - bi * pyy;
      fyb = fyb + br * pyy + bi * pxx;
   }
}

__device__ static void mntmbd(const SlBasinDev &B, const real *P, const real *T, real *ub, real *vb,
                              int jump, int ijump, int isb, int jsb, real arj, real aii, uint32_t pij,
                              real ubj, real vbj, real &fxb, real &fyb, int ncatg, real &cshltr,
                              const Forcing &F)
{
   const uint32_t q = IDX(B, isb, jsb);
   const real aiul = aii * (ub[q] - ubj), aivb = aii * (vb[q] - vbj);
   if (jump == 2) {
      const real xttx = real(0.), xtty = real(0.);
      ub[q] = ub[pij] + aiul + arj * (vb[q] - vbj) + xttx;
      vb[q] = vb[pij] + aivb - arj * (ub[q] - ub[pij]) + xtty;
      return;
   }
   real fxbp, fybp;
   if (ijump != 2) {
      const real fxd = fxb, fyd = fyb;
      real xr = B.elpct[isb - 1] * B.cost[jsb - 1], yr = B.elpdt[isb - 1] * B.sint[jsb - 1];
      frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 1, F);
      xr = B.elpdt[isb - 1] * B.cost[jsb - 1];
      yr = B.elpct[isb - 1] * B.sint[jsb - 1];
      fxb = xr * fxbp + yr * fybp;
      fyb = xr * fybp - yr * fxbp;
      const real xttx = P[P_DELT] * (fxb - fxd), xtty = P[P_DELT] * (fyb - fyd);
      ub[q] = ub[pij] + aiul + arj * (vb[q] - vbj) + xttx;
      vb[q] = vb[pij] + aivb - arj * (ub[q] - ub[pij]) + xtty;
      return;
   }
   cshltr = real(1.);
   real xr = B.elpct[isb - 1] * B.cost[jsb - 1], yr = B.elpdt[isb - 1] * B.sint[jsb - 1];
   frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 0, F);
   xr = B.elpdt[isb - 1] * B.cost[jsb - 1];
   yr = B.elpct[isb - 1] * B.sint[jsb - 1];
   fxb = xr * fxbp + yr * fybp;
   fyb = xr * fybp - yr * fxbp;
   ub[q] = aii * ub[q] + arj * vb[q] + P[P_DELT] * fxb;
   vb[q] = aii * vb[q] - arj * ub[q] + P[P_DELT] * fyb;
}

// Momentum point (i,j) with F_MOM set; hb gives the final HB of the step on the
// four squares around it.  Writes UB/VB at (i,j) and at the boundary points
// this point owns.  ect, edt: the metrics ELPCT(I), ELPDT(I) of the point's
// column (loaded once by a caller that visits several rows of one column);
// u0, v0, zcor: UB, VB and the Coriolis factor at (i,j) as the step found them
// (no other point writes them, so a caller may load them early).
template <class HB, class Z>
__device__ __forceinline__ void momntm_pt(const SlBasinDev &B, const real *P, const real *T, uint32_t fl,
                                          int i, int j, const HB &hb, const Z &z, real *ub, real *vb, real ect,
                                          real edt, real u0, real v0, real zcor)
{
   const uint32_t p = IDX(B, i, j);
   const real h1 = hb(i - 1, j - 1), h2 = hb(i, j - 1), h3 = hb(i - 1, j), h4 = hb(i, j);
   const real za = z.b(i - 1, j - 1), zb2 = z.b(i, j - 1), zc = z.b(i - 1, j), zp = z.b(i, j);
   const real zbm = z.m(i, j);
   Forcing F;
   F.iik = (int) F_IIK(fl);
   const int ncatg = (fl & F_OCEAN) ? 2 : 1;
   int jump = 1;
   real dhx, dhy, aintrp, arj, aii, fxb = real(0.), fyb = real(0.), cshltr = real(0.), ubj, vbj;
   bool blocked = false;
   if (ncatg != 1 && (fl & F_IT2) && zbm < -real(50.)) {
      dhx = -h1 + h2 - h3 + h4;
      dhy = -h1 - h2 + h3 + h4;
      F.dph = real(.25) * (h1 + h2 + h3 + h4 + za + zb2 + zc + zp);
      const real dphz = fmin(real(599.), F.dph);
      F.id = (int) fmin(real(599.), dphz);
      if (F.id <= 0) F.id = 1;
      const real dphi = (real) F.id;
      const real *E = EREC(B, F.iik, F.id);
      aintrp = E[E_AI2] + E[E_SLP] * (dphz - dphi);
   } else {
      const real hmn = fmin(fmin(fmin(h4, h3), h2), h1);
      if (hmn <= zbm) {
         const real hmxx = fmax(fmax(fmax(h4, h3), h2), h1);
         if (hmxx <= zbm) {
            blocked = true;
         } else {
            const real hf1 = fmax(h1, zbm), hf2 = fmax(h2, zbm), hf3 = fmax(h3, zbm), hf4 = fmax(h4, zbm);
            jump = 2;
            dhx = -hf1 + hf2 - hf3 + hf4;
            dhy = -hf1 - hf2 + hf3 + hf4;
         }
      } else {
         dhx = -h1 + h2 - h3 + h4;
         dhy = -h1 - h2 + h3 + h4;
      }
      if (!blocked) {
         real hpd[4];
         hpd[0] = h1 + za;
         hpd[1] = h2 + zb2;
         hpd[2] = h3 + zc;
         hpd[3] = h4 + zp;
         F.dph = real(.25) * (hpd[0] + hpd[1] + hpd[2] + hpd[3]);
         F.id = (int) fmin(real(599.), F.dph);
         if (F.dph <= real(5.)) {
            if (jump != 2) {
               for (int k = 0; k < 4; k++) {
                  if (hpd[k] < real(5.)) hpd[k] = hpd[k] * (real(.4) + real(.12) * hpd[k] - real(.008) * hpd[k] * hpd[k]);
                  else hpd[k] = hpd[k] - real(1.);
               }
               F.dph = real(.25) * (hpd[0] + hpd[1] + hpd[2] + hpd[3]);
               F.id = (int) fmin(real(599.), F.dph);
            }
            if (F.id < 1) F.id = 1;
         }
         aintrp = real(0.);
         for (int k = 0; k < 4; k++) {
            const real dz = hpd[k];
            int idd = (int) fmin(real(599.), dz);
            if (idd < 1) idd = 1;
            const real dphi = (real) idd;
            const real *E = EREC(B, F.iik, idd);
            const real aitr = E[E_AI2] + E[E_SLP] * (dz - dphi);
            aintrp = aintrp + aitr;
         }
         aintrp = real(.25) * aintrp;
      }
   }
   real uij = real(0.);
   if (blocked) {
      ub[p] = real(0.);
      vb[p] = real(0.);
      ubj = real(0.);
      vbj = real(0.);
   } else {
      aii = real(1.) + P[P_X1B4] * aintrp;
      const real *E = EREC(B, F.iik, F.id);
      arj = E[E_AR2] * P[P_X1B4] * zcor;
      const real aiu = aii * u0, aiv = aii * v0;
      const real fct = P[P_X1B6] * F.dph;
      real xttx, xtty;
      if (jump != 2) {
         if (ncatg == 2) {
            cshltr = real(1.);
         } else {
            const real hf1 = h1 - zbm, hf2 = h2 - zbm, hf3 = h3 - zbm, hf4 = h4 - zbm;
            if (!(fl & F_TREE1)) {
               cshltr = real(.25) * (fmin(hf1, real(1.)) + fmin(hf2, real(1.)) + fmin(hf3, real(1.)) + fmin(hf4, real(1.)));
            } else {
               const real hfx[4] = {hf1, hf2, hf3, hf4};
               cshltr = real(0.);
               for (int kx = 0; kx < 4; kx++) {
                  real hfxx = real(1.);
                  if (hfx[kx] <= real(0.)) hfxx = real(0.);
                  else if (hfx[kx] < real(50.)) {
                     const real q = hfx[kx] / real(50.);
                     hfxx = q * q;
                  }
                  cshltr = cshltr + hfxx;
               }
               cshltr = real(.25) * cshltr;
            }
         }
         real fxbp, fybp;
         real xr = ect * B.cost[j - 1], yr = edt * B.sint[j - 1];
         frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 1, F);
         xr = edt * B.cost[j - 1];
         yr = ect * B.sint[j - 1];
         fxb = (xr * fxbp + yr * fybp);
         fyb = (xr * fybp - yr * fxbp);
         xttx = P[P_DELT] * fxb;
         xtty = P[P_DELT] * fyb;
      } else {
         xttx = real(0.);
         xtty = real(0.);
      }
      ubj = u0;
      vbj = v0;
      const real br = E[E_BR2], bi = E[E_BI2];
      uij = aiu - fct * (br * dhx - bi * dhy) + arj * v0 + xttx;
      ub[p] = uij;
      vb[p] = aiv - fct * (br * dhy + bi * dhx) - arj * uij + xtty;
   }
   // boundary transports (momntm.f:384-458)
   const uint32_t bop = F_BOP(fl);
   if (bop == 0) return;
   int ijump = 1;
   if (bop <= 4) {
      const int isb = i + cISBS[bop], jsb = j + cJSBS[bop];
      const uint32_t q = IDX(B, isb, jsb);
      if (uij == real(0.)) {
         ub[q] = real(0.);
         vb[q] = real(0.);
      } else {
         const uint32_t fq = B.flags[q];
         if (!(fq & F_BSKIP)) {
            if (fq & F_B8) ijump = 2;
            mntmbd(B, P, T, ub, vb, jump, ijump, isb, jsb, arj, aii, p, ubj, vbj, fxb, fyb, ncatg,
                   cshltr, F);
         }
      }
   } else {
      const int icor = (bop == 5) ? 1 : 3;
      for (int kk = 0; kk < 3; kk++) {
         const int k = cISHF[icor][kk];
         const int isb = i + cIH[icor] + cIP[k], jsb = j + cJH[icor] + cJP[k];
         const uint32_t q = IDX(B, isb, jsb);
         if (uij != real(0.)) {
            const uint32_t fq = B.flags[q];
            if (fq & F_BSKIP) break;
            if (fq & F_B8) ijump = 2;
            mntmbd(B, P, T, ub, vb, jump, ijump, isb, jsb, arj, aii, p, ubj, vbj, fxb, fyb, ncatg,
                   cshltr, F);
         } else {
            ub[q] = real(0.);
            vb[q] = real(0.);
         }
      }
   }
}

// ----------------------------------------------------------- staged kernels
// The counters feed the checks against the reference's list limits (the
// 30000 / 8000 entries of smpt2g.f:31-38, 57-62, 78-96); one atomic per wave.
// Both run for the storms whose own step runs SMPT2G (the flag k_smooth reads).
__global__ void __launch_bounds__(TX * TY) k_pflag(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (!step || !((int) par(R, s, step)[P_FLAGS] & SF_SMPT)) return;
   const bool in = (i <= B.imxb && j <= B.jmxb);
   const size_t so = (size_t) s * R.fstride;
   uint8_t f = 0;
   int nd = 0;
   if (in) f = pflag_pt(B, i, j, GUZ{R.ub + so, B.pitch}, GF{A.hc + so, B.pitch}, ZG{B.zb, B.zbm, B.flags, B.pitch}, &nd);
   const int np = wave_sum(f != 0);
   nd = wave_sum(nd);
   if (threadIdx.x == 0) {
      if (np) atomicAdd(&R.stat[s * 8 + 0], np);
      if (nd) atomicAdd(&R.stat[s * 8 + 1], nd);
   }
   if (in) R.pflag[so + IDX(B, i, j)] = f;
}

__global__ void __launch_bounds__(TX * TY) k_hclass(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (!step || !((int) par(R, s, step)[P_FLAGS] & SF_SMPT)) return;
   const bool in = (i <= B.imxb && j <= B.jmxb);
   const size_t so = (size_t) s * R.fstride;
   uint8_t c = 0;
   if (in) c = hclass_pt(B, i, j, GF{A.hc + so, B.pitch}, GB{R.pflag + so, B.pitch}, ZG{B.zb, B.zbm, B.flags, B.pitch});
   const int n1 = wave_sum(c == 1), n2 = wave_sum(c == 2);
   if (threadIdx.x == 0) {
      if (n1) atomicAdd(&R.stat[s * 8 + 2], n1);
      if (n2) atomicAdd(&R.stat[s * 8 + 3], n2);
   }
   if (in) R.hclass[so + IDX(B, i, j)] = c;
}

// FLT2G + BDRYHT + the CONTTY row copy HB(I,JMXB)=HB(I,1) (contty.f:218-219).
// hc = HB after CONTTY, hn = HB after TMSTPC1.
__global__ void __launch_bounds__(TX * TY) k_smooth(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (i > B.imxb || j > B.jmxb || !step) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const real *hs = A.hc + so;
   const real *P = par(R, s, step);
   if (j == B.jmxb) {
      A.hn[so + p] = (i <= B.imxb - 1) ? hs[IDX(B, i, 1)] : hs[p];
      return;
   }
   const bool smpt = (int) P[P_FLAGS] & SF_SMPT;
   A.hn[so + p] = smooth_pt(B, P, tabv(R, s, P), smpt, i, j, GF{hs, B.pitch}, GB{R.hclass + so, B.pitch},
                            ZG{B.zb, B.zbm, B.flags, B.pitch}, R.tidesh ? R.tidesh + (size_t) s * B.nbcpt : nullptr);
}

// ------------------------------------------------------- sub-grid 1-D flow
// The sides of 1-D flow squares are visited in list order with each storm's
// state changing from one side to the next (a square may border several
// sides), so one lane per storm walks them, as the reference loops do.
__constant__ int cIHH[5] = {0, -1, 1, 0, 0}, cJHH[5] = {0, 0, 0, -1, 1};
__constant__ real cSIGA[5] = {0, -1, 1, -1, 1};

__device__ __forceinline__ real zmet(const SlBasinDev &B, int i, int j)
{
   return B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
}

// SMPT2G's last pass over the 1-D sides (smpt2g.f:101-123), after the height
// classes: a side that did not carry water both ways at the last FLW1DM_C,
// with both squares classed and either at most 1 ft above the side's lower
// barrier, makes both squares special corner points (class 2, HSUB 888).
// hc = HB after CONTTY (the HSUB value of class-1 points).
__global__ void k_f1d_mark(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s >= R.nstorm) return;
   const int step = storm_step(R, s);
   if (!step || !((int) par(R, s, step)[P_FLAGS] & SF_SMPT)) return;
   const size_t so = (size_t) s * R.fstride;
   const real *hs = A.hc + so;
   uint8_t *cl = R.hclass + so;
   const uint8_t *act = R.f1dact + (size_t) s * B.nsqrwc;
   int moved = 0;
   for (int l = 0; l < B.nsqrw; l++) {
      if (act[l]) continue;
      const int *e = B.e_int + (size_t) l * SL_EI;
      const int i = e[E1_I], j = e[E1_J];
      if (!(i > 1 && i < B.imxb - 1 && j > 1 && j < B.jmxb - 1)) continue;
      const uint32_t p = IDX(B, i, j), q = IDX(B, e[E1_II], e[E1_JJ]);
      const uint8_t ci = cl[p];
      if (!ci) continue;
      const uint8_t ce = cl[q];
      if (!ce) continue;
      const real lim = B.e_real[(size_t) l * 8 + E2_ZBMIN] + real(1.);
      const real vi = (ci == 2) ? real(888.) : hs[p], ve = (ce == 2) ? real(888.) : hs[q];
      if (vi <= lim || ve <= lim) {
         moved += (ci == 1) + (ce == 1);
         cl[p] = 2;
         cl[q] = 2;
      }
   }
   if (moved) {   // the points change lists (the corner-point limits count them)
      R.stat[s * 8 + 2] -= moved;
      R.stat[s * 8 + 3] += moved;
   }
}

__device__ __forceinline__ real dim_(real a, real c)
{
   return a > c ? a - c : real(0.);
}

// Fractional powers of CHANNL: pow in double, rounded to real.  The FP32 step
// then gets the correctly rounded double result's rounding, as the FP32
// restatement computes it; the device double pow may differ from the host
// libm's in the last bit, so channels are the one place the FP64 step is
// not bit for bit the reference's.
__device__ __forceinline__ real powr(real x, real y)
{
   return (real) pow((double) x, (double) y);
}

// CHANNL (channl.f:74-186) on HB after BDRYHT, in place; Q of each channel
// is the storm's chq.
__device__ static void channl(const SlBasinDev &B, const real *P, real *hb, real *chq)
{
   const real g = real(32.2), cdely = real(.1);
   for (int np = 0; np < B.npss; np++) {
      const real qwdt = chq[np] * P[P_DELT];
      for (int k = 0; k < 2; k++) {
         const int *c = B.ch_pt[np];
         const uint32_t p0 = IDX(B, c[0 + 2 * k], c[1 + 2 * k]), pl = IDX(B, c[4 + 2 * k], c[5 + 2 * k]);
         hb[p0] = fmax(hb[p0] - real(.5) * qwdt / B.ch_a0[np], -B.zb[p0]);
         hb[pl] = fmax(hb[pl] + real(.5) * qwdt / B.ch_al[np], -B.zb[pl]);
      }
   }
   for (int np = 0; np < B.npss; np++) {
      const int *c = B.ch_pt[np];
      const real clngth = B.ch_lnh[np] * real(5280.), gl = g * clngth;
      real y0 = real(0.), yl = real(0.), sig = real(1.);
      for (int k = 0; k < 2; k++) y0 = y0 + real(.5) * hb[IDX(B, c[0 + 2 * k], c[1 + 2 * k])];
      y0 = y0 + B.ch_dph[np];
      for (int k = 0; k < 2; k++) yl = yl + real(.5) * hb[IDX(B, c[4 + 2 * k], c[5 + 2 * k])];
      yl = yl + B.ch_dph[np];
      if (y0 <= yl) {
         const real z = y0;
         y0 = yl;
         yl = z;
         sig = real(-1.);
      }
      const real dely = y0 - yl;
      if (dely < cdely) yl = y0 - cdely;
      real hit = yl, hcrit;
      const real y = y0;
      const real af = gl + B.gb * powr(y, real(4.) / real(3.)), af3 = real(3.) * af;
      const real ae = B.aman * powr(y, real(13.) / real(3.));
      for (int it = 0;;) {   // Newton for the critical depth, at most 10 corrections
         const real hito3 = powr(hit, real(1.) / real(3.)), hit2 = hit * hit, hit3 = hit * hit2;
         const real hit10 = hit3 * hito3, hit13 = hit * hit10;
         const real cnum = af * hit3 + B.agb * hit13 - ae, cdnm = af3 * hit2 + B.agb13 * hit10;
         hcrit = hit - cnum / cdnm;
         if (fabs(hcrit - hit) >= real(.001)) {
            it = it + 1;
            if (it <= 10) {
               hit = hcrit;
               continue;
            }
         }
         break;
      }
      real qflw;
      if (yl <= hcrit) {
         qflw = sqrt(g * hcrit) * hcrit;
      } else {
         const real y0trd = powr(y0, real(1.) / real(3.)), yltrd = powr(yl, real(1.) / real(3.));
         const real y04 = y0 * y0trd, yl4 = yl * yltrd;
         const real z0 = y0 * y0, zl = yl * yl;
         const real y013 = z0 * z0 * y0trd, yl13 = zl * zl * yltrd;
         const real qflwsq = B.aman * (y013 - yl13) / (clngth + B.bman * (y04 - yl4));
         qflw = sqrt(fabs(qflwsq));
      }
      real q = qflw * sig;
      q = q * B.ch_wth[np];
      if (dely < cdely) q = q * dely / cdely;
      chq[np] = q;
   }
}

// CHANNL, then FLW1DM_C (flw1dm_block_c.f:229-430) and FLW1DM_M
// (flw1dm_block_m.f:153-397), on HB after BDRYHT (in place).  HSUB of the
// squares the sides join lives in hsq.
__global__ void k_sides(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s >= R.nstorm) return;
   const int step = storm_step(R, s);
   if (!step) return;
   const size_t so = (size_t) s * R.fstride;
   real *hb = A.hn + so, *flw = R.flwsqr + (size_t) s * B.nsqrwc, *hq = R.hsq + (size_t) s * B.nsq;
   uint8_t *act = R.f1dact + (size_t) s * B.nsqrwc;
   const real *P = par(R, s, step);
   const real *T = tabv(R, s, P);
   if (B.npss) channl(B, P, hb, R.chq + (size_t) s * 5);
   if (!B.f1d_on) return;
   const int nl = B.nsqrwc;
   // ---- continuity
   for (int l = 0; l < nl; l++) {
      const int *e = B.e_int + (size_t) l * SL_EI;
      hq[e[E1_QI]] = real(0.);
      hq[e[E1_QE]] = real(0.);
      act[l] = 0;
   }
   for (int l = 0; l < nl; l++) {
      const real f0 = flw[l];
      if (f0 == real(0.)) continue;
      const int *e = B.e_int + (size_t) l * SL_EI;
      const real *E = B.e_real + (size_t) l * 8;
      const int i = e[E1_I], j = e[E1_J], k = e[E1_K], ii = e[E1_II], jj = e[E1_JJ];
      const uint32_t p = IDX(B, i, j), q = IDX(B, ii, jj);
      real *hqi = hq + e[E1_QI], *hqe = hq + e[E1_QE];
      const real z = zmet(B, i, j), z1 = zmet(B, ii, jj);
      real chnghi = -(cSIGA[k] * P[P_X1B5] * f0 / z);
      real chnghe = -(chnghi * z / z1);
      const bool cut = l >= B.nsqrw;
      const real *C = cut ? B.e_cut + (size_t) (l - B.nsqrw) * 4 : nullptr;
      bool shaped = true;
      if (cut) {
         if (C[1] >= real(1.) && C[2] >= real(1.)) {
            chnghi = chnghi * C[0] / C[1];
            chnghe = chnghe * C[0] / C[2];
            shaped = false;
         }
      } else if (E[E2_DELCUT] >= real(1.)) {
         shaped = false;
      }
      if (shaped) {
         // the draining square first, then the filling one
         const int leap = chnghi > real(0.) ? 2 : 1;
         for (int pass = 0; pass < 2; pass++) {
            const bool interior = (leap == 1) == (pass == 0);   // square being treated
            const bool out = pass == 0;                           // its water leaves
            real chng = interior ? chnghi : chnghe;
            const real cstrt = chng, ht = interior ? hb[p] : hb[q];
            const real bk = interior ? E[E2_BANK1] : E[E2_BANK2];
            const real cutw = cut ? (interior ? C[1] : C[2]) : real(0.);
            if (out) {
               if (ht <= bk) {
                  if (cut) chng = chng * C[0] / cutw;
               } else {
                  const real delb = ht - bk;
                  if (!cut) {
                     if (-chng * E[E2_DELCUT] <= delb) chng = chng * E[E2_DELCUT];
                     else chng = chng - delb + delb / E[E2_DELCUT];
                  } else {
                     if (-chng * C[0] <= delb) chng = chng * C[0];
                     else chng = (chng * C[0] + delb) / cutw - delb;
                  }
               }
            } else {
               if (ht >= bk) {
                  chng = cut ? chng * C[0] : chng * E[E2_DELCUT];
               } else {
                  const real delb = bk - ht;
                  if (!cut) {
                     if (!(chng <= delb)) chng = delb + (chng - delb) * E[E2_DELCUT];
                  } else {
                     if (chng * C[0] <= delb * cutw) chng = chng * C[0] / cutw;
                     else chng = chng * C[0] - delb * cutw + delb;
                  }
               }
            }
            if (interior) {
               chnghi = chng;
               *hqi = *hqi + fabs(cstrt - chng);
            } else {
               chnghe = chng;
               *hqe = *hqe + fabs(cstrt - chng);
            }
         }
      }
      const real hbj = hb[p] + chnghi, hbjj = hb[q] + chnghe;
      const real zbi = B.zb[p], zbe = B.zb[q];
      const real depli = hbj + zbi, deple = hbjj + zbe;
      if (depli < real(0.)) {
         const real cii = chnghi - depli, sube = -(cii * z / z1);
         hb[p] = -zbi;
         hb[q] = fmax(hb[q] - chnghe + sube, -zbe);
         flw[l] = f0 * cii / chnghi;
      } else if (deple < real(0.)) {
         const real cee = chnghe - deple, subi = -(cee * z1 / z);
         hb[p] = fmax(hb[p] - chnghi + subi, -zbi);
         hb[q] = -zbe;
         flw[l] = f0 * cee / chnghe;
      } else {
         hb[p] = hbj;
         hb[q] = hbjj;
         act[l] = 1;
      }
   }
   // ---- momentum
   const real slpa = EREC(B, 1, 2)[E_AI2] - EREC(B, 1, 1)[E_AI2];
   for (int l = 0; l < nl; l++) {
      const int *e = B.e_int + (size_t) l * SL_EI;
      const real *E = B.e_real + (size_t) l * 8;
      const int i = e[E1_I], j = e[E1_J], k = e[E1_K], ii = e[E1_II], jj = e[E1_JJ], iik = e[E1_IIK];
      const uint32_t p = IDX(B, i, j), q = IDX(B, ii, jj);
      real *hqi = hq + e[E1_QI], *hqe = hq + e[E1_QE];
      const real hint1 = hb[p], hext1 = hb[q], zl = E[E2_ZBMIN], zbi = B.zb[p], zbe = B.zb[q];
      if (fmin(hint1, hext1) >= zl || (hint1 == -zbi && hext1 == -zbe)) {
         flw[l] = real(0.);
         *hqi = real(0.);
         *hqe = real(0.);
         continue;
      }
      const real hovt = E[E2_HWEIR];
      if (fmax(hint1, hext1) <= hovt) {
         flw[l] = real(0.);
         *hqi = real(0.);
         *hqe = real(0.);
         continue;
      }
      const real hint = fmin(fmax(hint1, hovt), zl), hext = fmin(fmax(hext1, hovt), zl);
      const int jump = (fmax(hint, hext) >= zl) ? 1 : 2;
      const real head = cSIGA[k] * (hext - hint);
      const real d1 = hint + zbi, d2 = hext + zbe;
      const real dph = real(.5) * (d1 + d2);
      if (dph < real(0.00000001)) {
         flw[l] = real(0.);
         *hqi = real(0.);
         *hqe = real(0.);
         continue;
      }
      int id = (int) fmin(real(599.), dph);
      if (id < 1) id = 1;
      const real *Ed = EREC(B, iik, id);
      real biobr = Ed[E_BI2] / Ed[E_BR2];
      if (B.nsqrw != B.nsqrs && l >= B.nsqrs && l < B.nsqrw) biobr = real(0.);
      const real hcoff = Ed[E_BR2] + Ed[E_BI2] * biobr;
      const real flwt = -(P[P_X1B1] * hcoff * dph * head);
      real frc12t = real(0.);
      if (jump != 1) {
         real cshltr;
         if (e[E1_T]) {
            const real r = dph / real(50.);
            cshltr = r * r;
         } else {
            const real cshint = real(1.) - dim_(real(1.), dim_(hint1, hovt));
            const real cshext = real(1.) - dim_(real(1.), dim_(hext1, hovt));
            const real csh2i = real(1.) - dim_(real(1.), dim_(zl, hint1));
            const real csh2e = real(1.) - dim_(real(1.), dim_(zl, hext1));
            cshltr = real(.5) * fmin(cshint + cshext, csh2i + csh2e);
         }
         int iii = i, jjj = j;
         real xr0, yr0, elpcz, elpdz;
         if (k != 3 && k != 4) {   // force point on the side
            if (k == 2) iii = i + 1;
            xr0 = B.cosl[jjj - 1];
            yr0 = B.sinl[jjj - 1];
            elpcz = B.elpct[iii - 1];
            elpdz = B.elpdt[iii - 1];
         } else {
            if (k == 4) jjj = j + 1;
            xr0 = B.cost[jjj - 1];
            yr0 = B.sint[jjj - 1];
            elpcz = B.elpcl[iii - 1];
            elpdz = B.elpdl[iii - 1];
         }
         real fxb, fyb;
         const Forcing F{dph, id, iik};
         frcpnt(B, P, T, xr0 * elpcz, yr0 * elpdz, fxb, fyb, 1, cshltr, 1, F);
         const real xr = xr0 * elpdz, yr = yr0 * elpcz;
         const real fxbp = xr * fxb + yr * fyb, fybp = xr * fyb - yr * fxb;
         if (E[E2_COS] != real(1.)) {   // stress along the side's flow direction
            fxb = E[E2_COS] * fxbp - E[E2_SIN] * fybp;
            fyb = E[E2_SIN] * fxbp + E[E2_COS] * fybp;
         } else {
            fxb = fxbp;
            fyb = fybp;
         }
         real frc1 = real(.75) * fxb, frc2 = real(.75) * fyb;
         if (k == 3 || k == 4) {
            const real t = frc2;
            frc2 = -frc1;
            frc1 = t;
         }
         frc12t = P[P_DELT] * (frc1 + biobr * frc2);
      }
      real aintrp = real(0.);
      const real dd[2] = {d1, d2};
      for (int kk = 0; kk < 2; kk++) {
         const int idd = (int) fmin(real(299.), dd[kk]);
         real aitr;
         if (idd <= 1) {
            aitr = EREC(B, B.sfl ? 1 : iik, 1)[E_AI2] + slpa * (dd[kk] - real(1.));
         } else {
            const real *a0 = EREC(B, iik, idd), *a1 = EREC(B, iik, idd + 1);
            const real slpai = a1[E_AI2] - a0[E_AI2], dphi = (real) idd;
            aitr = a0[E_AI2] + slpai * (dd[kk] - dphi);
         }
         aintrp = aintrp + aitr;
      }
      aintrp = real(.5) * aintrp;
      const real flwcof = P[P_X1B4] * (aintrp - Ed[E_AR2] * biobr) + real(1.);
      const real fl = flwt + flwcof * flw[l] + frc12t;
      const real fwd = *hqe - *hqi;
      if (fwd != real(0.)) {
         real fd2 = fabs(-(cSIGA[k] * fwd * P[P_X1B2] / (dph * dph)));
         const real z = zmet(B, i, j), z1 = zmet(B, ii, jj);
         const real xy = real(.5) * (real(1.) / z + real(1.) / z1);
         fd2 = fd2 * xy;
         real f1 = fl;
         for (int itr = 0; itr < 20; itr++) f1 = fl / (real(1.) + fd2 * fabs(f1));
         flw[l] = f1;
      } else {
         flw[l] = fl;
      }
   }
}

__global__ void __launch_bounds__(TX * TY) k_hmxsv(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb - 1 || j > B.jmxb - 1 || !storm_step(R, s)) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const real hx = R.hmx[so + p];
   const real v = hmxsv_pt(B, i, j, GF{A.hc + so, B.pitch}, hx, ZG{B.zb, B.zbm, B.flags, B.pitch});
   if (!same_bits(v, hx)) R.hmx[so + p] = v;
}

__global__ void __launch_bounds__(TX * TY) k_filter(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb || j > B.jmxb || !storm_step(R, s)) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const real *hs = A.hc + so;
   real out = hs[p];
   if (i <= B.imxb - 1 && j <= B.jmxb - 1)
      out = filter_pt(B, i, j, GF{hs, B.pitch}, R.hsubtop + (size_t) s * B.pitch, ZG{B.zb, B.zbm, B.flags, B.pitch});
   A.hn[so + p] = out;
}

__global__ void __launch_bounds__(TX * TY) k_momntm(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (i > B.imxb || j > B.jmxb || !step) return;
   const uint32_t fl = B.flags[IDX(B, i, j)];
   if (!(fl & F_MOM)) return;
   const size_t so = (size_t) s * R.fstride;
   const real *P = par(R, s, step);
   const uint32_t p = IDX(B, i, j);
   momntm_pt(B, P, tabv(R, s, P), fl, i, j, GF{A.hc + so, B.pitch}, ZG{B.zb, B.zbm, B.flags, B.pitch}, R.ub + so, R.vb + so,
             B.elpct[i - 1], B.elpdt[i - 1], R.ub[so + p], R.vb[so + p], B.zcor[p]);
}

// ------------------------------------------------------ closed island basins
// Does storm s filter on its current step (dyn: its own flag; else filt)?
__device__ __forceinline__ bool filters(const SlRunDev &R, int s, int step, int dyn, int filt)
{
   return dyn ? ((int) par(R, s, step)[P_FLAGS] & SF_FILT) != 0 : filt != 0;
}

// FLTER2 pass: HB after HMXSV (hc, the HSUB snapshot) -> hn; the last column
// and row keep HSUB (row JMXB then takes the filtered row 1, k_island_hb).
__global__ void __launch_bounds__(TX * TY) k_filter2(SlBasinDev B, SlRunDev R, StepArgs A, int dyn, int filt)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (i > B.imxb || j > B.jmxb || !step || !filters(R, s, step, dyn, filt)) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const real *hs = A.hc + so;
   real out = hs[p];
   if (i <= B.imxb - 1 && j <= B.jmxb - 1) out = flter2_pt(B, i, j, GF{hs, B.pitch}, ZG{B.zb, B.zbm, B.flags, B.pitch});
   A.hn[so + p] = out;
}

// After FLTER2: HB(I,JMXB) = HB(I,1) (flter2.f:96-97), from the filtered
// field src into dst: the whole field with the row copy (fused path: back into
// the buffer the step's HB lives in), or row JMXB only (rowonly; staged path,
// src == dst).
__global__ void __launch_bounds__(TX * TY) k_island_hb(SlBasinDev B, SlRunDev R, const real *src, real *dst,
                                                       int dyn, int filt, int rowonly)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z, step = storm_step(R, s);
   if (i > B.imxb || j > B.jmxb || !step || !filters(R, s, step, dyn, filt)) return;
   if (rowonly && j != B.jmxb) return;
   const size_t so = (size_t) s * R.fstride;
   dst[so + IDX(B, i, j)] = src[so + (j == B.jmxb ? IDX(B, i, 1) : IDX(B, i, j))];
}

// After MOMNTM: UB(I,1) = UB(I,JMXB), VB(I,1) = VB(I,JMXB) (momntm.f:464-469).
__global__ void k_island_uv(SlBasinDev B, SlRunDev R)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x + 1, s = blockIdx.y;
   if (i > B.imxb || !storm_step(R, s)) return;
   const size_t so = (size_t) s * R.fstride;
   R.ub[so + IDX(B, i, 1)] = R.ub[so + IDX(B, i, B.jmxb)];
   R.vb[so + IDX(B, i, 1)] = R.vb[so + IDX(B, i, B.jmxb)];
}

// ------------------------------------------------------------ fused stage B
// One workgroup = one tile of one storm: 512 points, (512 / TH) x TH (64 x 8,
// 32 x 16 or 16 x 32; the autotuner picks).  Regions (halo below / above the
// tile, in points, each direction):
//   F  final HB MOMNTM reads                 1 / 0
//   S  HB after FLT2G+BDRYHT (HMXSV, filter) SLO / SHI  (2/2 with a filter, else 1/1)
//   C  height classes FLT2G reads            S + 1
//   PF momentum flags the classes read       C below, C + 1 above
//   UZ U==0 mask the flags read              PF + 1
//   HS HB after CONTTY                       PF + 1 below, PF above (S without SMPT2G)
//   Z  ZB and ZBM                            S + 2 (the XFILTER masks reach two
//                                            points up; the SMPT2G flags one more)
// Every stage computes its whole region (points inside the grid), so each
// point's inputs are in LDS; only the tile's own points are written back.
template <int LO, int HI, int TH>
struct Reg {
   static constexpr int lo = LO, hi = HI, w = 512 / TH + LO + HI, h = TH + LO + HI, n = w * h;
};

// SMPT / FILT size the regions and, with dyn == 0, decide the stages; with
// dyn != 0 (storms at different steps in one launch) each storm's own step
// flags decide, inside the regions of the <TH, true, true> instance.
// The FP32 instances without FILTER are held to 64 VGPRs, so that 8
// waves fit on a SIMD (their LDS allows it; the kernel is latency-bound and
// runs 7 % faster so); the FP64 instances to 96 VGPRs, 5 waves (their LDS
// allows five workgroups of four waves on a CU).
template <int TH, bool SMPT, bool FILT>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(real) == 8 ? 5 : !FILT ? 8 : 1)))
k_stepB(SlBasinDev B, SlRunDev R, StepArgs A, int ntx, int nty, int dyn, int isl, int post)
{
   constexpr int TW = 512 / TH;   // tile width: every tile has 512 points
   constexpr int SLO = FILT ? 2 : 1, SHI = FILT ? 2 : 1;
   using RS = Reg<SLO, SHI, TH>;
   using RC = Reg<SLO + 1, SHI + 1, TH>;
   using RP = Reg<SLO + 1, SHI + 2, TH>;
   using RU = Reg<SLO + 2, SHI + 3, TH>;
   using RH = Reg<SMPT ? SLO + 2 : SLO, SMPT ? SHI + 2 : SHI, TH>;
   using RF = Reg<1, 0, TH>;
   using RZ = Reg<SLO + 2, SHI + 2, TH>;
   // HB after CONTTY is read only up to the FLT2G stage; the filtered HB that
   // FILTER writes afterwards takes its place (less LDS: FP64 tiles fit five
   // workgroups on a CU instead of four)
   __shared__ real sh_hs[RH::n > RF::n ? RH::n : RF::n];
   __shared__ real sh_s[RS::n];
   real *const sh_f = sh_hs;
   __shared__ uint8_t sh_uz[SMPT ? RU::n : 1];
   __shared__ uint8_t sh_pf[SMPT ? RP::n : 1];
   __shared__ uint8_t sh_cl[SMPT ? RC::n : 1];
   __shared__ real sh_zb[RZ::n];
   __shared__ real sh_zbm[RZ::n];
   __shared__ uint32_t sh_fl[RZ::n];
   __shared__ int sh_cnt[4];

   int tx, ty, s;
   tile_of(blockIdx.x, gridDim.x, ntx, nty, R.sgroup, tx, ty, s);
   const int step = storm_step(R, s);
   if (!step) return;   // the whole workgroup: its storm's run is complete
   const int tid = threadIdx.x;
   const int I0 = tx * TW + 1, J0 = ty * TH + 1;
   const int IM = B.imxb, JM = B.jmxb;
   const size_t so = (size_t) s * R.fstride;
   const real *hs_g = A.hc + so;
   const real *P = par(R, s, step);
   const real *T = tabv(R, s, P);
   const int sf = (int) P[P_FLAGS];
   const real *tide = R.tidesh ? R.tidesh + (size_t) s * B.nbcpt : nullptr;
   bool smpt = SMPT && (!dyn || (sf & SF_SMPT));
   // a closed island basin filters (FLTER2) and runs MOMNTM in kernels of
   // their own: its periodic neighbours and row copies cross tiles
   const bool filt = FILT && !isl && (!dyn || (sf & SF_FILT));
   if (tid < 4) sh_cnt[tid] = 0;

   // HB after CONTTY (and the U==0 mask) with halo; the static ZB and ZBM.
   // The loops are unrolled over a compile-time element count so each
   // thread's loads are in flight together.  With post (basins with 1-D flow
   // sides or channels) A.hc already holds HB after TMSTPC1: SMPT2G, FLT2G,
   // BDRYHT and the side passes ran in kernels of their own, and this kernel
   // starts at HMXSV.
#pragma unroll
   for (int k = 0; k < (RH::n + 255) / 256; k++) {
      const int e = tid + 256 * k;
      if (e < RH::n && !post) {
         const int gi = I0 - RH::lo + e % RH::w, gj = J0 - RH::lo + e / RH::w;
         sh_hs[e] = (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM) ? hs_g[IDX(B, gi, gj)] : real(0.);
      }
   }
#pragma unroll
   for (int k = 0; k < (RZ::n + 255) / 256; k++) {
      const int e = tid + 256 * k;
      if (e < RZ::n) {
         const int gi = I0 - RZ::lo + e % RZ::w, gj = J0 - RZ::lo + e / RZ::w;
         const bool in = gi >= 1 && gi <= IM && gj >= 1 && gj <= JM;
         sh_zb[e] = in ? B.zb[IDX(B, gi, gj)] : real(0.);
         sh_zbm[e] = in ? B.zbm[IDX(B, gi, gj)] : real(0.);
         sh_fl[e] = in ? B.flags[IDX(B, gi, gj)] : 0u;
      }
   }
   // the envelope of the thread's two tile squares: only this thread reads or
   // writes them, so their loads go out now and land while SMPT2G runs
   static_assert(TW * TH == 512, "two tile points per thread");
   real *hmx = R.hmx + so;
   real hx2[2];
#pragma unroll
   for (int k = 0; k < 2; k++) {
      const int e = tid + 256 * k, gi = I0 + e % TW, gj = J0 + e / TW;
      hx2[k] = (gi <= IM - 1 && gj <= JM - 1) ? hmx[IDX(B, gi, gj)] : real(0.);
   }
   int anyz = 0;
   if (smpt) {
      const uint8_t *uz_g = R.uz + so;
#pragma unroll
      for (int k = 0; k < (RU::n + 255) / 256; k++) {
         const int e = tid + 256 * k;
         if (e < RU::n) {
            const int gi = I0 - RU::lo + e % RU::w, gj = J0 - RU::lo + e / RU::w;
            const uint8_t u = (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM) ? uz_g[IDX(B, gi, gj)] : 0;
            sh_uz[e] = u;
            anyz |= u;
         }
      }
   }
   // Without a U == 0 point in reach no momentum point is flagged, no height
   // point classed and FLT2G changes nothing: the tile skips SMPT2G/FLT2G
   // (open water, most of a basin).
   if (!__syncthreads_or(anyz)) smpt = false;
   const LF hsL{sh_hs, I0 - RH::lo, J0 - RH::lo, RH::w};
   const LB clL{sh_cl, I0 - RC::lo, J0 - RC::lo, RC::w};
   const ZL zL{sh_zb, sh_zbm, sh_fl, I0 - RZ::lo, J0 - RZ::lo, RZ::w};
   if (smpt) {
      const LB uzL{sh_uz, I0 - RU::lo, J0 - RU::lo, RU::w};
      int np = 0, ndt = 0;
      for (int e = tid; e < RP::n; e += 256) {
         const int gi = I0 - RP::lo + e % RP::w, gj = J0 - RP::lo + e / RP::w;
         uint8_t f = 0;
         if (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM) {
            int nd;
            f = pflag_pt(B, gi, gj, uzL, hsL, zL, &nd);
            if (gi >= I0 && gi < I0 + TW && gj >= J0 && gj < J0 + TH) {
               np += f != 0;
               ndt += nd;
            }
         }
         sh_pf[e] = f;
      }
      np = wave_sum(np);
      ndt = wave_sum(ndt);
      if ((tid & 63) == 0) {
         if (np) atomicAdd(&sh_cnt[0], np);
         if (ndt) atomicAdd(&sh_cnt[1], ndt);
      }
      __syncthreads();
      const LB pfL{sh_pf, I0 - RP::lo, J0 - RP::lo, RP::w};
      int n1 = 0, n2 = 0;
      for (int e = tid; e < RC::n; e += 256) {
         const int gi = I0 - RC::lo + e % RC::w, gj = J0 - RC::lo + e / RC::w;
         uint8_t c = 0;
         if (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM) {
            c = hclass_pt(B, gi, gj, hsL, pfL, zL);
            if (gi >= I0 && gi < I0 + TW && gj >= J0 && gj < J0 + TH) {
               n1 += c == 1;
               n2 += c == 2;
            }
         }
         sh_cl[e] = c;
      }
      n1 = wave_sum(n1);
      n2 = wave_sum(n2);
      if ((tid & 63) == 0) {
         if (n1) atomicAdd(&sh_cnt[2], n1);
         if (n2) atomicAdd(&sh_cnt[3], n2);
      }
      __syncthreads();
      if (tid < 4 && sh_cnt[tid]) atomicAdd(&R.stat[s * 8 + tid], sh_cnt[tid]);
   }
   // FLT2G + BDRYHT + row copy
   for (int e = tid; e < RS::n; e += 256) {
      const int gi = I0 - RS::lo + e % RS::w, gj = J0 - RS::lo + e / RS::w;
      real v = real(0.);
      if (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM) {
         if (post) v = hs_g[IDX(B, gi, gj)];
         else if (gj == JM) v = (gi <= IM - 1) ? hs_g[IDX(B, gi, 1)] : hsL(gi, gj);
         else v = smooth_pt(B, P, T, smpt, gi, gj, hsL, clL, zL, tide);
      }
      sh_s[e] = v;
   }
   __syncthreads();
   const LF sL{sh_s, I0 - RS::lo, J0 - RS::lo, RS::w};
   // UB, VB and the Coriolis factor of the thread's two momentum points: only
   // this thread reads or writes them (MNTMBD writes boundary points, which are
   // no momentum points), so they load now and land during HMXSV and FILTER
   real *ub = R.ub + so, *vb = R.vb + so;
   real u0[2], v0[2], zc0[2];
#pragma unroll
   for (int k = 0; k < 2; k++) {
      const int e = tid + 256 * k, gi = I0 + e % TW, gj = J0 + e / TW;
      u0[k] = v0[k] = zc0[k] = real(0.);
      if (!isl && gi <= IM && gj <= JM && (zL.f(gi, gj) & F_MOM)) {
         const uint32_t p = IDX(B, gi, gj);
         u0[k] = ub[p];
         v0[k] = vb[p];
         zc0[k] = B.zcor[p];
      }
   }
   // HMXSV on the tile's own squares (HMX loaded at the start)
#pragma unroll
   for (int k = 0; k < 2; k++) {
      const int e = tid + 256 * k, gi = I0 + e % TW, gj = J0 + e / TW;
      if (gi <= IM - 1 && gj <= JM - 1) {
         const real v = hmxsv_pt(B, gi, gj, sL, hx2[k], zL);
         if (!same_bits(v, hx2[k])) hmx[IDX(B, gi, gj)] = v;
      }
   }
   real *hn_g = A.hn + so;
   if (filt) {
      const real *top = R.hsubtop + (size_t) s * B.pitch;
      for (int e = tid; e < RF::n; e += 256) {
         const int gi = I0 - RF::lo + e % RF::w, gj = J0 - RF::lo + e / RF::w;
         real v = real(0.);
         if (gi >= 1 && gi <= IM && gj >= 1 && gj <= JM)
            v = (gi <= IM - 1 && gj <= JM - 1) ? filter_pt(B, gi, gj, sL, top, zL) : sL(gi, gj);
         sh_f[e] = v;
      }
      __syncthreads();
   }
   const LF fL{filt ? sh_f : sh_s, I0 - (filt ? RF::lo : RS::lo), J0 - (filt ? RF::lo : RS::lo),
               filt ? RF::w : RS::w};
#pragma unroll
   for (int k = 0; k < 2; k++) {
      const int e = tid + 256 * k;
      int gi = I0 + e % TW;
      // the column is the same on both passes; kept opaque so the compiler
      // loads its metrics on each pass instead of holding their addresses
      // across MOMNTM (which spills them at 64 VGPRs)
      asm volatile("" : "+v"(gi));
      const int gj = J0 + e / TW;
      if (gi <= IM && gj <= JM) {
         hn_g[IDX(B, gi, gj)] = fL(gi, gj);
         const uint32_t fl = zL.f(gi, gj);
         if (!isl && (fl & F_MOM))
            momntm_pt(B, P, T, fl, gi, gj, fL, zL, ub, vb, B.elpct[gi - 1], B.elpdt[gi - 1], u0[k], v0[k], zc0[k]);
      }
   }
}

// ------------------------------------------------------------------ launch
template <int TH>
static void launch_B(int smpt, int filt, int dyn, int post, dim3 g, hipStream_t st, const SlBasinDev &B,
                     const SlRunDev &R, StepArgs a, int ntx, int nty)
{
   const int isl = B.island;
   if (isl) filt = 0;   // FLTER2 runs in its own kernel
   if (post) {          // from HMXSV on: no SMPT2G regions
      if (dyn || filt)
         hipLaunchKernelGGL((k_stepB<TH, false, true>), g, dim3(256), 0, st, B, R, a, ntx, nty, dyn, isl, 1);
      else hipLaunchKernelGGL((k_stepB<TH, false, false>), g, dim3(256), 0, st, B, R, a, ntx, nty, 0, isl, 1);
   } else if (dyn || (smpt && filt))
      hipLaunchKernelGGL((k_stepB<TH, true, true>), g, dim3(256), 0, st, B, R, a, ntx, nty, dyn, isl, 0);
   else if (smpt) hipLaunchKernelGGL((k_stepB<TH, true, false>), g, dim3(256), 0, st, B, R, a, ntx, nty, 0, isl, 0);
   else if (filt) hipLaunchKernelGGL((k_stepB<TH, false, true>), g, dim3(256), 0, st, B, R, a, ntx, nty, 0, isl, 0);
   else hipLaunchKernelGGL((k_stepB<TH, false, false>), g, dim3(256), 0, st, B, R, a, ntx, nty, 0, isl, 0);
}

// ------------------------------------------------------------------ products
// Envelope of storm s at grid point (i,j) as EnvSave writes it (slosh2.c:
// 1611-1671): int16 tenths of a foot, (short)(HMX*10+.5) in double, 999 where
// HMX+ZB == 0 (dry), 0 on the last row and column.  zbd is ZB in double, as
// the reference's C driver holds it.
template <class T>
__device__ __forceinline__ short env_val(const SlBasinDev &B, const T *hmx, const double *zbd, int i, int j)
{
   if (i == B.imxb || j == B.jmxb) return 0;
   const uint32_t p = IDX(B, i, j);
   const double h = (double) hmx[p];
   if (h + zbd[p] == 0.0) return 999;
   return (short) (int) (h * 10 + .5);
}

// Envelopes of storms s0..s0+ns-1 (env, [ns][jmxb][imxb], may be null) and
// their MOM, the maximum of MEOWs (gui/tclsrc/meowgen.tcl:34-110): per point
// the largest envelope value below 50 ft over the storms, 999 where none.
// With acc the MOM continues from the values already in mom (storms taken in
// chunks).
__global__ void k_envelopes(SlBasinDev B, SlRunDev R, const double *hx2, const double *zbd, int s0, int ns,
                            short *env, short *mom, int acc)
{
   const int cell = blockIdx.x * blockDim.x + threadIdx.x, n = B.imxb * B.jmxb;
   if (cell >= n) return;
   const int i = cell % B.imxb + 1, j = cell / B.imxb + 1;
   int m = 999;
   bool any = false;
   if (mom && acc) {
      m = mom[cell];
      any = m != 999;
   }
   for (int s = s0; s < s0 + ns; s++) {
      const short v = hx2 ? env_val(B, hx2 + (size_t) s * R.fstride, zbd, i, j)
                          : env_val(B, R.hmx + (size_t) s * R.fstride, zbd, i, j);
      if (env) env[(size_t) (s - s0) * n + cell] = v;
      if (v < 500 && (!any || v > m)) {
         m = v;
         any = true;
      }
   }
   if (mom) mom[cell] = (short) (any ? m : 999);
}

// A serving slot takes a new run whose initial water (INTLHT) was copied into
// hb0: HB in both buffers, HMX = HB on squares and 0 on the last row and column
// (intlht.f:172-177), UB = VB = 0, the slot's counters cleared, and its step
// count and first radial-table set set.  Queued on the pool's stream after the
// steps of the slot's previous run, so no launch sees a half-made slot.
__global__ void k_slot_init(SlBasinDev B, SlRunDev R, int slot, int tstart, int nsteps)
{
   const int p = blockIdx.x * blockDim.x + threadIdx.x;
   const size_t so = (size_t) slot * R.fstride;
   if (p < B.npts) {
      const int i = p % B.pitch + 1, j = p / B.pitch + 1;
      const real h = R.hb0[so + p];
      R.hb1[so + p] = h;
      R.hmx[so + p] = (i < B.imxb && j < B.jmxb) ? h : real(0.);
      R.ub[so + p] = real(0.);
      R.vb[so + p] = real(0.);
   }
   if (p < B.pitch) R.hsubtop[(size_t) slot * B.pitch + p] = real(0.);
   if (p < 8) R.stat[slot * 8 + p] = 0;
   // 1-D flow state: FLWSQR = 0 (intlht.f:190-191), F1DACT not 'T', Q = 0
   // (channl.f:73)
   if (p < B.nsqrwc) {
      R.flwsqr[(size_t) slot * B.nsqrwc + p] = real(0.);
      R.f1dact[(size_t) slot * B.nsqrwc + p] = 0;
   }
   if (p < B.nsq) R.hsq[(size_t) slot * B.nsq + p] = real(0.);
   if (p < 5) R.chq[slot * 5 + p] = real(0.);
   if (p == 0) {
      R.evcnt[slot] = 0;
      R.cur[slot] = 0;
      ((int *) R.nsteps)[slot] = nsteps;
      ((int *) R.tbase)[slot] = tstart;
   }
}

// End of a step: every storm still running has completed one more step.
__global__ void k_advance(SlRunDev R)
{
   const int s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s < R.nstorm && R.cur[s] < R.nsteps[s]) R.cur[s]++;
}

// ------------------------------------------------------------- storm schedule
// The per-step storm values of every run, generated here from each storm's
// program (layout in sl_dev.h) instead of on the host.  One thread per storm
// walks its steps in order.  Step n+1 reads the state STMVAL left after its
// n-th call (stmval.f:95-236): the call's hourly advance and table refresh
// arrive as events (the whole state after that call), and every call moves
// the storm, AX = AX + X12(11), AY = AY + X12(12) (stmval.f:175-178), in
// double and in that order, so the sums are the host's bit for bit.  BT is
// TTIMER's ramp (ttimer.f:36-41) from the program's table; the step flags are
// TMSTPC1's (tmstpc1.f:55-92).  Values are rounded to R as the host rounds them.
__global__ void __launch_bounds__(64) k_sched(const double *prog, const long *off, int nstorm, real *params,
                                              int maxsteps, int ismpt2g, int incsm, int filt_always)
{
   const int s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s >= nstorm) return;
   const double *g = prog + off[s];
   const int nsteps = (int) g[0], ic19 = (int) g[1], nev = (int) g[2];
   double st[SQ_N];
#pragma unroll
   for (int k = 0; k < SQ_N; k++) st[k] = g[SQ_HDR + k];
   const double *bt = g + SQ_BT;                         // bt[ITIME - 1]
   const double *ev = bt + (ic19 > 1 ? ic19 - 1 : 0);
   int e = 0, next = nev > 0 ? (int) ev[0] : -1;
   real *P = params + (size_t) s * maxsteps * SL_NPAR;
   for (int n = 0; n < nsteps; n++) {
      if (n > 0) {   // STMVAL call n
         if (n == next) {
#pragma unroll
            for (int k = 0; k < SQ_AX; k++) st[k] = ev[(size_t) e * SQ_EV + 1 + k];
            e++;
            next = e < nev ? (int) ev[(size_t) e * SQ_EV] : -1;
         }
         st[SQ_AX] = st[SQ_AX] + st[SQ_X11];
         st[SQ_AY] = st[SQ_AY] + st[SQ_X12];
      }
      const int itime = n + 1;
      int fl = 0;
      if (ismpt2g && itime > 5 && incsm != 1) fl |= SF_SMPT;
      if ((itime + 1) % incsm == 0 || filt_always) fl |= SF_FILT;
      real v[SL_NPAR];
      v[P_BT] = (real) (itime < ic19 ? bt[itime - 1] : 1.);
      v[P_DELT] = (real) st[SQ_DELT];
      v[P_X1B4] = (real) st[SQ_X1B3];
      v[P_X1B5] = (real) st[SQ_X1B4];
      v[P_X1B6] = (real) st[SQ_X1B5];
      v[P_C1] = (real) st[SQ_C1];
      v[P_C2] = (real) st[SQ_C2];
      v[P_AX] = (real) st[SQ_AX];
      v[P_AY] = (real) st[SQ_AY];
      v[P_X4] = (real) st[SQ_X4];
      v[P_X9] = (real) st[SQ_X9];
      v[P_X10] = (real) st[SQ_X10];
      v[P_X15] = (real) st[SQ_X15];
      v[P_X20] = (real) st[SQ_X20];
      v[P_X21] = (real) st[SQ_X21];
      v[P_X23] = (real) st[SQ_X23];
      v[P_X24] = (real) st[SQ_X24];
      v[P_W0] = (real) st[SQ_W0];
      v[P_W1] = (real) st[SQ_W1];
      v[P_SEA] = (real) st[SQ_SEA];
      v[P_ITIME] = (real) itime;
      v[P_TVER] = (real) st[SQ_TV];
      v[P_FLAGS] = (real) fl;
      v[P_X1B1] = (real) st[SQ_X1B0];
      v[P_X1B2] = (real) st[SQ_X1B1];
      v[P_PAD] = real(0.);
      real *p = P + (size_t) n * SL_NPAR;
#pragma unroll
      for (int k = 0; k < SL_NPAR; k++) p[k] = v[k];
   }
}

// ------------------------------------------------------------------ rex frames
// One rex v2 frame per storm: the values of RexSaveStep's basin loop
// (rex.c:453-499, cells I-outer and J-inner over (IMXB-1) x (JMXB-1)) as the
// 10-bit codes of Stuff2_xxx (pack.c:185-190), grouped into run-length records
// of at most 256 (8 bits run-1, then the code), packed most significant bit
// first.  A workgroup of 1024 threads per storm finds the run starts with a
// block scan, then the records (a run of L values makes ceil(L/256)), then
// writes each output byte from the one or two records under it.
#define RX_T 1024

// code of frame cell k: HHB = HMX where HMX < HB, else HB (tmstpc2.f:74-85),
// HMX alone (hbs == nullptr), or the windowed maxima (hx2, double); in double,
// as the C driver's arrays
__device__ __forceinline__ int rex_code(const SlBasinDev &B, const real *hbs, const real *hmx, const double *hx2,
                                        const double *zbd, int k, int J1)
{
   const int i = k / J1, j = k - i * J1;
   const uint32_t p = IDX(B, i + 1, j + 1);
   double hd;
   if (hx2) {
      hd = hx2[p];
   } else {
      real h = hmx[p];
      if (hbs) {
         const real b = hbs[p];
         if (!(h < b)) h = b;
      }
      hd = (double) h;
   }
   const double zd = zbd[p];
   int v = (hd + zd == 0.0) ? 999 : (int) (hd * 10 + .5);
   if (v < -320) v = -320;
   else if (v > 700 && v != 999) v = 700;
   return (v == 999 ? 701 : v) + 320;
}

// exclusive prefix sum over the workgroup's 1024 threads; *total = the sum
__device__ static int blk_scan(int v, int *sh, int *total)
{
   const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
   int x = v;
   for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
   }
   if (lane == 63) sh[w] = x;
   __syncthreads();
   if (w == 0) {
      int t = lane < RX_T / 64 ? sh[lane] : 0;
      for (int o = 1; o < RX_T / 64; o <<= 1) {
         const int y = __shfl_up(t, o);
         if (lane >= o) t += y;
      }
      if (lane < RX_T / 64) sh[lane] = t;
   }
   __syncthreads();
   const int base = w ? sh[w - 1] : 0;
   *total = sh[RX_T / 64 - 1];
   __syncthreads();
   return base + x - v;
}

// storms s0 + blockIdx.x; scr holds 2N+1 words per storm; frame bytes to
// out + blockIdx.x * cap, their count to nbytes (-1: over cap)
__global__ __launch_bounds__(RX_T) void k_rex(SlBasinDev B, SlRunDev R, const real *hbs, const double *hmax,
                                             const double *zbd, int s0, uint8_t *out, long cap, int *nbytes,
                                             uint32_t *scr)
{
   __shared__ int sh[RX_T / 64];
   const int s = s0 + blockIdx.x, t = threadIdx.x;
   const int J1 = B.jmxb - 1, N = (B.imxb - 1) * J1;
   const size_t so = (size_t) s * R.fstride;
   const real *hx = R.hmx + so, *hb = hbs ? hbs + so : nullptr;
   const double *h2 = hmax ? hmax + so : nullptr;
   uint32_t *starts = scr + (size_t) blockIdx.x * (2 * (size_t) N + 1), *rec = starts + N + 1;
   uint8_t *o = out + (size_t) blockIdx.x * cap;
   int nrun = 0, tot;
   for (int c = 0; c < N; c += RX_T) {
      const int k = c + t;
      int f = 0;
      if (k < N) f = k == 0 || rex_code(B, hb, hx, h2, zbd, k, J1) != rex_code(B, hb, hx, h2, zbd, k - 1, J1);
      const int pos = blk_scan(f, sh, &tot);
      if (f) starts[nrun + pos] = (uint32_t) k;
      nrun += tot;
   }
   if (t == 0) starts[nrun] = (uint32_t) N;
   __syncthreads();
   int nrec = 0;
   for (int c = 0; c < nrun; c += RX_T) {
      const int r = c + t;
      int len = 0, n = 0;
      if (r < nrun) {
         len = (int) (starts[r + 1] - starts[r]);
         n = (len + 255) >> 8;
      }
      const int pos = blk_scan(n, sh, &tot);
      if (r < nrun) {
         const uint32_t code = (uint32_t) rex_code(B, hb, hx, h2, zbd, (int) starts[r], J1);
         for (int q = 0; q < n; q++) {
            const int l = len - 256 * q < 256 ? len - 256 * q : 256;
            rec[nrec + pos + q] = ((uint32_t) (l - 1) << 10) | code;
         }
      }
      nrec += tot;
   }
   __syncthreads();
   const long nbits = 18L * nrec, nb = (nbits + 7) / 8;
   if (nb > cap) {
      if (t == 0) nbytes[blockIdx.x] = -1;
      return;
   }
   for (long b = t; b < nb; b += RX_T) {
      unsigned v = 0;
      for (int m = 0; m < 8; m++) {
         const long q = 8 * b + m;
         unsigned bit = 0;
         if (q < nbits) {
            const long r = q / 18;
            bit = (rec[r] >> (17 - (int) (q - 18 * r))) & 1u;
         }
         v = (v << 1) | bit;
      }
      o[b] = (uint8_t) v;
   }
   if (t == 0) nbytes[blockIdx.x] = (int) nb;
}

// ------------------------------------------------------ windowed maxima
// P-Surge windowed maxima of every storm (slosh2.c:1227-1316), in double as
// the C driver holds them, on the squares (I <= IMXB-1, J <= JMXB-1):
// MaxStatInitGrid sets -ZB, MaxStatUpdateGrid takes the maximum with HB after
// each step.
__global__ void k_maxstat(SlBasinDev B, SlRunDev R, const real *hb, const double *zbd, double *hmax,
                          double *hmax2, int init)
{
   const int s = blockIdx.y;
   const int i = (blockIdx.x % ((B.imxb + 63) / 64)) * 64 + threadIdx.x + 1;
   const int j = blockIdx.x / ((B.imxb + 63) / 64) + 1;
   if (s >= R.nstorm || i > B.imxb - 1 || j > B.jmxb - 1) return;
   const uint32_t p = IDX(B, i, j);
   const size_t q = (size_t) s * R.fstride + p;
   if (init) {
      if (hmax) hmax[q] = -1 * zbd[p];
      if (hmax2) hmax2[q] = -1 * zbd[p];
      return;
   }
   // both windows take the same HB (MaxStatUpdateGrid with f_max2,
   // slosh2.c:1295-1306)
   const double h = (double) hb[q];
   if (hmax && hmax[q] < h) hmax[q] = h;
   if (hmax2 && hmax2[q] < h) hmax2[q] = h;
}

// frame k (nb[k] bytes at src + k*cap) to dst + off[k]: the frames of a chunk
// back to back, for one copy to the host
__global__ void k_rex_pack(const uint8_t *src, long cap, const int *nb, const long *off, uint8_t *dst)
{
   const int k = blockIdx.x;
   const uint8_t *s = src + (size_t) k * cap;
   uint8_t *d = dst + off[k];
   for (long b = threadIdx.x; b < nb[k]; b += blockDim.x) d[b] = s[b];
}

// ------------------------------------------------------------------ gridded tide
// The tide of every square of clock group g at the group's step n: the sum of
// CalcTideGrid2 (slosh2.c:875-885, 946-956), in double and in its order:
// -NAVD88_MTL, then for each constituent amplitude x (CA cos - SA sin).
__global__ void __launch_bounds__(256) k_tide_field(SlTideDev T, int n)
{
   const int c = blockIdx.x * 256 + threadIdx.x, g = blockIdx.y;
   if (c >= T.ncell || n > T.gmax[g]) return;
   const int call = T.call[T.goff[g] / (2 * SL_NTIDE) + n];
   if (call < 0) return;
   const double *ca = T.casa + T.goff[g] + (size_t) n * 2 * SL_NTIDE, *sa = ca + SL_NTIDE;
   const size_t yo = (size_t) T.gyear[g] * SL_NTIDE * T.ncell + c;
   const double *a = T.amp + yo, *cv = T.cv + yo, *sv = T.sv + yo;
   double t = -T.mtl[c];
   for (int k = 0; k < SL_NTIDE; k++) {
      const size_t o = (size_t) k * T.ncell;
      t = t + a[o] * (ca[k] * cv[o] - sa[k] * sv[o]);
   }
   T.field[((size_t) g * 2 + (call & 1)) * T.ncell + c] = t;
}

// STATICHT (staticht.f:14-41) at square (i,j) as the tide coupling calls it
// after step n: from the storm values and radial table of the run's next step
// (after its last step, of that step)
__device__ real static_height(const SlBasinDev &B, const SlRunDev &R, int s, int n, int i, int j)
{
   const int k0 = min(n, R.nsteps[s] - 1);
   const real *P = par(R, s, k0 + 1), *Tb = tabv(R, s, P);
   const real xr = B.elpcl[i - 1] * B.cosl[j - 1], yr = B.elpdl[i - 1] * B.sinl[j - 1];
   const real x = xr - P[P_C1] - P[P_AX], y = yr - P[P_C2] - P[P_AY];
   const real rsq = x * x + y * y, r1 = sqrt(rsq) / real(5280.) + real(1.);
   int k = (int) r1;
   const real r2 = (real) k, dr = r1 - r2;
   if (k > 790) k = 790;
   const real *Rk = Tb + (uint32_t) k * 8, *Rk1 = Rk + 8;
   return Rk[R_DELP] + dr * (Rk1[R_DELP] - Rk[R_DELP]) + P[P_SEA];
}

// CalcTideGrid2's change to HB and TIDESH (slosh2.c:859-1025) for every run
// still running at step n; first: the call before the first step, on the
// initial water.  hb is the buffer that holds each run's HB.  HB goes through
// double as the driver's COMMON copy does, and is rounded to real where the
// device keeps it.  The
// old tide of a square is the group's field of step n-1: the driver's stored
// value and its recomputation from the previous call's angles are that sum.
// A step's threads take the listed squares only (those SetTideUsableFlag or
// the deep rule flag), the first call's every square.
// One thread takes one square (or BDRYHT point) of SL_TSPT runs: the square's
// ZB and flags are read once, and the runs' HB loads are in flight together.
#define SL_TSPT 4
__global__ void __launch_bounds__(256) k_tide_apply(SlBasinDev B, SlRunDev R, SlTideDev T, real *hb, int n,
                                                    int first)
{
   const int nsq = first ? T.ncell : T.nlist, t0 = blockIdx.x * 256 + threadIdx.x, s0 = blockIdx.y * SL_TSPT;
   if (t0 >= nsq + T.nbcpt) return;
   const int ns = min(SL_TSPT, R.nstorm - s0);
   // the field slots of this call and the one before, per run: -1 where its
   // group makes no call after step n
   int cs[SL_TSPT];
#pragma unroll
   for (int t = 0; t < SL_TSPT; t++) {
      const int s = s0 + t;
      cs[t] = -1;
      if (t < ns && (first || R.nsteps[s] >= n)) {
         const int g = T.sgrp[s];
         cs[t] = T.call[T.goff[g] / (2 * SL_NTIDE) + n];
      }
   }
   if (t0 >= nsq) {   // the boundary squares: TIDESH = the tide of this call
      const int b = t0 - nsq, c = T.bccell[b];
      for (int t = 0; t < ns; t++) {
         const int s = s0 + t;
         if (cs[t] >= 0)
            R.tidesh[(size_t) s * T.nbcpt + b] =
               (real) T.field[((size_t) T.sgrp[s] * 2 + (cs[t] & 1)) * T.ncell + c];
      }
      return;
   }
   const int p = first ? t0 : T.list[t0];    // a step: only the flagged squares
   const uint8_t f = first ? 1 : T.cfl[p];   // the first call: every wet square (slosh2.c:859-896)
   if (!(f & 3)) return;                      // not usable: HB keeps its value
   const int i = p % T.ni, j = p / T.ni;
   const size_t q = (size_t) j * B.pitch + i;
   const double zb = T.zb[p];
   double h[SL_TSPT];
   bool on[SL_TSPT];
#pragma unroll
   for (int t = 0; t < SL_TSPT; t++) {
      const int s = s0 + t;
      on[t] = cs[t] >= 0;
      h[t] = 0.;
      if (on[t] && !(f & 2)) h[t] = (double) hb[(size_t) s * R.fstride + q];
   }
#pragma unroll
   for (int t = 0; t < SL_TSPT; t++) {
      if (!on[t]) continue;
      const int s = s0 + t;
      const size_t o = (size_t) s * R.fstride + q;
      const double *fg = T.field + (size_t) T.sgrp[s] * 2 * T.ncell + p;
      const size_t now = (size_t) (cs[t] & 1) * T.ncell, old = (size_t) ((cs[t] - 1) & 1) * T.ncell;
      if (f & 2) {   // V2.1 / V2.2 deep squares (slosh2.c:904-931)
         hb[o] = (real) ((double) static_height(B, R, s, n, i + 1, j + 1) + fg[now]);
      } else if (h[t] + zb != 0.) {   // wet: subtract the old tide (none before the first call), add the new
         const double x = first ? h[t] + fg[now] : h[t] - fg[old] + fg[now];
         hb[o] = (x + zb >= 0.) ? (real) x : (real) (-zb);
      }
   }
}

}   // namespace SL_NS

using namespace SL_NS;

// One time step of every storm still running (each on its own step cur+1).
// smpt / filt say whether SMPT2G and FILTER run this step when the storms are
// in lockstep; dyn != 0 makes each storm's own step flags decide (fused path).
// path 0 = fused (events: 0 before CONTTY, 1 after it, 2 after the replay, 3
// after stage B); path 1 = staged (events around each of the 8 stage kernels).
// Returns which HB buffer holds the step's result.
#ifdef SL_F64
#define SL_ENVS sl_dev_envelopes_f64
#else
#define SL_ENVS sl_dev_envelopes_f32
#endif
extern "C" int SL_ENVS(const SlBasinDev *B, const SlRunDev *R, const double *hx2, const double *zbd, int s0,
                       int ns, short *env, short *mom, int acc, hipStream_t st)
{
   const int n = B->imxb * B->jmxb;
   hipLaunchKernelGGL(k_envelopes, dim3((n + 255) / 256), dim3(256), 0, st, *B, *R, hx2, zbd, s0, ns, env, mom,
                      acc);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifdef SL_F64
#define SL_SLOTINIT sl_dev_slot_init_f64
#else
#define SL_SLOTINIT sl_dev_slot_init_f32
#endif
extern "C" int SL_SLOTINIT(const SlBasinDev *B, const SlRunDev *R, int slot, int tstart, int nsteps, hipStream_t st)
{
   const int n = std::max(std::max(B->npts, B->pitch), std::max(B->nsqrwc, B->nsq));
   hipLaunchKernelGGL(k_slot_init, dim3((n + 255) / 256), dim3(256), 0, st, *B, *R, slot, tstart, nsteps);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

// INTLHT (intlht.f:53-185 without the Okeechobee branch; 161, 172-177) for
// every run, on the device.  ocn [jmxb][imxb] says how the host's reading of
// the basin fills each square (sl_intlht_class): squares in reach of ocean
// winds take the static height of the first hour's storm at the square, at
// least -ZB; the others a datum (the lake's, or the sea's on MSY's open
// squares), at least -ZB; listed squares dry out (-ZB).  Row JMXB repeats row
// 1.  In double (the host's operations in its order), rounded to R into HB
// (both buffers) and HMX (squares only).  bm: ELPCL[imxb], ELPDL[imxb],
// COSL[jmxb], SINL[jmxb], then ZB[jmxb][imxb] (packed, I fastest).
__global__ void __launch_bounds__(256) k_intlht(const double *prog, const long *off, const double *bm,
                                                const uint8_t *ocn, int imxb, int jmxb, int imxb1, int jmxb1,
                                                int pitch, SlRunDev R)
{
   const int s = blockIdx.y, p = blockIdx.x * 256 + threadIdx.x;
   if (p >= imxb * jmxb) return;
   const int i = p % imxb + 1, j = p / imxb + 1;
   const double *g = prog + off[s];
   const double *elpcl = bm, *elpdl = bm + imxb, *cosl = bm + 2 * imxb, *sinl = cosl + jmxb, *zbv = sinl + jmxb;
   const int jj = (j == jmxb) ? 1 : j;
   double hb = 0.;
   if (i <= imxb1 && jj <= jmxb1) {
      const int q = (jj - 1) * imxb + (i - 1);
      const double zb = zbv[q];
      const int cls = ocn[q];
      if (cls & 1) {
         const double *st = g + SQ_HDR, *delp = g + SQ_DELP;
         const double xr = elpcl[i - 1] * cosl[jj - 1], yr = elpdl[i - 1] * sinl[jj - 1];
         const double x = xr - st[SQ_C1] - st[SQ_AX], y = yr - st[SQ_C2] - st[SQ_AY];
         const double rsq = x * x + y * y, r1 = sqrt(rsq) / 5280. + 1.;
         int k = (int) r1;
         const double r2 = k, dr = r1 - r2;
         if (k > 790) k = 790;
         const double h = delp[k] + dr * (delp[k + 1] - delp[k]) + st[SQ_SEA];
         hb = h > -zb ? h : -zb;
      } else {
         const double dtm = (cls & 4) ? g[SQ_HDR + SQ_SEA] : g[3];
         const double t = dtm > -zb ? dtm : -zb;
         hb = -zb > t ? -zb : t;
      }
      if (cls & 2) hb = -zb;
   }
   const real v = (real) hb;
   const size_t o = (size_t) s * R.fstride + (size_t) (j - 1) * pitch + (i - 1);
   R.hb0[o] = v;
   R.hb1[o] = v;
   R.hmx[o] = (i == imxb || j == jmxb) ? real(0.) : v;
}

#ifdef SL_F64
#define SL_INTLHT sl_dev_intlht_f64
#else
#define SL_INTLHT sl_dev_intlht_f32
#endif
extern "C" int SL_INTLHT(const double *prog, const long *off, const double *bm, const unsigned char *ocn, int imxb,
                         int jmxb, int imxb1, int jmxb1, int pitch, const SlRunDev *R, hipStream_t st)
{
   if (R->nstorm < 1 || R->nstorm > 65535 || imxb < 2 || jmxb < 2 || pitch < imxb ||
       R->fstride < (size_t) pitch * jmxb)
      return -1;
   hipLaunchKernelGGL(k_intlht, dim3((imxb * jmxb + 255) / 256, R->nstorm), dim3(256), 0, st, prog, off, bm, ocn,
                      imxb, jmxb, imxb1, jmxb1, pitch, *R);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifdef SL_F64
#define SL_SCHED sl_dev_sched_f64
#else
#define SL_SCHED sl_dev_sched_f32
#endif
// per-step storm values [nstorm][maxsteps][SL_NPAR] of every run from the
// storm programs (prog + off[s], each run's nsteps <= maxsteps)
extern "C" int SL_SCHED(const double *prog, const long *off, int nstorm, real *params, int maxsteps, int ismpt2g,
                        int incsm, int filt_always, hipStream_t st)
{
   if (nstorm < 1 || incsm < 1) return -1;
   hipLaunchKernelGGL(k_sched, dim3((nstorm + 63) / 64), dim3(64), 0, st, prog, off, nstorm, params, maxsteps,
                      ismpt2g, incsm, filt_always);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifdef SL_F64
#define SL_REX sl_dev_rex_f64
#else
#define SL_REX sl_dev_rex_f32
#endif
// rex frames of storms s0 .. s0+ns-1 from HHB (hbs: the HB buffer before
// FILTER) or from HMX (hbs == nullptr); scr: ns * (2N+1) words
extern "C" int SL_REX(const SlBasinDev *B, const SlRunDev *R, const real *hbs, const double *hmax,
                      const double *zbd, int s0, int ns, unsigned char *out, long cap, int *nbytes, unsigned *scr,
                      hipStream_t st)
{
   hipLaunchKernelGGL(k_rex, dim3(ns), dim3(RX_T), 0, st, *B, *R, hbs, hmax, zbd, s0, out, cap, nbytes, scr);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

// rex values at listed squares, without the frames: storm s0 + blockIdx.y,
// square ks[p] in frame order ((I-1)(JMXB-1) + J-1); the value a rex frame
// holds there (tenths of a foot, -320..700, 999 dry), as rexout reads it back
// (sorc/rexout.cd/rexout.c:707-740)
__global__ __launch_bounds__(256) void k_rex_points(SlBasinDev B, SlRunDev R, const real *hbs, const double *hmax,
                                                    const double *zbd, int s0, const int *ks, int npts,
                                                    int16_t *out)
{
   const int p = blockIdx.x * 256 + threadIdx.x;
   if (p >= npts) return;
   const size_t so = (size_t) (s0 + blockIdx.y) * R.fstride;
   const int c = rex_code(B, hbs ? hbs + so : nullptr, R.hmx + so, hmax ? hmax + so : nullptr, zbd, ks[p],
                          B.jmxb - 1) - 320;
   out[(size_t) blockIdx.y * npts + p] = (int16_t) (c == 701 ? 999 : c);
}

#ifdef SL_F64
#define SL_REXPTS sl_dev_rex_points_f64
#else
#define SL_REXPTS sl_dev_rex_points_f32
#endif
extern "C" int SL_REXPTS(const SlBasinDev *B, const SlRunDev *R, const real *hbs, const double *hmax,
                         const double *zbd, int s0, int ns, const int *ks, int npts, int16_t *out, hipStream_t st)
{
   hipLaunchKernelGGL(k_rex_points, dim3((npts + 255) / 256, ns), dim3(256), 0, st, *B, *R, hbs, hmax, zbd, s0, ks,
                      npts, out);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifdef SL_F64
#define SL_MAXSTAT sl_dev_maxstat_f64
#else
#define SL_MAXSTAT sl_dev_maxstat_f32
#endif
// windowed maxima of the storms of R (either buffer may be null): init (-ZB)
// or the maximum with hb
extern "C" int SL_MAXSTAT(const SlBasinDev *B, const SlRunDev *R, const real *hb, const double *zbd, double *hmax,
                          double *hmax2, int init, hipStream_t st)
{
   const int nx = (B->imxb + 63) / 64;
   hipLaunchKernelGGL(k_maxstat, dim3(nx * (B->jmxb - 1), R->nstorm), dim3(64), 0, st, *B, *R, hb, zbd, hmax,
                      hmax2, init);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifdef SL_F64
#define SL_TIDE sl_dev_tide_f64
#else
#define SL_TIDE sl_dev_tide_f32
#endif
// the tide after step n (first: before the first step) of every run of R
extern "C" int SL_TIDE(const SlBasinDev *B, const SlRunDev *R, const SlTideDev *T, real *hb, int n, int first,
                       hipStream_t st)
{
   if (T->ngrp < 1 || T->ngrp > 65535 || R->nstorm < 1 || R->nstorm > 65535 || T->ncell < 1 || !R->tidesh ||
       T->ni < 1 || T->ncell % T->ni || T->ni > B->imxb - 1 || T->ncell / T->ni > B->jmxb - 1)
      return -1;
   hipLaunchKernelGGL(k_tide_field, dim3((T->ncell + 255) / 256, T->ngrp), dim3(256), 0, st, *T, n);
   const int nsq = first ? T->ncell : T->nlist;
   if (nsq + T->nbcpt > 0)
      hipLaunchKernelGGL(k_tide_apply, dim3((nsq + T->nbcpt + 255) / 256, (R->nstorm + SL_TSPT - 1) / SL_TSPT),
                         dim3(256), 0, st, *B, *R, *T, hb, n, first);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}

#ifndef SL_F64
// frames of one chunk back to back (both precisions use this one)
extern "C" int sl_dev_rex_pack(const unsigned char *src, long cap, const int *nbytes, const long *off, int ns,
                               unsigned char *dst, hipStream_t st)
{
   hipLaunchKernelGGL(k_rex_pack, dim3(ns), dim3(256), 0, st, src, cap, nbytes, off, dst);
   hipError_t e = hipGetLastError();
   return e == hipSuccess ? 0 : -(int) e;
}
#endif

extern "C" int SL_LAUNCH(const SlBasinDev *B, const SlRunDev *R, int smpt, int filt, int dyn, int parity,
                         int path, int bth, hipStream_t st, hipEvent_t *ev)
{
   real *h0 = parity ? R->hb1 : R->hb0, *h1 = parity ? R->hb0 : R->hb1;
   const dim3 gadv((R->nstorm + 255) / 256), badv(256);
   // basins with 1-D flow sides or channels: their passes sit between SMPT2G
   // and HMXSV, so the fused path runs SMPT2G/FLT2G/BDRYHT into hb2, the side
   // passes there, and the step kernel from HMXSV on
   const bool sides = B->nsqrw > 0 || B->f1d_on || B->npss;
   StepArgs a;
#define MARK(k) \
   if (ev) (void) hipEventRecord(ev[k], st)
   if (path == 0) {
      const int ntx = (B->imxb + TX - 1) / TX;
      int nty = (B->jmxb + ATH - 1) / ATH;
      a.hc = h0;
      a.hn = h1;
      MARK(0);
      hipLaunchKernelGGL(k_conttyA, dim3(ntx * nty * R->nstorm), dim3(TX, TYA), 0, st, *B, *R, a, ntx, nty);
      MARK(1);
      hipLaunchKernelGGL(k_replay, dim3(R->nstorm), dim3(64), 0, st, *B, *R, a, 1);
      MARK(2);
      a.hc = h1;
      a.hn = h0;
      if (sides) {
         // HB after CONTTY (h1) -> hb2: SMPT2G's classes and side pass, FLT2G,
         // BDRYHT and the row copy; then CHANNL and FLW1DM_C/M on hb2; the
         // step kernel takes hb2 from HMXSV on into h0
         const dim3 blk(TX, TY), grd((B->imxb + TX - 1) / TX, (B->jmxb + TY - 1) / TY, R->nstorm);
         StepArgs c;
         c.hc = h1;
         c.hn = R->hb2;
         if (smpt || dyn) {
            hipLaunchKernelGGL(k_pflag, grd, blk, 0, st, *B, *R, c);
            hipLaunchKernelGGL(k_hclass, grd, blk, 0, st, *B, *R, c);
            if (B->nsqrw > 0)
               hipLaunchKernelGGL(k_f1d_mark, dim3((R->nstorm + 63) / 64), dim3(64), 0, st, *B, *R, c);
         }
         hipLaunchKernelGGL(k_smooth, grd, blk, 0, st, *B, *R, c);
         if (B->f1d_on || B->npss)
            hipLaunchKernelGGL(k_sides, dim3((R->nstorm + 63) / 64), dim3(64), 0, st, *B, *R, c);
         a.hc = R->hb2;
      }
      const int th = (bth == 8 || bth == 32) ? bth : 16;
      const int nbx = (B->imxb + 512 / th - 1) / (512 / th);
      nty = (B->jmxb + th - 1) / th;
      const dim3 g(nbx * nty * R->nstorm);
      if (th == 8) launch_B<8>(smpt, filt, dyn, sides, g, st, *B, *R, a, nbx, nty);
      else if (th == 32) launch_B<32>(smpt, filt, dyn, sides, g, st, *B, *R, a, nbx, nty);
      else launch_B<16>(smpt, filt, dyn, sides, g, st, *B, *R, a, nbx, nty);
      if (B->island) {
         // HB after HMXSV is in h0: FLTER2 into h1 and back (with the row
         // copy) for the storms that filter, then MOMNTM and the periodic row
         const dim3 blk(TX, TY), grd((B->imxb + TX - 1) / TX, (B->jmxb + TY - 1) / TY, R->nstorm);
         StepArgs f;
         f.hc = h0;
         f.hn = h1;
         if (dyn || filt) {
            hipLaunchKernelGGL(k_filter2, grd, blk, 0, st, *B, *R, f, dyn, filt);
            hipLaunchKernelGGL(k_island_hb, grd, blk, 0, st, *B, *R, (const real *) h1, h0, dyn, filt, 0);
         }
         hipLaunchKernelGGL(k_momntm, grd, blk, 0, st, *B, *R, f);
         hipLaunchKernelGGL(k_island_uv, dim3((B->imxb + 255) / 256, R->nstorm), dim3(256), 0, st, *B, *R);
      }
      MARK(3);
      hipLaunchKernelGGL(k_advance, gadv, badv, 0, st, *R);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return -(int) e;
      return parity;
   }
   dim3 blk(TX, TY), grd((B->imxb + TX - 1) / TX, (B->jmxb + TY - 1) / TY, R->nstorm);
   // CONTTY: h0 (HSUB) -> h1
   a.hc = h0;
   a.hn = h1;
   MARK(0);
   hipLaunchKernelGGL(k_contty, grd, blk, 0, st, *B, *R, a);
   MARK(1);
   hipLaunchKernelGGL(k_replay, dim3(R->nstorm), dim3(64), 0, st, *B, *R, a, 0);
   MARK(2);
   // SMPT2G classification on h1
   a.hc = h1;
   a.hn = h0;
   if (smpt) {
      hipLaunchKernelGGL(k_pflag, grd, blk, 0, st, *B, *R, a);
      MARK(3);
      hipLaunchKernelGGL(k_hclass, grd, blk, 0, st, *B, *R, a);
      MARK(4);
   } else {
      MARK(3);
      MARK(4);
   }
   // SMPT2G's pass over the 1-D flow sides
   if (smpt && B->nsqrw > 0) hipLaunchKernelGGL(k_f1d_mark, dim3((R->nstorm + 63) / 64), dim3(64), 0, st, *B, *R, a);
   // FLT2G + BDRYHT + row copy: h1 -> h0
   hipLaunchKernelGGL(k_smooth, grd, blk, 0, st, *B, *R, a);
   // CHANNL, FLW1DM_C, FLW1DM_M on h0
   if (B->f1d_on || B->npss) hipLaunchKernelGGL(k_sides, dim3((R->nstorm + 63) / 64), dim3(64), 0, st, *B, *R, a);
   MARK(5);
   // HMXSV on h0
   a.hc = h0;
   hipLaunchKernelGGL(k_hmxsv, grd, blk, 0, st, *B, *R, a);
   MARK(6);
   real *fin = h0;
   if (filt) {
      a.hc = h0;
      a.hn = h1;
      if (B->island) {   // FLTER2, then HB(I,JMXB) = HB(I,1)
         hipLaunchKernelGGL(k_filter2, grd, blk, 0, st, *B, *R, a, 0, 1);
         hipLaunchKernelGGL(k_island_hb, grd, blk, 0, st, *B, *R, (const real *) h1, h1, 0, 1, 1);
      } else {
         hipLaunchKernelGGL(k_filter, grd, blk, 0, st, *B, *R, a);
      }
      fin = h1;
   }
   MARK(7);
   a.hc = fin;
   hipLaunchKernelGGL(k_momntm, grd, blk, 0, st, *B, *R, a);
   if (B->island)
      hipLaunchKernelGGL(k_island_uv, dim3((B->imxb + 255) / 256, R->nstorm), dim3(256), 0, st, *B, *R);
   MARK(8);
   hipLaunchKernelGGL(k_advance, gadv, badv, 0, st, *R);
#undef MARK
   hipError_t e = hipGetLastError();
   if (e != hipSuccess) return -(int) e;
   return (fin == R->hb0) ? 0 : 1;   // which buffer holds HB after the step
}
