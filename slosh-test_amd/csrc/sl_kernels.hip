// sl_kernels.hip -- the SLOSH time step on CDNA4 (gfx950), FP32.
//
// One launch of each kernel advances every storm of a batch through one stage
// of the step: blockIdx.z (or the thread index of the per-storm kernels) is the
// storm, so a single storm is simply a batch of one.  Each stage restates a
// reference routine (file:line under sorc/slosh.fd):
//
//   k_contty     CONTTY  contty.f:85-221 (Jacobi pass; drying squares listed)
//   k_replay     CONTTY  contty.f:146-213 serial back-step cascade, replayed
//                exactly from the first drying square, visiting only squares
//                whose inputs the cascade changed
//   k_pflag      SMPT2G  smpt2g.f:23-45  (momentum points next to U=0 points)
//   k_hclass     SMPT2G  smpt2g.f:51-120 (height points: copy / 888 / 999)
//   k_smooth     FLT2G   flt2g.f:38-132 both passes + BDRYHT bdryht.f:72-86
//                + the HB(I,JMXB)=HB(I,1) copy contty.f:218-219
//   k_hmxsv      HMXSV   hmxsv.f:44-136
//   k_filter     FILTER  filter.f:121-211, XFILTER xfilter.f:176-713
//   k_momntm     MOMNTM  momntm.f:78-470 with FRCPNT frcpnt.f:63-200 and
//                MNTMBD mntmbd.f:58-117
//
// Bit-exactness with the FP32 restatement of the reference depends on the
// operation order being the one written here: the file is compiled with
// -ffp-contract=off and correctly rounded fp32 division and square root.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "sl_dev.h"

#define TX 64
#define TY 4
#define IDX(B, i, j) ((size_t) ((j) - 1) * (size_t) (B).pitch + (size_t) ((i) - 1))
#define TABI(B, t, iik, id) (B).tab[(size_t) (t) * 147000 + (size_t) ((iik) - 1) * 600 + ((id) - 1)]
enum { T_AR2, T_AI2, T_BR2, T_BI2, T_CR2, T_CI2 };
enum { T_S, T_C, T_P, T_DELP, T_SW, T_CW };

// BLOCK DATA shift tables (flw1dm_block_c.f:37-38), 1-based
__constant__ int cIP[5] = {0, 0, 1, 1, 0}, cJP[5] = {0, 0, 0, 1, 1};
__constant__ int cIH[5] = {0, -1, 0, -1, 0}, cJH[5] = {0, -1, -1, 0, 0};
__constant__ int cIIH[5] = {0, 0, 1, 0, -1}, cJJH[5] = {0, -1, 0, 1, 0};

struct StepArgs {
   const float *hc;     // HB entering the stage (per storm base = hc + s*fstride)
   float *hn;           // HB leaving the stage
   int step;            // 1-based step number
};

__device__ __forceinline__ const float *par(const SlRunDev &R, int s, int step)
{
   return R.params + ((size_t) s * R.maxsteps + (step - 1)) * SL_NPAR;
}

__device__ __forceinline__ const float *tabv(const SlRunDev &R, int s, const float *P)
{
   int v = __float_as_int(P[P_TVER]);
   return R.tables + (size_t) (R.tbase[s] + v) * 6 * SL_NTABV;
}

__device__ __forceinline__ bool zr(const SlBasinDev &B, int i, int j)
{
   return i >= 2 && i <= B.imxb - 1 && j >= 2 && j <= B.jmxb - 1;
}

// ---------------------------------------------------------------------- CONTTY
// Jacobi pass over all squares.  HSUB is the entering HB (hc).  UB values with
// |UB| <= 1e-10 inside 2..IMXB1 x 2..JMXB1 count as 0 (and are flushed by their
// owner thread), as contty.f:90-96 does before the sweep.
__global__ void __launch_bounds__(TX * TY) k_contty(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb || j > B.jmxb) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const float *hc = A.hc + so;
   float *hn = A.hn + so, *ub = R.ub + so, *vb = R.vb + so;
   if (j == B.jmxb) {
      if (i <= B.imxb - 1) R.hsubtop[(size_t) s * B.pitch + (i - 1)] = hc[p];
      hn[p] = hc[p];
      return;
   }
   if (i == B.imxb) {
      hn[p] = hc[p];
      return;
   }
   const uint32_t fl = B.flags[p];
   if (fl & F_ZR) {
      if (fabsf(ub[p]) <= 1.E-10f) {
         ub[p] = 0.f;
         vb[p] = 0.f;
      }
   }
   float out = hc[p];
   if (fl & F_ACT) {
      const size_t p2 = p + 1, p3 = p + B.pitch + 1, p4 = p + B.pitch;
      float u1 = ub[p], u2 = ub[p2], u3 = ub[p3], u4 = ub[p4];
      float v1 = vb[p], v2 = vb[p2], v3 = vb[p3], v4 = vb[p4];
      if (zr(B, i, j) && fabsf(u1) <= 1.E-10f) { u1 = 0.f; v1 = 0.f; }
      if (zr(B, i + 1, j) && fabsf(u2) <= 1.E-10f) { u2 = 0.f; v2 = 0.f; }
      if (zr(B, i + 1, j + 1) && fabsf(u3) <= 1.E-10f) { u3 = 0.f; v3 = 0.f; }
      if (zr(B, i, j + 1) && fabsf(u4) <= 1.E-10f) { u4 = 0.f; v4 = 0.f; }
      const float dux = -u1 + u2 + u3 - u4;
      const float zb = B.zb[p];
      if (!(zb <= 30.f && dux == 0.f)) {
         const float *P = par(R, s, A.step);
         const float hpst = hc[p];
         const float z = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
         const float fct = P[P_X1B5] / z;
         const float dvy = -v1 - v2 + v3 + v4;
         const float hbj = hpst - fct * (dux + dvy);
         const float chng = hbj + zb;
         if (chng < 0.f) {
            if (-chng < 1E-3f) out = -zb;
            else if (hc[p] == -zb) out = hc[p];
            else {
               out = -zb;   // final value comes from the replay
               int k = atomicAdd(&R.evcnt[s], 1);
               if (k < SL_EVCAP) R.evkey[(size_t) s * SL_EVCAP + k] = ((uint32_t) j << 16) | (uint32_t) i;
            }
         } else {
            out = hbj;
         }
      }
   }
   hn[p] = out;
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
struct Replay {
   const SlBasinDev *B;
   const float *hc;
   float *hn, *ub, *vb;
   uint32_t *key;
   int n, full;
   uint32_t frontier;
   float x1b5;
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

__device__ static void go_full(Replay &X)
{
   const SlBasinDev &B = *X.B;
   X.full = 1;
   for (int j = 1; j <= B.jmxb - 1; j++)
      for (int i = 1; i <= B.imxb - 1; i++) {
         uint32_t k = ((uint32_t) j << 16) | (uint32_t) i;
         size_t p = IDX(B, i, j);
         if (k > X.frontier && (B.flags[p] & F_ACT)) X.hn[p] = X.hc[p];
      }
}

__device__ static void dirty(Replay &X, int i, int j)
{
   const SlBasinDev &B = *X.B;
   if (i < 1 || j < 1 || i > B.imxb - 1 || j > B.jmxb - 1) return;
   const size_t p = IDX(B, i, j);
   if (!(B.flags[p] & F_ACT)) return;
   const uint32_t k = ((uint32_t) j << 16) | (uint32_t) i;
   if (X.full) return;
   int pos = lbound(X.key, X.n, k);
   if (pos < X.n && X.key[pos] == k) return;
   if (k > X.frontier) X.hn[p] = X.hc[p];   // not visited yet: HB is still HSUB
   if (X.n >= SL_EVCAP) {
      go_full(X);
      return;
   }
   for (int q = X.n; q > pos; q--) X.key[q] = X.key[q - 1];
   X.key[pos] = k;
   X.n++;
}

// One visit of square (i,j) (the DO 320 body).  Returns 1 if it dried with a
// transport rescale (the caller then back-steps), else 0.
__device__ static int visit(Replay &X, int i, int j, int &leap, int ipass)
{
   const SlBasinDev &B = *X.B;
   const size_t p = IDX(B, i, j);
   if (!(B.flags[p] & F_ACT)) return 0;
   const uint32_t k = ((uint32_t) j << 16) | (uint32_t) i;
   if (k > X.frontier) X.frontier = k;
   const float zb = B.zb[p];
   if (leap == 2 && i == ipass) {
      X.hn[p] = -zb;
      leap = 1;
      return 0;
   }
   float *ub = X.ub, *vb = X.vb;
   const size_t p2 = p + 1, p3 = p + B.pitch + 1, p4 = p + B.pitch;
   const float dux = -ub[p] + ub[p2] + ub[p3] - ub[p4];
   if (zb <= 30.f && dux == 0.f) return 0;
   const float hpst = X.hc[p];
   const float z = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
   const float fct = X.x1b5 / z;
   const float dvy = -vb[p] - vb[p2] + vb[p3] + vb[p4];
   const float hbj = hpst - fct * (dux + dvy);
   const float chng = hbj + zb;
   if (chng < 0.f) {
      if (-chng < 1E-3f) {
         X.hn[p] = -zb;
         return 0;
      }
   } else {
      X.hn[p] = hbj;
      return 0;
   }
   if (X.hc[p] == -zb) return 0;
   float denm = hbj - hpst;
   denm = copysignf(fmaxf(1.E-5f, fabsf(denm)), denm);
   const float anum = -zb - hpst;
   float rfct = anum / denm;
   if (fabsf(rfct) < 1.E-5f) rfct = 0.f;
   const size_t c[4] = {p, p2, p3, p4};
   for (int q = 0; q < 4; q++) {
      ub[c[q]] = rfct * ub[c[q]];
      vb[c[q]] = rfct * vb[c[q]];
   }
   for (int jj = j - 1; jj <= j + 1; jj++)
      for (int ii = i - 1; ii <= i + 1; ii++) dirty(X, ii, jj);
   return 1;
}

__global__ void k_replay(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int s = blockIdx.x * blockDim.x + threadIdx.x;
   if (s >= R.nstorm) return;
   int n = R.evcnt[s];
   if (n == 0) return;
   const size_t so = (size_t) s * R.fstride;
   Replay X;
   X.B = &B;
   X.hc = A.hc + so;
   X.hn = A.hn + so;
   X.ub = R.ub + so;
   X.vb = R.vb + so;
   X.key = R.evkey + (size_t) s * SL_EVCAP;
   X.full = 0;
   X.x1b5 = par(R, s, A.step)[P_X1B5];
   if (n > SL_EVCAP) {
      n = 0;
      X.full = 1;
   } else {   // insertion sort of the drying squares, row-major
      for (int a = 1; a < n; a++) {
         uint32_t v = X.key[a];
         int b = a - 1;
         while (b >= 0 && X.key[b] > v) {
            X.key[b + 1] = X.key[b];
            b--;
         }
         X.key[b + 1] = v;
      }
   }
   X.n = n;
   const int jmxb1 = B.jmxb - 1;
   int j, ib, ifn, jump = 4, leap = 1, iset = 0, ipass = 0;
   if (X.full) {
      j = 1;
      X.frontier = 0;
      go_full(X);
   } else {
      j = (int) (X.key[0] >> 16);
      X.frontier = (uint32_t) j << 16;
   }
   ib = B.is[j - 1];
   ifn = B.ie[j - 1];
   // The reference sweep ends after a bounded number of visits; the budget
   // only guards the device against a malformed bound table.
   long budget = 64L * B.imxb * B.jmxb + 4096;
L130:
   if (--budget < 0) {
      atomicOr(&R.stat[s * 8 + 7], 1);
      goto L350;
   }
   if (ifn >= ib) {
      int pos = X.full ? 0 : lbound(X.key, X.n, ((uint32_t) j << 16) | (uint32_t) ib);
      int i = ib;
      for (;;) {
         if (X.full) {
            if (i > ifn) break;
         } else {
            if (pos >= X.n || X.key[pos] > (((uint32_t) j << 16) | (uint32_t) ifn)) break;
            i = (int) (X.key[pos] & 0xffffu);
            pos++;
         }
         if (visit(X, i, j, leap, ipass)) {
            j = j - 1;
            jump = 1;
            iset = i;
            leap = 1;
            if (j == 0) goto L330;
            if (!(B.flags[IDX(B, i, j)] & F_ACT)) goto L330;
            if (B.is[j - 1] > i) goto L330;
            ib = (i - 1 > B.is[j - 1]) ? i - 1 : B.is[j - 1];
            ifn = (i + 1 < B.ie[j - 1]) ? i + 1 : B.ie[j - 1];
            goto L130;
         }
         if (X.full) i++;
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
         int pos = lbound(X.key, X.n, (uint32_t) j << 16);
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
   R.evcnt[s] = 0;
   atomicAdd(&R.stat[s * 8 + 4], 1);
}

// ------------------------------------------------------------------ SMPT2G
// Sum of a per-lane count over the 64 lanes of a wave (every lane must call it);
// returns the total in every lane.
__device__ __forceinline__ int wave_sum(int v)
{
   for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
   return v;
}

// The counters feed the checks against the reference's list limits (the
// 30000 / 8000 entries of smpt2g.f:31-38, 57-62, 78-96); one atomic per wave.
__global__ void __launch_bounds__(TX * TY) k_pflag(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   const bool in = (i <= B.imxb && j <= B.jmxb);
   const size_t so = (size_t) s * R.fstride, p = in ? IDX(B, i, j) : 0;
   const float *ub = R.ub + so, *hn = A.hc + so;
   uint8_t f = 0;
   int np = 0, nd = 0;
   if (in && i >= 2 && i <= B.imxb - 2 && j >= 2 && j <= B.jmxb - 2) {
      if (ub[p] != 0.f && B.zbm[p] > -30.f) {
         if (ub[p - 1] == 0.f || ub[p - B.pitch] == 0.f || ub[p + B.pitch] == 0.f || ub[p + 1] == 0.f) {
            int ndry = 0;
            f = 1;
            for (int k = 1; k <= 4; k++) {
               size_t q = IDX(B, i + cIH[k], j + cJH[k]);
               if (hn[q] + B.zb[q] == 0.f) ndry++;
            }
            if (ndry) f |= 2;
            np = 1;
            nd = ndry;
         }
      }
   }
   np = wave_sum(np);
   nd = wave_sum(nd);
   if (threadIdx.x == 0) {
      if (np) atomicAdd(&R.stat[s * 8 + 0], np);
      if (nd) atomicAdd(&R.stat[s * 8 + 1], nd);
   }
   if (in) R.pflag[so + p] = f;
}

__global__ void __launch_bounds__(TX * TY) k_hclass(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   const bool in = (i <= B.imxb && j <= B.jmxb);
   const size_t so = (size_t) s * R.fstride, p = in ? IDX(B, i, j) : 0;
   uint8_t c = 0;
   if (in && i <= B.imxb - 1 && j <= B.jmxb - 1) {
      const float *hn = A.hc + so;
      if (hn[p] + B.zb[p] != 0.f) {
         const uint8_t *pf = R.pflag + so;
         // momentum points having (i,j) as a corner: (i+1,j+1) (i,j+1) (i+1,j) (i,j)
         uint8_t a = pf[p + B.pitch + 1] | pf[p + B.pitch] | pf[p + 1] | pf[p];
         c = (a & 2) ? 2 : (a & 1) ? 1 : 0;
      }
   }
   const int n1 = wave_sum(c == 1), n2 = wave_sum(c == 2);
   if (threadIdx.x == 0) {
      if (n1) atomicAdd(&R.stat[s * 8 + 2], n1);
      if (n2) atomicAdd(&R.stat[s * 8 + 3], n2);
   }
   if (in) R.hclass[so + p] = c;
}

// FLT2G (both passes) + BDRYHT + the CONTTY row copy.  hc = HB after CONTTY,
// hn = HB after TMSTPC1.
__device__ __forceinline__ float gamf_of(const SlBasinDev &B, int k, float gam0, int ii, int jj)
{
   if (B.ebsn_ell) {
      const float gamfk = B.elpdl2[ii - 1] + (B.elpcl2[ii - 1] - B.elpdl2[ii - 1]) * B.sinl2[jj - 1];
      const float z = gamfk / gam0;
      return 2.f * z / (1.f + z) - 1.f;
   }
   return B.gamfp[k - 1];
}

__global__ void __launch_bounds__(TX * TY) k_smooth(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb || j > B.jmxb) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const float *hs = A.hc + so;
   float *hb = A.hn + so;
   const float *P = par(R, s, A.step);
   const int flags = __float_as_int(P[P_FLAGS]);
   if (j == B.jmxb) {
      hb[p] = (i <= B.imxb - 1) ? hs[IDX(B, i, 1)] : hs[p];
      return;
   }
   float out = hs[p];
   if (i <= B.imxb - 1) {
      if (flags & SF_SMPT) {
         const uint8_t *cl = R.hclass + so;
         const uint8_t c = cl[p];
         if (c) {
            const float hct = (c == 1) ? 1.f : 0.f;
            const float hsub = hs[p], zb = B.zb[p];
            if (!(hsub + zb <= hct)) {
               float hsm = 0.f, hdif = 0.f, hzz, gamf;
               const float gam0 = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
               for (int k = 1; k <= 4; k++) {
                  const int ii = i + cIIH[k], jj = j + cJJH[k];
                  bool self = true;
                  if (ii == 0 || jj == 0) {
                     gamf = 0.f;
                  } else {
                     gamf = gamf_of(B, k, gam0, ii, jj);
                     if (!(ii == B.imxb || jj == B.jmxb)) {
                        const size_t q = IDX(B, ii, jj);
                        if (cl[q] == c) {                     // HSUB(n) != 999
                           const float hq = hs[q];
                           if (!(hq + B.zb[q] <= hct)) {
                              const int kk = k % 4 + 1;
                              const float z1 = B.zbm[IDX(B, i + cIP[k], j + cJP[k])];
                              const float z2 = B.zbm[IDX(B, i + cIP[kk], j + cJP[kk])];
                              const float zzz = fminf(z1, z2) + hct;
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
               const float hbtemp = (1.f / 3.f) * hsub + (hsm + hdif) / 6.f;
               out = fmaxf(hbtemp, -zb);
            }
         }
      }
      if (B.flags[p] & F_BDRY) {
         const float *T = tabv(R, s, P);
         const float *DELP = T + T_DELP * SL_NTABV;
         const float xr = B.elpcl[i - 1] * B.cosl[j - 1], yr = B.elpdl[i - 1] * B.sinl[j - 1];
         const float x = xr - P[P_C1] - P[P_AX], y = yr - P[P_C2] - P[P_AY];
         const float rsq = x * x + y * y, r1 = sqrtf(rsq) / 5280.f + 1.f;
         int k = (int) r1;
         const float r2 = (float) k, dr = r1 - r2;
         if (k > 790) k = 790;
         out = DELP[k] + dr * (DELP[k + 1] - DELP[k]) + P[P_SEA] + 0.f;
      }
   }
   hb[p] = out;
}

// ------------------------------------------------------------------- HMXSV
__global__ void __launch_bounds__(TX * TY) k_hmxsv(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb - 1 || j > B.jmxb - 1) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const float *hb = A.hc + so;
   float *hmx = R.hmx + so;
   const float hx = hmx[p], h = hb[p];
   if (!(hx < h)) return;
   const float zb = B.zb[p];
   if (zb > 40.f || B.ebsn1_amp) {
      hmx[p] = h;
      return;
   }
   if (h + zb < 1.f) {
      hmx[p] = h;
      return;
   }
   float hsm = 0.f, hdif = 0.f, hzz, gamf;
   const float gam0 = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
   for (int k = 1; k <= 4; k++) {
      const int ii = i + cIIH[k], jj = j + cJJH[k];
      if (ii == 0 || jj == 0) {
         gamf = 0.f;
         hsm = hsm + h;
         hzz = h;
      } else {
         const float gamfk = B.elpdl2[ii - 1] + (B.elpcl2[ii - 1] - B.elpdl2[ii - 1]) * B.sinl2[jj - 1];
         const float z = gamfk / gam0;
         gamf = 2.f * z / (1.f + z) - 1.f;
         const size_t q = IDX(B, ii, jj);
         if ((jj != B.jmxb && ii != B.imxb) && hb[q] + B.zb[q] >= 1.f) {
            const int kk = k % 4 + 1;
            const float z1 = B.zbm[IDX(B, i + cIP[k], j + cJP[k])];
            const float z2 = B.zbm[IDX(B, i + cIP[kk], j + cJP[kk])];
            const float zzz = fminf(z1, z2) + 1.f;
            if (h < zzz || hb[q] < zzz) {
               hsm = hsm + h;
               hzz = h;
            } else {
               hsm = hsm + hb[q];
               hzz = hb[q];
            }
         } else {
            hsm = hsm + h;
            hzz = h;
         }
      }
      hdif = hdif + hzz * gamf;
   }
   const float htemp = (4.f * h + hsm + hdif) / 8.f;
   hmx[p] = fmaxf(hx, htemp);
}

// ------------------------------------------------------------ FILTER/XFILTER
__global__ void __launch_bounds__(TX * TY) k_filter(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb || j > B.jmxb) return;
   const size_t so = (size_t) s * R.fstride, p = IDX(B, i, j);
   const float *hs = A.hc + so;
   float *hb = A.hn + so;
   float out = hs[p];
   if (i <= B.imxb - 1 && j <= B.jmxb - 1) {
      const float zb = B.zb[p], hsub = hs[p];
      if (!(zb <= -B.hter)) {
         if (B.ismth == 1) {
            const float hcrt = 0.5f;
            if (!(hsub + zb < hcrt)) {
               float hsm = 0.f, hdif = 0.f, hzz, gamf;
               const float gam0 = B.elpdl2[i - 1] + (B.elpcl2[i - 1] - B.elpdl2[i - 1]) * B.sinl2[j - 1];
               for (int k = 1; k <= 4; k++) {
                  const int ii = i + cIIH[k], jj = j + cJJH[k];
                  if (ii == 0 || jj == 0) {
                     gamf = 0.f;
                     hsm = hsm + hsub;
                     hzz = hsub;
                  } else {
                     gamf = gamf_of(B, k, gam0, ii, jj);
                     const size_t q = IDX(B, ii, jj);
                     if (ii == B.imxb || jj == B.jmxb || hs[q] + B.zb[q] < hcrt) {
                        hsm = hsm + hsub;
                        hzz = hsub;
                     } else {
                        const int kk = k % 4 + 1;
                        const float z1 = B.zbm[IDX(B, i + cIP[k], j + cJP[k])];
                        const float z2 = B.zbm[IDX(B, i + cIP[kk], j + cJP[kk])];
                        const float zzz = fminf(z1, z2) + hcrt;
                        if (hsub < zzz || hs[q] < zzz) {
                           hsm = hsm + hsub;
                           hzz = hsub;
                        } else {
                           hsm = hsm + hs[q];
                           hzz = hs[q];
                        }
                     }
                  }
                  hdif = hdif + hzz * gamf;
               }
               const float htemp = 0.5f * hsub + (hsm + hdif) / 8.f;
               out = fmaxf(-zb, htemp);
            }
         } else if (!(hsub + zb < B.exclude)) {
            // XFILTER: HSUB(IMXB,.) is never written (0); HSUB(.,JMXB) is the
            // entering row JMXB saved by k_contty.
            const float hcrt = B.hcrt;
            const float *top = R.hsubtop + (size_t) s * B.pitch;
#define HX(a, b) ((a) == B.imxb ? 0.f : ((b) == B.jmxb ? top[(a) - 1] : hs[IDX(B, a, b)]))
#define ZX(a, b) (B.zbm[IDX(B, a, b)])
            const int IIX[5] = {0, -1, 1, 1, -1}, JJX[5] = {0, -1, -1, 1, 1};
            const int IP1[5] = {0, -1, 1, 2, 0}, JP1[5] = {0, 0, -1, 1, 2};
            const int IP2[5] = {0, 0, 2, 1, -1}, JP2[5] = {0, -1, 0, 2, 1};
            int maskii[5] = {0, 0, 0, 0, 0}, maskix[5] = {0, 0, 0, 0, 0};
            for (int k = 1; k <= 4; k++) {
               const int ii = i + cIIH[k], jj = j + cJJH[k], ix = i + IIX[k], jx = j + JJX[k];
               if (ii == 0 || jj == 0 || ix == 0 || jx == 0) continue;
               if (!(ii == B.imxb || jj == B.jmxb || HX(ii, jj) + B.zb[IDX(B, ii, jj)] < hcrt)) {
                  const int kk2 = k % 4 + 1;
                  const float zzz = fminf(ZX(i + cIP[k], j + cJP[k]), ZX(i + cIP[kk2], j + cJP[kk2])) + hcrt;
                  if (!(hsub < zzz && HX(ii, jj) < zzz)) maskii[k] = 1;
               }
               if (!(ix == B.imxb || jx == B.jmxb || HX(ix, jx) + B.zb[IDX(B, ix, jx)] < hcrt)) {
                  const int kk = k % 4 + 1, km1 = (k + 2) % 4 + 1;
                  const float z = ZX(i + cIP[k], j + cJP[k]), zz = ZX(i + cIP[kk], j + cJP[kk]);
                  const float zzz = fminf(z, zz) + hcrt;
                  const float yyy = fminf(z, ZX(i + IP2[k], j + JP2[k])) + hcrt;
                  const float xxx = fminf(z, ZX(i + IP1[k], j + JP1[k])) + hcrt;
                  const float www = fminf(z, ZX(i + cIP[km1], j + cJP[km1])) + hcrt;
                  const float hii = HX(ii, jj), hix = HX(ix, jx), ha1 = HX(i + IP1[k], j + JP1[k]);
                  if (((hsub >= zzz && hii >= yyy) || (hix >= yyy && hii >= zzz)) ||
                      ((hsub >= www && ha1 >= xxx) || (hix >= xxx && ha1 >= www)))
                     maskix[k] = 1;
               }
            }
            float hsm2 = 0.f, xsm2 = 0.f;
            for (int km = 1; km <= 4; km++) {
               const int kkm = km % 4 + 1, km1m = (km + 2) % 4 + 1;
               const int ix = i + IIX[km], jx = j + JJX[km], ii = i + cIIH[km], jj = j + cJJH[km];
               if (ii == 0 || jj == 0 || ix == 0 || jx == 0 || i + IIX[kkm] == 0 || j + JJX[kkm] == 0 ||
                   i + IIX[km1m] == 0 || j + JJX[km1m] == 0 || i + cIIH[kkm] == 0 ||
                   j + cJJH[kkm] == 0 || i + cIIH[km1m] == 0 || j + cJJH[km1m] == 0)
                  continue;
               hsm2 = hsm2 + (float) maskii[km] * HX(ii, jj) - (float) maskii[km] * hsub;
               if (B.xver == 1) {
                  const float w = (float) (maskix[km] * maskii[km] * maskix[kkm]);
                  xsm2 = xsm2 + 0.5f * (w * HX(ix, jx) - w * hsub);
               } else {
                  xsm2 = xsm2 + 0.5f * 1.f * (float) maskix[km] * (float) maskii[km] *
                                (float) maskii[km1m] * (HX(ix, jx) - hsub);
               }
            }
            const float htemp = hsub + 1.f / 8.f * (hsm2 - 1.f * (xsm2));
            out = fmaxf(-zb, htemp);
#undef HX
#undef ZX
         }
      }
   }
   hb[p] = out;
}

// ------------------------------------------------------------------ MOMNTM
struct Forcing {
   float dph;
   int id, iik;
};

__device__ static void frcpnt(const SlBasinDev &B, const float *P, const float *T, float x, float y,
                              float &fxb, float &fyb, int ncatg, float &cshltr, int npls,
                              const Forcing &F)
{
   const float xp = x - P[P_C1] - P[P_AX], yp = y - P[P_C2] - P[P_AY];
   const float rsq = xp * xp + yp * yp, rs = sqrtf(rsq), r1 = rs / 5280.f + 1.f;
   int k = (int) r1;
   const float r2 = (float) k, dr = r1 - r2;
   if (k > 790) k = 790;
   const float ccn = P[P_X15] + rsq;
   const float x1218 = (ncatg == 1) ? P[P_W0] : P[P_W1];
   float ck1, sk1;
   if (ncatg == 1) {
      const float *CW = T + T_CW * SL_NTABV, *SW = T + T_SW * SL_NTABV;
      ck1 = CW[k] + dr * (CW[k + 1] - CW[k]);
      sk1 = SW[k] + dr * (SW[k + 1] - SW[k]);
   } else {
      const float *C = T + T_C * SL_NTABV, *S = T + T_S * SL_NTABV;
      cshltr = 1.f;
      ck1 = C[k] + dr * (C[k + 1] - C[k]);
      sk1 = S[k] + dr * (S[k + 1] - S[k]);
   }
   float a = rs * P[P_X9] - x1218 * (yp * ck1 + xp * sk1);
   float b = rs * P[P_X10] + x1218 * (xp * ck1 - yp * sk1);
   if (ncatg == 1) {
      const float rhol = fabsf(xp * P[P_X20] + yp * P[P_X21]);
      const float ccc = ccn * rhol / (P[P_X15] + rhol * rhol);
      a = a + ccc * P[P_X23];
      b = b + ccc * P[P_X24];
   }
   a = a * P[P_X4] / ccn;
   b = b * P[P_X4] / ccn;
   float fctt = 3.E-6f * sqrtf(a * a + b * b);
   fctt = fctt * cshltr * P[P_BT];
   const float fxx = fctt * a, fyy = fctt * b;
   const float cr = TABI(B, T_CR2, F.iik, F.id), ci = TABI(B, T_CI2, F.iik, F.id);
   fxb = cr * fxx - ci * fyy;
   fyb = cr * fyy + ci * fxx;
   if (npls != 0) {
      const float *Pt = T + T_P * SL_NTABV;
      const float aa = -F.dph * (Pt[k] + dr * (Pt[k + 1] - Pt[k]));
      const float pxx = aa * xp, pyy = aa * yp;
      const float br = TABI(B, T_BR2, F.iik, F.id), bi = TABI(B, T_BI2, F.iik, F.id);
      fxb = fxb + br * pxx - bi * pyy;
      fyb = fyb + br * pyy + bi * pxx;
   }
}

__device__ static void mntmbd(const SlBasinDev &B, const float *P, const float *T, float *ub, float *vb,
                              int jump, int ijump, int isb, int jsb, float arj, float aii, size_t pij,
                              float ubj, float vbj, float &fxb, float &fyb, int ncatg, float &cshltr,
                              const Forcing &F)
{
   const size_t q = IDX(B, isb, jsb);
   const float aiul = aii * (ub[q] - ubj), aivb = aii * (vb[q] - vbj);
   if (jump == 2) {
      const float xttx = 0.f, xtty = 0.f;
      ub[q] = ub[pij] + aiul + arj * (vb[q] - vbj) + xttx;
      vb[q] = vb[pij] + aivb - arj * (ub[q] - ub[pij]) + xtty;
      return;
   }
   float fxbp, fybp;
   if (ijump != 2) {
      const float fxd = fxb, fyd = fyb;
      float xr = B.elpct[isb - 1] * B.cost[jsb - 1], yr = B.elpdt[isb - 1] * B.sint[jsb - 1];
      frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 1, F);
      xr = B.elpdt[isb - 1] * B.cost[jsb - 1];
      yr = B.elpct[isb - 1] * B.sint[jsb - 1];
      fxb = xr * fxbp + yr * fybp;
      fyb = xr * fybp - yr * fxbp;
      const float xttx = P[P_DELT] * (fxb - fxd), xtty = P[P_DELT] * (fyb - fyd);
      ub[q] = ub[pij] + aiul + arj * (vb[q] - vbj) + xttx;
      vb[q] = vb[pij] + aivb - arj * (ub[q] - ub[pij]) + xtty;
      return;
   }
   cshltr = 1.f;
   float xr = B.elpct[isb - 1] * B.cost[jsb - 1], yr = B.elpdt[isb - 1] * B.sint[jsb - 1];
   frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 0, F);
   xr = B.elpdt[isb - 1] * B.cost[jsb - 1];
   yr = B.elpct[isb - 1] * B.sint[jsb - 1];
   fxb = xr * fxbp + yr * fybp;
   fyb = xr * fybp - yr * fxbp;
   ub[q] = aii * ub[q] + arj * vb[q] + P[P_DELT] * fxb;
   vb[q] = aii * vb[q] - arj * ub[q] + P[P_DELT] * fyb;
}

__global__ void __launch_bounds__(TX * TY) k_momntm(SlBasinDev B, SlRunDev R, StepArgs A)
{
   const int i = blockIdx.x * TX + threadIdx.x + 1, j = blockIdx.y * TY + threadIdx.y + 1;
   const int s = blockIdx.z;
   if (i > B.imxb || j > B.jmxb) return;
   const size_t p = IDX(B, i, j);
   const uint32_t fl = B.flags[p];
   if (!(fl & F_MOM)) return;
   const size_t so = (size_t) s * R.fstride;
   const float *hb = A.hc + so;
   float *ub = R.ub + so, *vb = R.vb + so;
   const float *P = par(R, s, A.step);
   const float *T = tabv(R, s, P);
   const size_t pa = p - B.pitch - 1, pb = p - B.pitch, pc = p - 1;   // (i-1,j-1) (i,j-1) (i-1,j)
   const float h1 = hb[pa], h2 = hb[pb], h3 = hb[pc], h4 = hb[p];
   const float zbm = B.zbm[p];
   Forcing F;
   F.iik = (int) F_IIK(fl);
   const int ncatg = (fl & F_OCEAN) ? 2 : 1;
   int jump = 1;
   float dhx, dhy, aintrp, arj, aii, fxb = 0.f, fyb = 0.f, cshltr = 0.f, ubj, vbj;
   bool blocked = false;
   if (ncatg != 1 && (fl & F_IT2) && zbm < -50.f) {
      dhx = -h1 + h2 - h3 + h4;
      dhy = -h1 - h2 + h3 + h4;
      F.dph = .25f * (h1 + h2 + h3 + h4 + B.zb[pa] + B.zb[pb] + B.zb[pc] + B.zb[p]);
      const float dphz = fminf(599.f, F.dph);
      F.id = (int) fminf(599.f, dphz);
      if (F.id <= 0) F.id = 1;
      const float dphi = (float) F.id;
      aintrp = TABI(B, T_AI2, F.iik, F.id) + B.slpai[F.id] * (dphz - dphi);
   } else {
      const float hmn = fminf(fminf(fminf(h4, h3), h2), h1);
      if (hmn <= zbm) {
         const float hmxx = fmaxf(fmaxf(fmaxf(h4, h3), h2), h1);
         if (hmxx <= zbm) {
            blocked = true;
         } else {
            const float hf1 = fmaxf(h1, zbm), hf2 = fmaxf(h2, zbm), hf3 = fmaxf(h3, zbm), hf4 = fmaxf(h4, zbm);
            jump = 2;
            dhx = -hf1 + hf2 - hf3 + hf4;
            dhy = -hf1 - hf2 + hf3 + hf4;
         }
      } else {
         dhx = -h1 + h2 - h3 + h4;
         dhy = -h1 - h2 + h3 + h4;
      }
      if (!blocked) {
         float hpd[4];
         hpd[0] = h1 + B.zb[pa];
         hpd[1] = h2 + B.zb[pb];
         hpd[2] = h3 + B.zb[pc];
         hpd[3] = h4 + B.zb[p];
         F.dph = .25f * (hpd[0] + hpd[1] + hpd[2] + hpd[3]);
         F.id = (int) fminf(599.f, F.dph);
         if (F.dph <= 5.f) {
            if (jump != 2) {
               for (int k = 0; k < 4; k++) {
                  if (hpd[k] < 5.f) hpd[k] = hpd[k] * (.4f + .12f * hpd[k] - .008f * hpd[k] * hpd[k]);
                  else hpd[k] = hpd[k] - 1.f;
               }
               F.dph = .25f * (hpd[0] + hpd[1] + hpd[2] + hpd[3]);
               F.id = (int) fminf(599.f, F.dph);
            }
            if (F.id < 1) F.id = 1;
         }
         aintrp = 0.f;
         for (int k = 0; k < 4; k++) {
            const float dz = hpd[k];
            int idd = (int) fminf(599.f, dz);
            if (idd < 1) idd = 1;
            const float dphi = (float) idd;
            const float aitr = TABI(B, T_AI2, F.iik, idd) + B.slpai[idd] * (dz - dphi);
            aintrp = aintrp + aitr;
         }
         aintrp = .25f * aintrp;
      }
   }
   if (blocked) {
      ub[p] = 0.f;
      vb[p] = 0.f;
      ubj = 0.f;
      vbj = 0.f;
   } else {
      aii = 1.f + P[P_X1B4] * aintrp;
      arj = TABI(B, T_AR2, F.iik, F.id) * P[P_X1B4] * B.zcor[p];
      const float aiu = aii * ub[p], aiv = aii * vb[p];
      const float fct = P[P_X1B6] * F.dph;
      float xttx, xtty;
      if (jump != 2) {
         if (ncatg == 2) {
            cshltr = 1.f;
         } else {
            const float hf1 = h1 - zbm, hf2 = h2 - zbm, hf3 = h3 - zbm, hf4 = h4 - zbm;
            if (!(fl & F_TREE1)) {
               cshltr = .25f * (fminf(hf1, 1.f) + fminf(hf2, 1.f) + fminf(hf3, 1.f) + fminf(hf4, 1.f));
            } else {
               const float hfx[4] = {hf1, hf2, hf3, hf4};
               cshltr = 0.f;
               for (int kx = 0; kx < 4; kx++) {
                  float hfxx = 1.f;
                  if (hfx[kx] <= 0.f) hfxx = 0.f;
                  else if (hfx[kx] < 50.f) {
                     const float q = hfx[kx] / 50.f;
                     hfxx = q * q;
                  }
                  cshltr = cshltr + hfxx;
               }
               cshltr = .25f * cshltr;
            }
         }
         float fxbp, fybp;
         float xr = B.elpct[i - 1] * B.cost[j - 1], yr = B.elpdt[i - 1] * B.sint[j - 1];
         frcpnt(B, P, T, xr, yr, fxbp, fybp, ncatg, cshltr, 1, F);
         xr = B.elpdt[i - 1] * B.cost[j - 1];
         yr = B.elpct[i - 1] * B.sint[j - 1];
         fxb = (xr * fxbp + yr * fybp);
         fyb = (xr * fybp - yr * fxbp);
         xttx = P[P_DELT] * fxb;
         xtty = P[P_DELT] * fyb;
      } else {
         xttx = 0.f;
         xtty = 0.f;
      }
      ubj = ub[p];
      vbj = vb[p];
      const float br = TABI(B, T_BR2, F.iik, F.id), bi = TABI(B, T_BI2, F.iik, F.id);
      ub[p] = aiu - fct * (br * dhx - bi * dhy) + arj * vb[p] + xttx;
      vb[p] = aiv - fct * (br * dhy + bi * dhx) - arj * ub[p] + xtty;
   }
   // boundary transports (momntm.f:384-458)
   const uint32_t bop = F_BOP(fl);
   if (bop == 0) return;
   const float uij = ub[p];
   int ijump = 1;
   if (bop <= 4) {
      const int ISBS[5] = {0, 0, -1, 0, 1}, JSBS[5] = {0, -1, 0, 1, 0};
      const int isb = i + ISBS[bop], jsb = j + JSBS[bop];
      const size_t q = IDX(B, isb, jsb);
      if (uij == 0.f) {
         ub[q] = 0.f;
         vb[q] = 0.f;
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
      const int ishf[4][3] = {{0, 0, 0}, {1, 2, 4}, {0, 0, 0}, {1, 3, 4}};
      for (int kk = 0; kk < 3; kk++) {
         const int k = ishf[icor][kk];
         const int isb = i + cIH[icor] + cIP[k], jsb = j + cJH[icor] + cJP[k];
         const size_t q = IDX(B, isb, jsb);
         if (uij != 0.f) {
            const uint32_t fq = B.flags[q];
            if (fq & F_BSKIP) break;
            if (fq & F_B8) ijump = 2;
            mntmbd(B, P, T, ub, vb, jump, ijump, isb, jsb, arj, aii, p, ubj, vbj, fxb, fyb, ncatg,
                   cshltr, F);
         } else {
            ub[q] = 0.f;
            vb[q] = 0.f;
         }
      }
   }
}

// ------------------------------------------------------------------ launch
// Kernel slots of one step, for per-kernel timing (SL_NKERN in sl_dev.h).
extern "C" int sl_dev_launch_step(const SlBasinDev *B, const SlRunDev *R, int step, int smpt,
                                  int filt, int parity, hipStream_t st, hipEvent_t *ev)
{
   dim3 blk(TX, TY), grd((B->imxb + TX - 1) / TX, (B->jmxb + TY - 1) / TY, R->nstorm);
   float *h0 = parity ? R->hb1 : R->hb0, *h1 = parity ? R->hb0 : R->hb1;
   StepArgs a;
   a.step = step;
#define MARK(k) \
   if (ev) (void) hipEventRecord(ev[k], st)
   // CONTTY: h0 (HSUB) -> h1
   a.hc = h0;
   a.hn = h1;
   MARK(0);
   hipLaunchKernelGGL(k_contty, grd, blk, 0, st, *B, *R, a);
   MARK(1);
   hipLaunchKernelGGL(k_replay, dim3((R->nstorm + 63) / 64), dim3(64), 0, st, *B, *R, a);
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
   // FLT2G + BDRYHT + row copy: h1 -> h0
   hipLaunchKernelGGL(k_smooth, grd, blk, 0, st, *B, *R, a);
   MARK(5);
   // HMXSV on h0
   a.hc = h0;
   hipLaunchKernelGGL(k_hmxsv, grd, blk, 0, st, *B, *R, a);
   MARK(6);
   float *fin = h0;
   if (filt) {
      a.hc = h0;
      a.hn = h1;
      hipLaunchKernelGGL(k_filter, grd, blk, 0, st, *B, *R, a);
      fin = h1;
   }
   MARK(7);
   a.hc = fin;
   hipLaunchKernelGGL(k_momntm, grd, blk, 0, st, *B, *R, a);
   MARK(8);
#undef MARK
   hipError_t e = hipGetLastError();
   if (e != hipSuccess) return -(int) e;
   return (fin == R->hb0) ? 0 : 1;   // which buffer holds HB after the step
}
