// sl_capi.cpp -- C ABI of the SLOSH time-stepping core on MI355X.
//
// A basin is loaded once (host restatement of the reference's basin init, then
// float static fields uploaded to HBM); a batch holds one or more storm runs on
// that basin with their fields resident on the device.  Each storm's forcing
// depends only on its track, so the whole per-step storm schedule (the values
// STMVAL leaves for the next step, and every radial table FORCE1 produces) is
// fixed before the first step.  The host walks each storm once (track,
// STERON, SETCMP, STMVAL with its FORCE1 table refreshes) and keeps a compact
// storm program: the state after each STMVAL call that changed it.  At upload
// the device expands the programs into the per-step values of every run
// (k_sched); stepping then needs no host round trip.
#include <map>
#include <mutex>
#include "sl_int.h"

// --------------------------------------------------------------------------
// basin

template <class R>
static uint32_t flag_bits(const slg_basin *B, int i, int j)
{
   const sl_basin *b = &B->h;
   uint32_t f = 0;
   char it = b->itree[PIX(b, i, j)];
   if (i <= b->imxb1 && j <= b->jmxb1 && b->kskp[PIX(b, i, j)] == '1') f |= F_ACT;
   if (it == '1') f |= F_TREE1;
   if (it == '2' || it == '5') f |= F_OCEAN;
   if (it == '2') f |= F_IT2;
   if (it == '6' || it == '4') f |= F_BSKIP;
   if (it == '8') f |= F_B8;
   if (i >= 2 && i <= b->imxb1 && j >= 2 && j <= b->jmxb1) f |= F_ZR;
   // SMPT2G candidate momentum point (smpt2g.f:23-30): interior and ZBM > -30
   if (i >= 2 && i <= b->imxb - 2 && j >= 2 && j <= b->jmxb - 2 && (R) b->zbm[PIX(b, i, j)] > (R) -30.)
      f |= F_PFC;
   // IIK of FRCPNT/MOMNTM (frcpnt.f:75-118)
   {
      int iik;
      if (b->zb[PIX(b, i, j)] >= 0) iik = (b->iver == 201903) ? 4 : 1;
      else if (b->iver == 201903) {
         int m = b->imann[PIX(b, i, j)];
         iik = (m >= 100) ? ((m - 5 > 245) ? 245 : m - 5) : 95;
      } else
         iik = !strncmp(b->sta, "SOUTH FL", 8) ? 245 : 1;
      f |= (uint32_t) iik << 8;
   }
   // MOMNTM interior points and their boundary duty (momntm.f:96-130, 384-458)
   if (j >= 2 && j <= B->jend && i >= b->ms[j - 1] && i <= b->me[j - 1] && it != '4' && it != '6') {
      int iskip = 2, bop = 0;
      f |= F_MOM;
      if (!B->island) {   // an island basin has no side boundaries in J (momntm.f:105-112)
         if (j == 2) iskip = 1;
         if (j == B->jend) iskip = 3;
      }
      if (iskip == 3) bop = (i != b->ms[j - 1]) ? 3 : 6;
      else if (iskip == 1) bop = (i != b->ms[j - 1]) ? 1 : 5;
      else if (i == b->ms[j - 1]) bop = 2;
      else if (i == b->me[j - 1]) bop = 4;
      f |= (uint32_t) bop << 16;
   }
   return f;
}

// Bounds of the grid loops (shared by both precisions).
static void basin_bounds(slg_basin *B)
{
   sl_basin *b = &B->h;
   const int NJ = b->jmxb;
   B->is.assign(b->is, b->is + NJ);
   B->ie.assign(b->ie, b->ie + NJ);
   B->ms.assign(b->ms, b->ms + NJ);
   B->me.assign(b->me, b->me + NJ);
}

// Static fields in precision R: every derived value is evaluated in R, as the
// reference does in its REAL (double in the default build; float for the FP32
// step, as the FP32 restatement evaluates them).
template <class R>
static void basin_images(slg_basin *B)
{
   sl_basin *b = &B->h;
   BasinImg<R> &I = img(B, R());
   const int P = B->pitch, NJ = b->jmxb;
   size_t np = (size_t) P * NJ;
   I.zb.assign(np, R(0));
   I.zbm.assign(np, R(0));
   I.zcor.assign(np, R(0));
   I.flags.assign(np, 0);
   R corl[90];
   for (int l = 0; l < 90; l++) corl[l] = (R) b->corl[l];
   for (int j = 1; j <= NJ; j++)
      for (int i = 1; i <= b->imxb; i++) {
         size_t q = (size_t) (j - 1) * P + (i - 1);
         I.zb[q] = (R) b->zb[PIX(b, i, j)];
         I.zbm[q] = (R) b->zbm[PIX(b, i, j)];
         I.flags[q] = flag_bits<R>(B, i, j);
         // Coriolis factor at the point as momntm.f:294-295 evaluates it
         R ylt = (R) b->ylt[PIX(b, i, j)];
         int llat = (int) ylt;
         if (llat >= 0 && llat <= 88)
            I.zcor[q] = corl[llat] + (ylt - (R) llat) * (corl[llat + 1] - corl[llat]);
         else if ((I.flags[q] & F_MOM) && !I.lat_i) {   // outside CORL(1..90): refused at upload
            I.lat_i = i;
            I.lat_j = j;
            I.lat_v = (double) ylt;
         }
      }
   // BDRYHT squares (the ISH/JSH list DEPSFC builds, depsfc.f:376-390)
   for (int n = 0; n < b->nbcpt; n++)
      I.flags[(size_t) (b->jsh[n] - 1) * P + (b->ish[n] - 1)] |= F_BDRY;
   double *vsrc[6] = {b->elpct, b->elpdt, b->elpcl, b->elpdl, b->elpcl2, b->elpdl2};
   for (int k = 0; k < 6; k++) {
      I.vi[k].assign(P, R(0));
      for (int i = 0; i < b->imxb; i++) I.vi[k][i] = (R) vsrc[k][i];
   }
   double *jsrc[5] = {b->sint, b->cost, b->sinl, b->cosl, b->sinl2};
   for (int k = 0; k < 5; k++) {
      I.vj[k].assign(NJ, R(0));
      for (int j = 0; j < NJ; j++) I.vj[k][j] = (R) jsrc[k][j];
   }
   double *tsrc[6] = {b->ar2, b->ai2, b->br2, b->bi2, b->cr2, b->ci2};
   I.tab.assign((size_t) 6 * 245 * 600, R(0));
   for (int t = 0; t < 6; t++)
      for (int ii = 1; ii <= 245; ii++)
         for (int n = 1; n <= 600; n++)
            I.tab[(size_t) t * 147000 + (size_t) (ii - 1) * 600 + (n - 1)] =
                (R) tsrc[t][(size_t) (n - 1) * 245 + (ii - 1)];
   // SLPAI of momntm.f:81-89, from the AI2 column in R
   I.slpai.assign(601, R(0));
   int col = (b->iver == 201903) ? 4 : 1;
   for (int i = 1; i <= 599; i++)
      I.slpai[i] = I.tab[(size_t) T_AI2_OFF + (size_t) (col - 1) * 600 + i] -
                   I.tab[(size_t) T_AI2_OFF + (size_t) (col - 1) * 600 + (i - 1)];
   I.slpai[600] = I.slpai[599];
   // Device layout: one record per (IIK, depth): AI2 SLPAI AR2 BR2 BI2 CR2 CI2 0
   I.tab8.assign((size_t) 245 * 600 * 8, R(0));
   for (int ii = 1; ii <= 245; ii++)
      for (int n = 1; n <= 600; n++) {
         R *e = &I.tab8[((size_t) (ii - 1) * 600 + (n - 1)) * 8];
         const size_t o = (size_t) (ii - 1) * 600 + (n - 1);
         e[0] = I.tab[147000 * 1 + o];   // AI2
         e[1] = I.slpai[n];
         e[2] = I.tab[147000 * 0 + o];   // AR2
         e[3] = I.tab[147000 * 2 + o];   // BR2
         e[4] = I.tab[147000 * 3 + o];   // BI2
         e[5] = I.tab[147000 * 4 + o];   // CR2
         e[6] = I.tab[147000 * 5 + o];   // CI2
      }
   // Metric weights GAMF of the four neighbours of every square, as HMXSV
   // (hmxsv.f:84-92) and, on elliptic/hyperbolic grids, FLT2G and FILTER
   // evaluate them: z = GAMFK/GAM0, GAMF = 2z/(1+z) - 1, in R from the metric
   // images (the same operations the kernels would do every step).
   I.g4.assign((size_t) 4 * np, R(0));
   {
      const std::vector<R> &cl2 = I.vi[4], &dl2 = I.vi[5], &sl2 = I.vj[4];
      static const int iih[4] = {0, 1, 0, -1}, jjh[4] = {-1, 0, 1, 0};
      for (int j = 1; j <= NJ - 1; j++)
         for (int i = 1; i <= b->imxb - 1; i++) {
            const R gam0 = dl2[i - 1] + (cl2[i - 1] - dl2[i - 1]) * sl2[j - 1];
            for (int k = 0; k < 4; k++) {
               const int ii = i + iih[k], jj = j + jjh[k];
               if (ii == 0 || jj == 0) continue;
               const R gamfk = dl2[ii - 1] + (cl2[ii - 1] - dl2[ii - 1]) * sl2[jj - 1];
               const R z = gamfk / gam0;
               I.g4[(size_t) k * np + (size_t) (j - 1) * P + (i - 1)] = R(2) * z / (R(1) + z) - R(1);
            }
         }
   }
   // sub-grid 1-D flow sides: indices, friction class, KTREE 'T', and the
   // compact index of each joined square (HSUB of those squares lives in a
   // per-storm array of nsq values); every side joins two interior squares
   {
      static const int IHH[5] = {0, -1, 1, 0, 0}, JHH[5] = {0, 0, 0, -1, 1};
      const int nl = b->nsqrwc, nc = b->ncut;
      const bool sfl = !strncmp(b->sta, "SOUTH FL", 8);
      std::map<long, int> sq;
      auto qidx = [&](int i, int j) {
         auto it = sq.emplace((long) j * 100000 + i, (int) sq.size()).first;
         return it->second;
      };
      I.eint.assign((size_t) std::max(nl, 1) * SL_EI, 0);
      I.ereal.assign((size_t) std::max(nl, 1) * 8, R(0));
      I.ecut.assign((size_t) std::max(nc, 1) * 4, R(0));
      for (int l = 0; l < nl; l++) {
         int *e = &I.eint[(size_t) l * SL_EI];
         const int i = b->isqr[l], j = b->jsqr[l], k = b->iside[l];
         e[E1_I] = i;
         e[E1_J] = j;
         e[E1_K] = k;
         e[E1_II] = i + IHH[k];
         e[E1_JJ] = j + JHH[k];
         e[E1_IIK] = (b->zb[PIX(b, i, j)] >= 0 || !sfl) ? 1 : 245;   // flw1dm_block_m.f:173-222
         e[E1_T] = b->ktree[l] == 'T';
         e[E1_QI] = qidx(i, j);
         e[E1_QE] = qidx(e[E1_II], e[E1_JJ]);
         R *r = &I.ereal[(size_t) l * 8];
         r[E2_HWEIR] = (R) b->hweir[l];
         r[E2_ZBMIN] = (R) b->zbmin[l];
         r[E2_BANK1] = (R) b->bank1[l];
         r[E2_BANK2] = (R) b->bank2[l];
         r[E2_DELCUT] = (R) b->delcut[l];
         r[E2_COS] = (R) b->cos1d[l];
         r[E2_SIN] = (R) b->sin1d[l];
      }
      for (int n = 0; n < nc; n++) {
         I.ecut[(size_t) n * 4 + 0] = (R) b->cutl[n];
         I.ecut[(size_t) n * 4 + 1] = (R) b->cutli[n];
         I.ecut[(size_t) n * 4 + 2] = (R) b->cutle[n];
      }
      I.nsq = (int) sq.size();
   }
   I.built = true;
}

template <class R>
static int basin_upload(slg_basin *B, char *err, int errlen)
{
   BasinImg<R> &I = img(B, R());
   if (!I.built) basin_images<R>(B);
   if (I.lat_i) {
      // MOMNTM takes the Coriolis factor from CORL(LLAT+1), CORL(LLAT+2) with
      // LLAT = INT(YLT) (momntm.f:294-295): a momentum point whose latitude
      // truncates outside 0..88 reads outside the table and has no defined value
      if (err && errlen > 0)
         snprintf(err, errlen, "grid point (%d,%d) lies at latitude %.3f, outside the Coriolis table (0 to 89)",
                  I.lat_i, I.lat_j, I.lat_v);
      return SLG_EINPUT;
   }
   sl_basin *b = &B->h;
   const int P = B->pitch, NJ = b->jmxb;
   const size_t np = (size_t) P * NJ, rs = sizeof(R);
   size_t off[32], tot = 0;
   int n = 0;
   auto add = [&](size_t bytes) {
      off[n++] = tot;
      tot += (bytes + 255) & ~(size_t) 255;
   };
   add(np * rs);   // zb
   add(np * rs);   // zbm
   add(np * rs);   // zcor
   add(np * 4);    // flags
   for (int k = 0; k < 4; k++) add((size_t) NJ * 4);   // is ie ms me
   for (int k = 0; k < 6; k++) add((size_t) P * rs);
   for (int k = 0; k < 5; k++) add((size_t) NJ * rs);
   add(I.tab8.size() * rs);
   add(4 * np * rs);   // g4
   add(I.eint.size() * 4);
   add(I.ereal.size() * rs);
   add(I.ecut.size() * rs);
   add(np * 8);        // ZB in double
   add(np * 4);        // BDRYHT list index per point
   char *d = nullptr;
   HIPOK(hipMalloc((void **) &d, tot));
   I.dmem = d;
   int m = 0;
   hipError_t upe = hipSuccess;
   auto up = [&](const void *src, size_t bytes) -> char * {
      char *dst = d + off[m++];
      const hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
      if (upe == hipSuccess) upe = e;
      return dst;
   };
   SlBasinDevT<R> &D = I.dev;
   memset(&D, 0, sizeof D);
   D.imxb = b->imxb;
   D.jmxb = b->jmxb;
   D.pitch = P;
   D.npts = (int) np;
   D.jend = B->jend;
   D.island = B->island;
   D.ebsn_ell = (b->ebsn == '$' || b->ebsn == '+');
   D.ebsn1_amp = (b->ebsn1 == '&');
   D.ismth = b->ismth;
   D.xver = strncmp(b->sta, "CH2 (2008)", 10) == 0 ? 1 : 2;
   D.hter = (R) b->hter;
   {
      // HCRT and EXCLUDE of XFILTER by smoothing code (xfilter.f:176-207)
      static const double hc[12] = {0, 0, .5, .1, .2, .3, .4, 1.E-5, .2, .3, .4, .5};
      D.hcrt = (b->ismth >= 2 && b->ismth <= 11) ? (R) hc[b->ismth] : R(0);
      D.exclude = (b->ismth >= 7) ? (R) 1.E-5 : D.hcrt;
   }
   for (int k = 0; k < 4; k++) D.gamfp[k] = (R) b->gamfp[k];
   D.nbcpt = b->nbcpt;
   D.zb = (const R *) up(I.zb.data(), np * rs);
   D.zbm = (const R *) up(I.zbm.data(), np * rs);
   D.zcor = (const R *) up(I.zcor.data(), np * rs);
   D.flags = (const uint32_t *) up(I.flags.data(), np * 4);
   D.is = (const int *) up(B->is.data(), (size_t) NJ * 4);
   D.ie = (const int *) up(B->ie.data(), (size_t) NJ * 4);
   D.ms = (const int *) up(B->ms.data(), (size_t) NJ * 4);
   D.me = (const int *) up(B->me.data(), (size_t) NJ * 4);
   const R **vi[6] = {&D.elpct, &D.elpdt, &D.elpcl, &D.elpdl, &D.elpcl2, &D.elpdl2};
   for (int k = 0; k < 6; k++) *vi[k] = (const R *) up(I.vi[k].data(), (size_t) P * rs);
   const R **vj[5] = {&D.sint, &D.cost, &D.sinl, &D.cosl, &D.sinl2};
   for (int k = 0; k < 5; k++) *vj[k] = (const R *) up(I.vj[k].data(), (size_t) NJ * rs);
   D.tab8 = (const R *) up(I.tab8.data(), I.tab8.size() * rs);
   D.g4 = (const R *) up(I.g4.data(), 4 * np * rs);
   D.e_int = (const int *) up(I.eint.data(), I.eint.size() * 4);
   D.e_real = (const R *) up(I.ereal.data(), I.ereal.size() * rs);
   D.e_cut = (const R *) up(I.ecut.data(), I.ecut.size() * rs);
   D.nsqrs = b->nsqrs;
   D.nsqrw = b->nsqrw;
   D.nsqrwc = b->nsqrwc;
   D.nsq = I.nsq;
   D.f1d_on = b->nsqrwc != 0 && b->nof1d != '+';
   D.sfl = !strncmp(b->sta, "SOUTH FL", 8);
   D.npss = b->npss;
   for (int k = 0; k < 5; k++) {
      for (int e = 0; e < 2; e++) {
         D.ch_pt[k][0 + 2 * e] = b->ipt0[k][e];
         D.ch_pt[k][1 + 2 * e] = b->jpt0[k][e];
         D.ch_pt[k][4 + 2 * e] = b->iptl[k][e];
         D.ch_pt[k][5 + 2 * e] = b->jptl[k][e];
      }
      D.ch_lnh[k] = (R) b->chlnh[k];
      D.ch_wth[k] = (R) b->chwth[k];
      D.ch_dph[k] = (R) b->chdph[k];
      D.ch_a0[k] = (R) b->aresq0[k];
      D.ch_al[k] = (R) b->aresql[k];
   }
   D.aman = (R) b->aman;
   D.bman = (R) b->bman;
   D.gb = (R) b->gb;
   D.agb = (R) b->agb;
   D.agb13 = (R) b->agb13;
   {
      std::vector<double> zd(np, 0.);
      for (int j = 1; j <= NJ; j++)
         for (int i = 1; i <= b->imxb; i++) zd[(size_t) (j - 1) * P + (i - 1)] = b->zb[PIX(b, i, j)];
      I.zbd = (const double *) up(zd.data(), np * 8);
   }
   {
      std::vector<int> bi(np, -1);
      for (int n = 0; n < b->nbcpt; n++) bi[(size_t) (b->jsh[n] - 1) * P + (b->ish[n] - 1)] = n;
      D.bcidx = (const int *) up(bi.data(), np * 4);
   }
   HIPOK(upe);
   HIPOK(hipDeviceSynchronize());
   I.on_dev = true;
   return 0;
}

int sl_basin_ensure(slg_basin *B, int prec, char *err, int errlen)
{
   if (prec == 64) return B->i64.on_dev ? 0 : basin_upload<double>(B, err, errlen);
   return B->i32.on_dev ? 0 : basin_upload<float>(B, err, errlen);
}

extern "C" int slg_basin_open(const char *dta, const char *bnt_line, slg_basin **out, char *err,
                              int errlen)
{
   return slg_basin_open_opts(dta, bnt_line, 0, out, err, errlen);
}

extern "C" int slg_basin_open_opts(const char *dta, const char *bnt_line, int flags, slg_basin **out, char *err,
                                   int errlen)
{
   char e[256];
   *out = nullptr;
   if (flags & ~SLG_OPEN_FULL_DEPTH) {
      seterr(err, errlen, "%s", "unknown basin option");
      return SLG_EINPUT;
   }
   slg_basin *B = new slg_basin();
   if (sl_basin_load_opts(dta, &B->h, (flags & SLG_OPEN_FULL_DEPTH) != 0, e)) {
      seterr(err, errlen, "%s", e);
      delete B;
      return SLG_EINPUT;
   }
   sl_basin *b = &B->h;
   if (bnt_line && bnt_line[0]) {
      if (sl_basin_llx(b, bnt_line, e)) {
         seterr(err, errlen, "%s", e);
         sl_basin_free(b);
         delete B;
         return SLG_EINPUT;
      }
      strncpy(B->bnt, bnt_line, sizeof B->bnt - 1);
   }
   B->pitch = (b->imxb + 63) & ~63;
   // MOMNTM's last row: JMXB on a closed island basin, whose row 1 repeats row
   // JMXB (momntm.f:98-99, 464-469)
   B->island = !strcmp(b->dollar, "$ ");
   B->jend = B->island ? b->jmxb : b->jmxb1;
   // MOMNTM reads the four squares around a point and writes the boundary point
   // beside it; the momentum rows must leave column 1 and column IMXB for those.
   for (int j = 2; j <= B->jend; j++)
      if (b->ms[j - 1] <= b->me[j - 1] && (b->ms[j - 1] < 2 || b->me[j - 1] > b->imxb - 1)) {
         seterr(err, errlen, "momentum row %d spans columns outside 2..IMXB-1", "", j);
         sl_basin_free(b);
         delete B;
         return SLG_EUNSUP;
      }
   char e1 = b->ebsn1;
   B->incsm = 20;
   if (e1 == '+' || e1 == '=') B->incsm = 1;
   if (e1 == '!' || e1 == '#') B->incsm = 5;
   if (e1 == '@' || e1 == '%') B->incsm = 10;
   if (e1 == '^' || e1 == '*') B->incsm = 30;
   B->ismpt2g = !(e1 == '+' || e1 == '&' || e1 == '#' || e1 == '%' || e1 == '*');
   basin_bounds(B);
   *out = B;
   return 0;
}

extern "C" int slg_basin_set_ylt(slg_basin *B, const double *ylt)
{
   sl_basin *b = &B->h;
   if (B->i32.on_dev || B->i64.on_dev) return SLG_EINPUT;   // before the first upload only
   memcpy(b->ylt, ylt, sizeof(double) * (size_t) b->imxb * b->jmxb);
   b->have_ylt = 1;
   B->i32.built = B->i64.built = false;
   return 0;
}

extern "C" int slg_basin_sides(const slg_basin *B, int *nsqrs, int *nsqrw, int *nsqrwc)
{
   if (nsqrs) *nsqrs = B->h.nsqrs;
   if (nsqrw) *nsqrw = B->h.nsqrw;
   if (nsqrwc) *nsqrwc = B->h.nsqrwc;
   return 0;
}

extern "C" int slg_basin_dims(const slg_basin *B, int *imxb, int *jmxb, int *nbcpt)
{
   *imxb = B->h.imxb;
   *jmxb = B->h.jmxb;
   if (nbcpt) *nbcpt = B->h.nbcpt;
   return 0;
}

extern "C" void slg_basin_close(slg_basin *B)
{
   if (!B) return;
   if (B->i32.dmem) (void) hipFree(B->i32.dmem);
   if (B->i64.dmem) (void) hipFree(B->i64.dmem);
   sl_basin_free(&B->h);
   delete B;
}

// basins with sub-grid 1-D flow sides or channels (their passes run in kernels
// of their own between SMPT2G and HMXSV, on either path)
static bool has_sides(const slg_basin *B)
{
   return B->h.nsqrw > 0 || (B->h.nsqrwc > 0 && B->h.nof1d != '+') || B->h.npss > 0;
}

// --------------------------------------------------------------------------
// storm schedule: ref-format rows (oracle/ref_driver.c dump_storm layout)

static void storm_row(const sl_storm *s, const sl_basin *b, double *v)
{
   memset(v, 0, 90 * sizeof(double));
   v[0] = s->itime;   v[1] = s->bt;       v[2] = s->delt;   v[3] = s->c1;     v[4] = s->c2;
   v[5] = s->ax;      v[6] = s->ay;       v[7] = s->c21;    v[8] = s->c22;    v[9] = s->ydelp;
   v[10] = s->zdelp;  v[11] = s->pnn;     v[12] = s->yc24;  v[13] = s->zc24;  v[14] = s->treal;
   v[15] = s->w1218[0]; v[16] = s->w1218[1]; v[17] = s->wmax[0]; v[18] = s->wmax[1];
   v[19] = s->itmadv; v[20] = s->nhrad;   v[21] = s->ptency; v[22] = s->rtency; v[23] = s->mhalt;
   v[24] = s->x1b[3]; v[25] = s->x1b[4];  v[26] = s->x1b[5]; v[27] = s->ic19;  v[28] = s->seadtm;
   v[29] = s->jhr;    v[30] = s->zlato;  v[31] = s->x1b[0]; v[32] = s->x1b[1];
   for (int k = 1; k <= 50; k++) v[40 + k - 1] = s->x12[k];
   (void) b;
}

template <class R>
static void push_tables(std::vector<R> &T, const sl_storm *s)
{
   // device record k: C S CW SW P DELP 0 0
   const double *src[6] = {s->C, s->S, s->CW, s->SW, s->P, s->DELP};
   size_t base = T.size();
   T.resize(base + (size_t) SL_NTABV * 8, R(0));
   for (int k = 1; k <= 800; k++)
      for (int t = 0; t < 6; t++) T[base + (size_t) k * 8 + t] = (R) src[t][k];
}

static void push_tables(storm_sched &S, const sl_storm *s, int step, int prec)
{
   if (prec == 64) push_tables(S.t64, s);
   else push_tables(S.t32, s);
   S.tab_step.push_back(step);
}

template <class R>
static void step_params(const slg_basin *B, const double *r, int tv, R *P);

// the storm state the per-step values come from (SQ_* order, sl_dev.h)
static void sched_state(const sl_storm *s, int tv, double *q)
{
   q[SQ_DELT] = s->delt;
   q[SQ_X1B0] = s->x1b[0];
   q[SQ_X1B1] = s->x1b[1];
   q[SQ_X1B3] = s->x1b[3];
   q[SQ_X1B4] = s->x1b[4];
   q[SQ_X1B5] = s->x1b[5];
   q[SQ_C1] = s->c1;
   q[SQ_C2] = s->c2;
   q[SQ_X4] = s->x12[4];
   q[SQ_X9] = s->x12[9];
   q[SQ_X10] = s->x12[10];
   q[SQ_X11] = s->x12[11];
   q[SQ_X12] = s->x12[12];
   q[SQ_X15] = s->x12[15];
   q[SQ_X20] = s->x12[20];
   q[SQ_X21] = s->x12[21];
   q[SQ_X23] = s->x12[23];
   q[SQ_X24] = s->x12[24];
   q[SQ_W0] = s->w1218[0];
   q[SQ_W1] = s->w1218[1];
   q[SQ_SEA] = s->seadtm;
   q[SQ_TV] = tv;
   q[SQ_AX] = s->ax;
   q[SQ_AY] = s->ay;
}

static double ttimer_bt(int itime, int ic19);

int sl_make_schedule(const slg_basin *B, const char *trk, storm_sched &S, bool keep_rows, int prec,
                     char *err, int errlen, bool hostp, const slg_track *track)
{
   char e[256];
   double row[90], q[SQ_N], q0[SQ_N];
   int rc;
   if (track) {
      sl_trk *t = (sl_trk *) malloc(sizeof *t);
      if (!t) {
         seterr(err, errlen, "%s", "out of host memory");
         return SLG_EINPUT;
      }
      rc = sl_track_from_mem(track, t, e);
      if (!rc) rc = sl_storm_init(t, &B->h, &S.s0, e);
      free(t);
   } else
      rc = sl_storm_load(trk, &B->h, &S.s0, e);
   if (rc) {
      seterr(err, errlen, "%s", e);
      return SLG_EINPUT;
   }
   sl_storm s = S.s0;
   S.rows.clear();
   S.p32.clear();
   S.p64.clear();
   S.t32.clear();
   S.t64.clear();
   S.tab_step.clear();
   S.prog.clear();
   S.hostp = hostp;
   storm_row(&s, &B->h, row);
   if (keep_rows) S.rows.insert(S.rows.end(), row, row + 90);
   push_tables(S, &s, 0, prec);
   // storm program: header, the state after SETCMP, TTIMER's ramp; events follow
   const int ic19 = s.ic19;
   S.prog.assign(SQ_HDR, 0.);
   S.prog[1] = ic19;
   S.prog[3] = s.dtmlak;
   sched_state(&s, 0, q0);
   S.prog.insert(S.prog.end(), q0, q0 + SQ_N);
   S.prog.insert(S.prog.end(), s.DELP, s.DELP + SL_NTABV);   // INTLHT's table (the first hour's)
   for (int it = 1; it < ic19; it++) S.prog.push_back(ttimer_bt(it, ic19));
   int gen = s.table_gen, n = 0, nev = 0;
   if (hostp) {
      if (prec == 64) S.p64.reserve((size_t) (s.mhalt + 1) * SL_NPAR);
      else S.p32.reserve((size_t) (s.mhalt + 1) * SL_NPAR);
   }
   while (s.itime < s.mhalt) {
      // values step n+1 reads: the row after step n, and the table set in force
      const int tv = (int) S.tab_step.size() - 1;
      if (hostp && prec == 64) {
         size_t o = S.p64.size();
         S.p64.resize(o + SL_NPAR);
         step_params(B, row, tv, &S.p64[o]);
      } else if (hostp) {
         size_t o = S.p32.size();
         S.p32.resize(o + SL_NPAR);
         step_params(B, row, tv, &S.p32[o]);
      }
      sl_storm_ttimer(&s);
      if (sl_storm_stmval(&s, &B->h)) {
         seterr(err, errlen, "%s", "FORCE1 did not converge");
         return SLG_EINPUT;
      }
      n++;
      if (keep_rows || hostp) storm_row(&s, &B->h, row);
      if (keep_rows) S.rows.insert(S.rows.end(), row, row + 90);
      if (s.table_gen != gen) {
         push_tables(S, &s, n, prec);
         gen = s.table_gen;
      }
      // an event wherever this STMVAL call changed more than AX and AY
      sched_state(&s, (int) S.tab_step.size() - 1, q);
      if (memcmp(q, q0, sizeof(double) * SQ_AX)) {
         S.prog.push_back(n);
         S.prog.insert(S.prog.end(), q, q + SQ_AX);
         memcpy(q0, q, sizeof q0);
         nev++;
      }
   }
   S.nsteps = n;
   S.prog[0] = n;
   S.prog[2] = nev;
   return 0;
}

// device parameters of a step from the ref-format row left by the step before
// (the storm schedule is computed in double and rounded to R where it is used)
// TTIMER's ramp, ttimer.f:36-41
static double ttimer_bt(int itime, int ic19)
{
   double bt = 1.;
   if (itime < ic19) {
      double time = itime, ang = time / ic19;
      bt = .5 * (1. - cos(ang * 3.14159265358979323846));
   }
   return bt;
}

template <class R>
static void step_params(const slg_basin *B, const double *r, int tv, R *P)
{
   int itime = (int) r[0] + 1, ic19 = (int) r[27];
   const double bt = ttimer_bt(itime, ic19);
   for (int k = 0; k < SL_NPAR; k++) P[k] = R(0);
   P[P_BT] = (R) bt;
   P[P_DELT] = (R) r[2];
   P[P_X1B4] = (R) r[24];
   P[P_X1B5] = (R) r[25];
   P[P_X1B6] = (R) r[26];
   P[P_C1] = (R) r[3];
   P[P_C2] = (R) r[4];
   P[P_AX] = (R) r[5];
   P[P_AY] = (R) r[6];
   P[P_X4] = (R) r[40 + 3];
   P[P_X9] = (R) r[40 + 8];
   P[P_X10] = (R) r[40 + 9];
   P[P_X15] = (R) r[40 + 14];
   P[P_X20] = (R) r[40 + 19];
   P[P_X21] = (R) r[40 + 20];
   P[P_X23] = (R) r[40 + 22];
   P[P_X24] = (R) r[40 + 23];
   P[P_W0] = (R) r[15];
   P[P_W1] = (R) r[16];
   P[P_SEA] = (R) r[28];
   P[P_X1B1] = (R) r[31];
   P[P_X1B2] = (R) r[32];
   int fl = 0;
   if (B->ismpt2g && itime > 5 && B->incsm != 1) fl |= SF_SMPT;
   if ((itime + 1) % B->incsm == 0 || B->h.ebsn1 == '+') fl |= SF_FILT;
   // small integers, exact in either precision
   P[P_ITIME] = (R) itime;
   P[P_TVER] = (R) tv;
   P[P_FLAGS] = (R) fl;
}

// --------------------------------------------------------------------------
// batch

// A track given in memory as the rows INPUTN2 would read from a file
// (inputn2.f:560-605; PressureType 0: the deficit itself), with its header.
extern "C" int sl_track_from_mem(const slg_track *m, sl_trk *t, char e[256])
{
   sl_trk_init(t);
   e[0] = 0;
   if (!m || m->npts < 0 || (m->npts > 0 && (!m->hour || !m->lat || !m->lon || !m->delp || !m->rmw))) {
      snprintf(e, 256, "track: rows missing");
      return -1;
   }
   t->date_yr = m->year;
   t->date_mo = m->month;
   t->date_dy = m->day;
   t->date_hr = m->hour_of_day;
   t->date_mn = m->minute;
   t->dhr = m->date_hour;
   t->land = m->landfall_hour;
   t->beg = m->begin_hour;
   t->end = m->end_hour;
   t->ocean = m->ocean_datum;
   t->lake = m->lake_datum;
   t->windf = m->wind_factor;
   t->rmwf = m->rmw_factor;
   t->spline = m->spline != 0;
   for (int k = 0; k < m->npts; k++) {
      if ((int) m->hour[k] == 999) break;   // the file's terminator row
      const int rc = sl_trk_row(t, m->hour[k], m->lat[k], m->lon[k], m->delp[k], m->rmw[k], e);
      if (rc < 0) return -1;
      if (rc > 0) break;
   }
   return 0;
}

static int batch_open(slg_basin *B, int nstorm, const char *const *trk, const slg_track *tracks, int precision,
                      int flags, slg_batch **out, char *err, int errlen)
{
   *out = nullptr;
   if (!B || (!trk && !tracks)) {
      seterr(err, errlen, "%s", "a basin and nstorm tracks");
      return SLG_EINPUT;
   }
   if (flags & ~SLG_SCHED_HOST) {
      seterr(err, errlen, "%s", "unknown batch flags");
      return SLG_EINPUT;
   }
   if (nstorm < 1 || nstorm > SLG_MAX_STORMS) {
      seterr(err, errlen, "nstorm must be 1 .. %s%d (storms index a launch grid dimension)", "", SLG_MAX_STORMS);
      return SLG_EINPUT;
   }
   if (precision != 32 && precision != 64) {
      seterr(err, errlen, "%s", "precision must be 32 or 64");
      return SLG_EINPUT;
   }
   slg_batch *X = new slg_batch();
   X->b = B;
   X->prec = precision;
   X->nstorm = nstorm;
   X->sc.resize(nstorm);
   X->maxsteps = 0;
   X->dmem = nullptr;
   // storm schedules are independent: compute them on up to 16 host threads
   {
      const int nt = std::max(1, std::min<int>({16, nstorm, (int) std::thread::hardware_concurrency()}));
      std::vector<int> rcs(nstorm, 0);
      std::vector<std::string> msgs(nstorm);
      std::vector<std::thread> pool;
      for (int t = 0; t < nt; t++)
         pool.emplace_back([&, t]() {
            char e[512];
            for (int s = t; s < nstorm; s += nt) {
               e[0] = 0;
               rcs[s] = sl_make_schedule(B, trk ? trk[s] : nullptr, X->sc[s], nstorm <= 8, precision, e,
                                         sizeof e, (flags & SLG_SCHED_HOST) != 0, tracks ? &tracks[s] : nullptr);
               if (rcs[s]) msgs[s] = e;
            }
         });
      for (auto &th : pool) th.join();
      for (int s = 0; s < nstorm; s++) {
         if (rcs[s]) {
            seterr(err, errlen, "%s", msgs[s].c_str());
            delete X;
            return rcs[s];
         }
         if (X->sc[s].nsteps > X->maxsteps) X->maxsteps = X->sc[s].nsteps;
      }
   }
   // initial water (INTLHT) of every run: kept in double on the host with the
   // host schedule, else computed on the device at upload (k_intlht)
   if (flags & SLG_SCHED_HOST) {
      const size_t np = (size_t) B->h.imxb * B->h.jmxb;
      X->hb0.resize((size_t) nstorm * np);
      for (int s = 0; s < nstorm; s++) sl_initial_water(&B->h, &X->sc[s].s0, &X->hb0[(size_t) s * np]);
   }
   X->step = 0;
   X->parity = 0;
   X->path = SLG_PATH_FUSED;
   X->tile_rows = 16;
   X->hbs_buf = -1;
   X->last_par.assign(nstorm, 0);
   *out = X;
   return 0;
}

extern "C" int slg_batch_open_opts(slg_basin *B, int nstorm, const char *const *trk, int precision, int flags,
                                   slg_batch **out, char *err, int errlen)
{
   return batch_open(B, nstorm, trk, nullptr, precision, flags, out, err, errlen);
}

extern "C" int slg_batch_open_tracks(slg_basin *B, int nstorm, const slg_track *tracks, int precision, int flags,
                                     slg_batch **out, char *err, int errlen)
{
   return batch_open(B, nstorm, nullptr, tracks, precision, flags, out, err, errlen);
}

extern "C" int slg_batch_open_prec(slg_basin *B, int nstorm, const char *const *trk, int precision,
                                   slg_batch **out, char *err, int errlen)
{
   return batch_open(B, nstorm, trk, nullptr, precision, 0, out, err, errlen);
}

extern "C" int slg_batch_open(slg_basin *B, int nstorm, const char *const *trk, slg_batch **out,
                              char *err, int errlen)
{
   return slg_batch_open_prec(B, nstorm, trk, 32, out, err, errlen);
}

extern "C" int slg_batch_precision(const slg_batch *X)
{
   return X->prec;
}

static void drop_graph(slg_batch *X);

extern "C" int slg_batch_set_path(slg_batch *X, int path, int tile_rows)
{
   if ((path != SLG_PATH_FUSED && path != SLG_PATH_STAGED) ||
       (tile_rows != 8 && tile_rows != 16 && tile_rows != 32))
      return SLG_EINPUT;
   X->path = path;
   X->tile_rows = tile_rows;
   drop_graph(X);
   return 0;
}

// --------------------------------------------------------------------------
// MEOW track sets

static int meow_check(const slg_basin *B, const slg_meow *m, int cap, char *err, int errlen)
{
   if (!B || !m || cap < 1) {
      seterr(err, errlen, "%s", "a basin, MEOW parameters and room for at least one track");
      return SLG_EINPUT;
   }
   if (!B->h.proj.have) {
      seterr(err, errlen, "%s", "MEOW seeds need the basin's ?basins.dta line (slg_basin_open with bnt_line)");
      return SLG_EINPUT;
   }
   if (!(m->apart > 0.) || !(m->r34 >= 0.) || (m->xyflag != 0 && m->xyflag != 1) || m->num == 0 ||
       m->num < -1) {
      seterr(err, errlen, "%s", "MEOW: apart > 0, r34 >= 0, xyflag 0 or 1, num -1 or positive");
      return SLG_EINPUT;
   }
   return 0;
}

extern "C" int slg_meow_seeds(const slg_basin *B, const slg_meow *m, int cap, int *idx, double *lat, double *lon,
                              char *err, int errlen)
{
   int rc = meow_check(B, m, cap, err, errlen);
   if (rc) return rc;
   const int n = sl_meow_seeds(&B->h, m->xyflag, m->lat, m->lon, m->dir, m->apart, m->num, m->r34, cap, idx, lat,
                               lon);
   if (n < 0) {
      seterr(err, errlen, "MEOW: more seed points than the %s%d given room for", "", cap);
      return SLG_EINPUT;
   }
   return n;
}

extern "C" int slg_meow_tracks(const slg_basin *B, const slg_meow *m, int cap, slg_track *tracks, double *rows,
                               char *err, int errlen)
{
   int rc = meow_check(B, m, cap, err, errlen);
   if (rc) return rc;
   if (!tracks || !rows || !(m->speed >= 0.) || m->landfall_hour < 1 || m->landfall_hour > SLG_MEOW_HOURS ||
       m->begin_hour < 1 || m->end_hour > SLG_MEOW_HOURS - 1 || m->begin_hour >= m->end_hour) {
      seterr(err, errlen, "MEOW tracks: speed >= 0 and 1 <= begin < end < landfall ... hours within 1..%s%d", "",
             SLG_MEOW_HOURS);
      return SLG_EINPUT;
   }
   std::vector<int> idx(cap);
   std::vector<double> la(cap), lo(cap);
   const int n = sl_meow_seeds(&B->h, m->xyflag, m->lat, m->lon, m->dir, m->apart, m->num, m->r34, cap, idx.data(),
                               la.data(), lo.data());
   if (n < 0) {
      seterr(err, errlen, "MEOW: more tracks than the %s%d given room for", "", cap);
      return SLG_EINPUT;
   }
   const int H = SLG_MEOW_HOURS;
   for (int t = 0; t < n; t++) {
      double *h = rows + (size_t) t * 5 * H, *a = h + H, *o = a + H, *d = o + H, *r = d + H;
      for (int j = 1; j <= H; j++) {
         // the point speed * (j - landfall) miles along the heading from the
         // seed (CreateMEOWTrk, meow.tcl:327-331), longitude east positive
         double lat2, lon2;
         sl_earth_jump(la[t], -lo[t], m->speed * (j - m->landfall_hour), m->dir, &lat2, &lon2);
         h[j - 1] = j;
         a[j - 1] = lat2;
         o[j - 1] = lon2;
         d[j - 1] = m->delp;
         r[j - 1] = m->rmw;
      }
      slg_track &T = tracks[t];
      memset(&T, 0, sizeof T);
      T.year = 2000;
      T.month = 1;
      T.day = 1;
      T.date_hour = 1;
      T.landfall_hour = m->landfall_hour;
      T.begin_hour = m->begin_hour;
      T.end_hour = m->end_hour;
      T.ocean_datum = m->ocean_datum;
      T.lake_datum = m->lake_datum;
      T.wind_factor = 1.;
      T.rmw_factor = 1.;
      T.spline = 0;
      T.npts = H;
      T.hour = h;
      T.lat = a;
      T.lon = o;
      T.delp = d;
      T.rmw = r;
   }
   return n;
}

template <class R>
static int batch_schedule(const slg_batch *X, int storm, int from_device, double *dst, char *err, int errlen)
{
   const storm_sched &S = X->sc[storm];
   const size_t n = (size_t) S.nsteps * SL_NPAR;
   std::vector<R> v(n);
   if (from_device) {
      const SlRunDevT<R> &D = X->prec == 64 ? *(const SlRunDevT<R> *) &X->d64 : *(const SlRunDevT<R> *) &X->d32;
      HIPOK(hipMemcpy(v.data(), D.params + (size_t) storm * D.maxsteps * SL_NPAR, n * sizeof(R),
                      hipMemcpyDeviceToHost));
   } else {
      if (!S.hostp) {
         seterr(err, errlen, "%s", "the host keeps no per-step values (open with SLG_SCHED_HOST)");
         return SLG_EINPUT;
      }
      memcpy(v.data(), sched_params(S, R()).data(), n * sizeof(R));
   }
   for (size_t k = 0; k < n; k++) dst[k] = (double) v[k];
   return 0;
}

extern "C" int slg_batch_schedule(const slg_batch *X, int storm, int from_device, double *dst, char *err,
                                  int errlen)
{
   if (!X || storm < 0 || storm >= X->nstorm || !dst || (from_device && !X->dmem)) {
      seterr(err, errlen, "%s", "an uploaded batch (from_device) and a storm of it");
      return SLG_EINPUT;
   }
   return X->prec == 64 ? batch_schedule<double>(X, storm, from_device, dst, err, errlen)
                        : batch_schedule<float>(X, storm, from_device, dst, err, errlen);
}

extern "C" int slg_batch_nsteps(const slg_batch *X, int storm)
{
   return X->sc[storm].nsteps;
}


template <class R>
static int batch_upload(slg_batch *X, char *err, int errlen)
{
   slg_basin *B = X->b;
   if (!img(B, R()).on_dev) {
      int rc = basin_upload<R>(B, err, errlen);
      if (rc) return rc;
   }
   const size_t rs = sizeof(R);
   const sl_basin *b = &B->h;
   const int P = B->pitch, NJ = b->jmxb, S = X->nstorm;
   const size_t fstride = (size_t) P * NJ, np = fstride * S;
   size_t ntab = 0;
   for (auto &sc : X->sc) ntab += sc.tab_step.size();
   size_t off[24], tot = 0;
   int n = 0;
   auto add = [&](size_t bytes) {
      off[n++] = tot;
      tot += (bytes + 255) & ~(size_t) 255;
   };
   add(np * rs); add(np * rs); add(np * rs); add(np * rs); add(np * rs);   // hb0 hb1 ub vb hmx
   add((size_t) S * P * rs);                                           // hsubtop
   add(np); add(np); add(np);                                          // pflag hclass uz
   add((size_t) S * SL_EVCAP * 4);                                     // evkey
   add((size_t) S * 4);                                                // evcnt
   add((size_t) S * 8 * 4);                                            // stat
   add((size_t) S * X->maxsteps * SL_NPAR * rs);                       // params
   add(ntab * SL_NTABV * 8 * rs);                                      // tables
   add((size_t) S * 4);                                                // tbase
   add((size_t) S * 4);                                                // nsteps
   add((size_t) S * 4);                                                // cur
   const BasinImg<R> &I = img(B, R());
   const size_t nl = (size_t) std::max(b->nsqrwc, 1), nq = (size_t) std::max(I.nsq, 1);
   add((size_t) S * nl * rs);                                          // flwsqr
   add((size_t) S * nl);                                               // f1dact
   add((size_t) S * nq * rs);                                          // hsq
   add((size_t) S * 5 * rs);                                           // chq
   const bool sides = has_sides(B);
   add(sides ? np * rs : 0);                                           // hb2
   char *d = nullptr;
   HIPOK(hipMalloc((void **) &d, tot));
   X->dmem = d;
   X->bytes = tot;
   SlRunDevT<R> &D = run(X, R());
   memset(&D, 0, sizeof D);
   D.nstorm = S;
   D.fstride = fstride;
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
   D.maxsteps = X->maxsteps;
   D.flwsqr = (R *) (d + off[17]);    // FLWSQR = 0 (intlht.f:190-191), F1DACT not 'T'
   D.f1dact = (uint8_t *) (d + off[18]);
   D.hsq = (R *) (d + off[19]);
   D.chq = (R *) (d + off[20]);       // Q = 0 (channl.f:73)
   D.hb2 = sides ? (R *) (d + off[21]) : nullptr;
   D.evcap = SL_EVCAP;
   D.vbudget = 0;
   D.sgroup = SL_SGROUP;
   HIPOK(hipMemset(d, 0, tot));
   // initial water (INTLHT) per storm: HB, HMX=HB, UB=VB=0 (from the host's,
   // or on the device below)
   const bool hostw = !X->hb0.empty();
   std::vector<R> f(hostw ? fstride : 0, R(0));
   std::vector<R> params(X->sc[0].hostp ? (size_t) S * X->maxsteps * SL_NPAR : 0, R(0));
   std::vector<R> tables;
   std::vector<int> tbase(S), nst(S);
   for (int s = 0; s < S; s++) {
      if (hostw) {
         const double *hb = &X->hb0[(size_t) s * b->imxb * NJ];
         for (int j = 1; j <= NJ; j++)
            for (int i = 1; i <= b->imxb; i++) f[(size_t) (j - 1) * P + (i - 1)] = (R) hb[PIX(b, i, j)];
         HIPOK(hipMemcpy(D.hb0 + (size_t) s * fstride, f.data(), fstride * rs, hipMemcpyHostToDevice));
         HIPOK(hipMemcpy(D.hb1 + (size_t) s * fstride, f.data(), fstride * rs, hipMemcpyHostToDevice));
         for (int j = 1; j <= NJ; j++)       // HMX = HB on squares only (intlht.f:172-177)
            for (int i = 1; i <= b->imxb; i++)
               if (i == b->imxb || j == NJ) f[(size_t) (j - 1) * P + (i - 1)] = R(0);
         HIPOK(hipMemcpy(D.hmx + (size_t) s * fstride, f.data(), fstride * rs, hipMemcpyHostToDevice));
      }
      const std::vector<R> &sp = sched_params(X->sc[s], R()), &st = sched_tables(X->sc[s], R());
      if (X->sc[s].hostp) memcpy(&params[(size_t) s * X->maxsteps * SL_NPAR], sp.data(), sp.size() * sizeof(R));
      tbase[s] = (int) (tables.size() / ((size_t) SL_NTABV * 8));
      tables.insert(tables.end(), st.begin(), st.end());
      nst[s] = X->sc[s].nsteps;
   }
   if (X->sc[0].hostp) HIPOK(hipMemcpy(d + off[12], params.data(), params.size() * rs, hipMemcpyHostToDevice));
   if (!X->sc[0].hostp || !hostw) {
      // the storm programs to the device, with (for the initial water) the
      // basin metrics ELPCL ELPDL COSL SINL, ZB and the squares in reach of
      // ocean winds, all in double; the per-step values and the initial water
      // are generated from them there
      std::vector<long> poff(S);
      size_t np2 = 0;
      for (int s = 0; s < S; s++) {
         poff[s] = (long) np2;
         np2 += X->sc[s].prog.size();
      }
      const int ni = b->imxb, nj = b->jmxb;
      const size_t nb = 2 * (size_t) ni + 2 * (size_t) nj + (size_t) ni * nj;
      std::vector<double> buf;
      buf.reserve(np2 + nb);
      for (int s = 0; s < S; s++) buf.insert(buf.end(), X->sc[s].prog.begin(), X->sc[s].prog.end());
      buf.insert(buf.end(), b->elpcl, b->elpcl + ni);
      buf.insert(buf.end(), b->elpdl, b->elpdl + ni);
      buf.insert(buf.end(), b->cosl, b->cosl + nj);
      buf.insert(buf.end(), b->sinl, b->sinl + nj);
      buf.insert(buf.end(), b->zb, b->zb + (size_t) ni * nj);
      std::vector<unsigned char> ocn((size_t) ni * nj, 0);   // how INTLHT fills each square
      for (int j = 1; j <= b->jmxb1; j++)
         for (int i = 1; i <= b->imxb1; i++) ocn[PIX(b, i, j)] = (unsigned char) sl_intlht_class(b, i, j);
      char *pd = nullptr;
      const size_t nbytes = buf.size() * sizeof(double) + (size_t) S * sizeof(long) + ocn.size();
      HIPOK(hipMalloc((void **) &pd, nbytes));
      long *offd = (long *) (pd + buf.size() * sizeof(double));
      unsigned char *ocnd = (unsigned char *) (offd + S);
      hipError_t he = hipMemcpy(pd, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice);
      if (he == hipSuccess) he = hipMemcpy(offd, poff.data(), (size_t) S * sizeof(long), hipMemcpyHostToDevice);
      if (he == hipSuccess) he = hipMemcpy(ocnd, ocn.data(), ocn.size(), hipMemcpyHostToDevice);
      int krc = he == hipSuccess ? 0 : -(int) he;
      if (!krc && !X->sc[0].hostp)
         krc = dev_sched((const double *) pd, offd, S, (R *) (d + off[12]), X->maxsteps, B->ismpt2g, B->incsm,
                         b->ebsn1 == '+', 0);
      if (!krc && !hostw)
         krc = dev_intlht((const double *) pd, offd, (const double *) pd + np2, ocnd, ni, nj, b->imxb1, b->jmxb1, P,
                          &D, 0);
      if (!krc) krc = -(int) hipDeviceSynchronize();
      (void) hipFree(pd);
      if (krc) {
         seterr(err, errlen, "%s", "device storm schedule / initial water failed");
         return SLG_EHIP;
      }
   }
   HIPOK(hipMemcpy(d + off[13], tables.data(), tables.size() * rs, hipMemcpyHostToDevice));
   HIPOK(hipMemcpy(d + off[14], tbase.data(), (size_t) S * 4, hipMemcpyHostToDevice));
   HIPOK(hipMemcpy(d + off[15], nst.data(), (size_t) S * 4, hipMemcpyHostToDevice));
   D.params = (const R *) (d + off[12]);
   D.tables = (const R *) (d + off[13]);
   D.tbase = (const int *) (d + off[14]);
   D.nsteps = (const int *) (d + off[15]);
   D.cur = (int *) (d + off[16]);   // zeroed with the rest
   HIPOK(hipStreamCreate(&X->st));
   X->step = 0;
   X->parity = 0;
   if (X->tide_mode) {
      // TIDESH of every run, then the tide's first call on the initial water
      // (RunInit, slosh2.c:1851-1868)
      const size_t nb = (size_t) std::max(b->nbcpt, 1) * S;
      R *t = nullptr;
      HIPOK(hipMalloc((void **) &t, nb * rs));
      D.tidesh = t;
      HIPOK(hipMemset(t, 0, nb * rs));
      const int krc = dev_tide(&I.dev, &D, &X->tdev, D.hb0, 0, 1, X->st);
      HIPOK(krc < 0 ? (hipError_t) -krc : (krc > 0 ? hipErrorInvalidValue : hipSuccess));
      HIPOK(hipStreamSynchronize(X->st));
      // every storm part's fields start from the same step-0 tide (slot 0)
      const SlTideDev &T = X->tdev;
      const size_t fg = (size_t) T.ngrp * 2 * T.ncell;
      for (int p = 1; p < SL_MAXSPLIT; p++)
         for (int g = 0; g < T.ngrp; g++)
            HIPOK(hipMemcpy(T.field + p * fg + (size_t) g * 2 * T.ncell, T.field + (size_t) g * 2 * T.ncell,
                            (size_t) T.ncell * 8, hipMemcpyDeviceToDevice));
   }
   return 0;
}

extern "C" int slg_batch_upload(slg_batch *X, char *err, int errlen)
{
   return X->prec == 64 ? batch_upload<double>(X, err, errlen) : batch_upload<float>(X, err, errlen);
}

// storms s0 .. s0+ns-1 of a batch as a batch of their own (same buffers)
template <class R>
static SlRunDevT<R> run_view(const SlRunDevT<R> &D, const SlBasinDevT<R> &B, int s0, int ns)
{
   SlRunDevT<R> V = D;
   const size_t s = (size_t) s0, fs = D.fstride;
   V.nstorm = ns;
   V.hb0 += s * fs;
   V.hb1 += s * fs;
   V.ub += s * fs;
   V.vb += s * fs;
   V.hmx += s * fs;
   V.hsubtop += s * (size_t) B.pitch;
   V.pflag += s * fs;
   V.hclass += s * fs;
   V.uz += s * fs;
   V.evkey += s * SL_EVCAP;
   V.evcnt += s;
   V.stat += s * 8;
   V.params += s * (size_t) D.maxsteps * SL_NPAR;
   V.tbase += s;
   V.nsteps += s;
   V.cur += s;
   V.flwsqr += s * (size_t) B.nsqrwc;
   V.f1dact += s * (size_t) B.nsqrwc;
   V.hsq += s * (size_t) B.nsq;
   V.chq += s * 5;
   if (V.hb2) V.hb2 += s * fs;
   if (V.tidesh) V.tidesh += s * (size_t) B.nbcpt;
   return V;
}

// Part p of N (storms p*n/N .. (p+1)*n/N - 1) steps on its own stream.  On the
// first step of a call, part p > 0 starts when part p-1's cascade replay of that
// step has finished: from then on the parts run staggered, each one's replay
// beside the others' stencil kernels.  (Part 1 also starts after everything
// queued on the batch stream before.)
template <class R>
static int launch_split(slg_batch *X, const SlBasinDevT<R> &B, const SlRunDevT<R> &D, int smpt, int filt,
                        hipEvent_t *ev, int dyn, bool first)
{
   const int N = X->nsplit, n = X->nstorm;
   hipEvent_t *prev = nullptr;
   int r = 0;
   for (int p = 0; p < N; p++) {
      const int s0 = (int) ((long) n * p / N), s1 = (int) ((long) n * (p + 1) / N);
      const SlRunDevT<R> V = run_view(D, B, s0, s1 - s0);
      hipStream_t st = p ? X->stx[p - 1] : X->st;
      hipEvent_t *evp = (p == 0 && ev) ? ev : (first ? X->ev_phase[p] : nullptr);
      if (p && first && hipStreamWaitEvent(st, prev[2], 0) != hipSuccess) return -(int) hipErrorUnknown;
      r = dev_launch(&B, &V, smpt, filt, dyn, X->parity, X->path, X->tile_rows, st, evp);
      if (r < 0) return r;
      if (X->tide_mode) {   // each part sums its own fields: the parts run staggered
         SlTideDev Tp = X->tdev;
         Tp.sgrp += s0;
         Tp.field += (size_t) p * Tp.ngrp * 2 * Tp.ncell;
         const int t = dev_tide(&B, &V, &Tp, r ? V.hb1 : V.hb0, X->step + 1, 0, st);
         if (t < 0) return t;
      }
      if (X->hbmax || X->hbmax2) {
         const size_t o = (size_t) s0 * D.fstride;
         const int m = dev_maxstat(&B, &V, r ? V.hb1 : V.hb0, img(X->b, R()).zbd, X->hbmax ? X->hbmax + o : nullptr,
                                   X->hbmax2 ? X->hbmax2 + o : nullptr, 0, st);
         if (m < 0) return m;
      }
      prev = evp;
   }
   return r;
}

// the whole batch on its stream; then the windowed maxima when they are on
template <class R>
static int launch_whole(slg_batch *X, const SlBasinDevT<R> &B, const SlRunDevT<R> &D, int smpt, int filt,
                        hipEvent_t *ev, int dyn)
{
   const int r = dev_launch(&B, &D, smpt, filt, dyn, X->parity, X->path, X->tile_rows, X->st, ev);
   if (r < 0) return r;
   if (X->tide_mode) {   // the driver's tide after the step, before the maxima see HB
      const int t = dev_tide(&B, &D, &X->tdev, r ? D.hb1 : D.hb0, X->step + 1, 0, X->st);
      if (t < 0) return t;
   }
   if (!(X->hbmax || X->hbmax2)) return r;
   const int m = dev_maxstat(&B, &D, r ? D.hb1 : D.hb0, img(X->b, R()).zbd, X->hbmax, X->hbmax2, 0, X->st);
   return m < 0 ? m : r;
}

static int launch_one(slg_batch *X, int smpt, int filt, hipEvent_t *ev, int dyn = 0, bool first = false)
{
   if (X->nsplit > 1 && X->path == SLG_PATH_FUSED) {
      if (X->prec == 64) return launch_split(X, X->b->i64.dev, X->d64, smpt, filt, ev, dyn, first);
      return launch_split(X, X->b->i32.dev, X->d32, smpt, filt, ev, dyn, first);
   }
   if (X->prec == 64) return launch_whole(X, X->b->i64.dev, X->d64, smpt, filt, ev, dyn);
   return launch_whole(X, X->b->i32.dev, X->d32, smpt, filt, ev, dyn);
}

// the batch stream waits for the other parts' steps: callers keep seeing one
// stream
static int join_streams(slg_batch *X)
{
   for (int p = 1; p < X->nsplit; p++) {
      if (hipEventRecord(X->ev_join[p - 1], X->stx[p - 1]) != hipSuccess) return -1;
      if (hipStreamWaitEvent(X->st, X->ev_join[p - 1], 0) != hipSuccess) return -1;
   }
   return 0;
}

extern "C" int slg_batch_set_streams(slg_batch *X, int nstreams)
{
   if (nstreams < 1 || nstreams > SL_MAXSPLIT || nstreams > X->nstorm) return SLG_EINPUT;
   if (nstreams > 1 && !X->dmem) return SLG_EINPUT;   // after slg_batch_upload
   for (int p = 1; p < nstreams; p++) {
      if (X->stx[p - 1]) continue;
      if (hipStreamCreate(&X->stx[p - 1]) != hipSuccess ||
          hipEventCreateWithFlags(&X->ev_join[p - 1], hipEventDisableTiming) != hipSuccess)
         return SLG_EHIP;
   }
   for (int p = 0; p < nstreams; p++)
      for (auto &e : X->ev_phase[p])
         if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return SLG_EHIP;
   // steps already queued on the batch stream come first
   if (join_streams(X)) return SLG_EHIP;
   if (X->tide_mode) {
      // the tide fields of every part as part 0 holds them (all parts are at
      // the same step; a part that did not step since holds older ones)
      const SlTideDev &T = X->tdev;
      const size_t fg = (size_t) T.ngrp * 2 * T.ncell;
      if (hipStreamSynchronize(X->st) != hipSuccess) return SLG_EHIP;
      for (int p = 1; p < SL_MAXSPLIT; p++)
         if (hipMemcpy(T.field + p * fg, T.field, fg * 8, hipMemcpyDeviceToDevice) != hipSuccess) return SLG_EHIP;
   }
   X->nsplit = nstreams;
   return 0;
}

extern "C" int slg_batch_streams(const slg_batch *X)
{
   return X->nsplit;
}

static void drop_graph(slg_batch *X)
{
   if (X->gexec) (void) hipGraphExecDestroy(X->gexec);
   if (X->graph) (void) hipGraphDestroy(X->graph);
   X->gexec = nullptr;
   X->graph = nullptr;
}

// Capture graph_steps fused steps in which every storm's own step flags decide
// (so one graph serves every step of the run) and instantiate it.
static int build_graph(slg_batch *X, char *err, int errlen)
{
   drop_graph(X);
   HIPOK(hipStreamBeginCapture(X->st, hipStreamCaptureModeThreadLocal));
   int rc = 0;
   for (int k = 0; k < X->graph_steps && rc >= 0; k++) rc = launch_one(X, 1, 1, nullptr, 1);
   hipGraph_t g = nullptr;
   HIPOK(hipStreamEndCapture(X->st, &g));
   if (rc < 0) {
      (void) hipGraphDestroy(g);
      seterr(err, errlen, "kernel launch failed in capture (%s%d)", "", -rc);
      return SLG_EHIP;
   }
   X->graph = g;
   HIPOK(hipGraphInstantiate(&X->gexec, g, nullptr, nullptr, 0));
   return 0;
}

// Replay the step loop as hipGraphs of n steps each (fused path; 0 = off).
extern "C" int slg_batch_set_graph(slg_batch *X, int steps_per_graph)
{
   if (steps_per_graph < 0) return SLG_EINPUT;
   drop_graph(X);
   X->graph_steps = steps_per_graph;
   return 0;
}

// Storms whose tiles follow each other in the fused kernels' workgroup order
// (1 .. 64; fields are identical for any value).
extern "C" int slg_batch_set_storm_group(slg_batch *X, int sg)
{
   if (sg < 1 || sg > 64 || !X->dmem) return SLG_EINPUT;
   X->d32.sgroup = X->d64.sgroup = sg;
   drop_graph(X);                     // captured launches hold the old value
   return 0;
}

extern "C" int slg_batch_storm_group(const slg_batch *X)
{
   return X->prec == 64 ? X->d64.sgroup : X->d32.sgroup;
}

// Lower the cascade replay's limits, for checking its two fallbacks on
// ordinary runs: evcap drying squares listed per storm (1 .. SL_EVCAP) before
// the walk visits every square, and a visit budget (0: the default bound).
extern "C" int slg_batch_set_replay_limits(slg_batch *X, int evcap, long budget)
{
   if (evcap < 1 || evcap > SL_EVCAP || budget < 0) return SLG_EINPUT;
   if (!X->dmem) return SLG_EINPUT;   // after slg_batch_upload
   X->d32.evcap = X->d64.evcap = evcap;
   X->d32.vbudget = X->d64.vbudget = budget;
   drop_graph(X);                     // captured launches hold the old values
   return 0;
}

// The cascade replay stops a run that exhausts its visit budget and flags it
// (stat[7]); such a run's fields are incomplete and are not handed out.
template <class R>
static int check_runs(slg_batch *X, int s0, int ns, char *err, int errlen)
{
   int st[64 * 8];   // in chunks of 64 storms: no host or device allocation
   HIPOK(hipStreamSynchronize(X->st));
   for (int c = 0; c < ns; c += 64) {
      const int m = std::min(64, ns - c);
      HIPOK(hipMemcpy(st, run(X, R()).stat + (size_t) (s0 + c) * 8, (size_t) m * 32, hipMemcpyDeviceToHost));
      for (int s = 0; s < m; s++)
         if (st[s * 8 + 7]) {
            seterr(err, errlen, "run %s%d: the drying-cascade replay ran out of visits; its fields are incomplete",
                   "", s0 + c + s);
            return SLG_ELIMIT;
         }
   }
   return 0;
}

static int check_runs_any(slg_batch *X, int s0, int ns, char *err, int errlen)
{
   return X->prec == 64 ? check_runs<double>(X, s0, ns, err, errlen) : check_runs<float>(X, s0, ns, err, errlen);
}

// after a lockstep step: where each storm still running left its HB
static void note_step(slg_batch *X, int step, int rc)
{
   for (int s = 0; s < X->nstorm; s++)
      if (X->sc[s].nsteps >= step) X->last_par[s] = rc;
}

// advance every storm of the batch by n steps (stops at the shortest run)
extern "C" int slg_batch_step(slg_batch *X, int n, char *err, int errlen)
{
   const slg_basin *B = X->b;
   const int lim = X->maxsteps;   // storms with shorter runs stop at their own end
   if (X->graph_steps > 0 && X->path == SLG_PATH_FUSED && X->nsplit == 1 && !X->tide_mode) {
      // whole graphs while they fit, then single steps with the same kernels
      const int step0 = X->step;
      if (!X->gexec) {
         int rc = build_graph(X, err, errlen);
         if (rc) return rc;
      }
      while (n >= X->graph_steps && X->step + X->graph_steps <= lim) {
         HIPOK(hipGraphLaunch(X->gexec, X->st));
         X->step += X->graph_steps;
         n -= X->graph_steps;
      }
      for (; n > 0 && X->step < lim; n--) {
         int rc = launch_one(X, 1, 1, nullptr, 1);
         if (rc < 0) {
            seterr(err, errlen, "kernel launch failed (%s%d)", "", -rc);
            return SLG_EHIP;
         }
         X->step++;
      }
      // a fused step leaves HB in the buffer it started from (X->parity, which
      // an earlier staged step, e.g. an autotuner trial, may have set to 1):
      // every storm that stepped in this call ends there
      X->hbs_buf = -1;
      for (int s = 0; s < X->nstorm; s++)
         if (X->sc[s].nsteps > step0) X->last_par[s] = X->parity;
      return 0;
   }
   for (int k = 0; k < n && X->step < lim; k++) {
      int step = X->step + 1;
      int itime = step;
      int smpt = (B->ismpt2g && itime > 5 && B->incsm != 1);
      int filt = ((itime + 1) % B->incsm == 0 || B->h.ebsn1 == '+');
      int rc = launch_one(X, smpt, filt, nullptr, 0, k == 0);
      if (rc < 0) {
         join_streams(X);
         seterr(err, errlen, "kernel launch failed (%s%d)", "", -rc);
         return SLG_EHIP;
      }
      // staged: FLT2G/HMXSV leave HB in the entering buffer, FILTER writes the other
      X->hbs_buf = !filt ? rc : (X->path == SLG_PATH_STAGED ? X->parity : -1);
      X->parity = rc;
      X->step = step;
      note_step(X, step, rc);
   }
   if (join_streams(X)) {
      seterr(err, errlen, "%s", "stream join failed");
      return SLG_EHIP;
   }
   return 0;
}

// As slg_batch_step, with HIP events recorded on the batch's stream around every
// kernel; ms[k] accumulates the time of kernel slot k (SLG_K_* order) and
// launches[k] the number of launches timed.
extern "C" int slg_batch_step_timed(slg_batch *X, int n, double *ms, int *launches, char *err,
                                    int errlen)
{
   // events around every kernel of a ring of RING steps; the host waits only
   // when the ring is full, so timing does not serialise the step loop
   const int RING = 64, NE = SLG_NKERN + 1;
   const slg_basin *B = X->b;
   std::vector<hipEvent_t> ev((size_t) RING * NE);
   struct Slot {
      int smpt, filt;
   };
   std::vector<Slot> ran(RING);
   for (auto &e : ev) HIPOK(hipEventCreate(&e));
   int pending = 0, rc_all = 0;
   auto drain = [&]() -> int {
      if (!pending) return 0;
      if (hipEventSynchronize(ev[(size_t) (pending - 1) * NE + (X->path == SLG_PATH_FUSED ? 3 : 8)]) != hipSuccess)
         return SLG_EHIP;
      for (int r = 0; r < pending; r++) {
         hipEvent_t *E = &ev[(size_t) r * NE];
         float t = 0.f;
         if (X->path == SLG_PATH_FUSED) {
            const int slot[3] = {SLG_K_CONTTY, SLG_K_REPLAY, SLG_K_STEPB};
            for (int q = 0; q < 3; q++) {
               if (hipEventElapsedTime(&t, E[q], E[q + 1]) != hipSuccess) return SLG_EHIP;
               ms[slot[q]] += t;
               launches[slot[q]]++;
            }
            continue;
         }
         const int on[8] = {1, 1, ran[r].smpt, ran[r].smpt, 1, 1, ran[r].filt, 1};
         for (int q = 0; q < 8; q++) {
            if (!on[q]) continue;
            if (hipEventElapsedTime(&t, E[q], E[q + 1]) != hipSuccess) return SLG_EHIP;
            ms[q] += t;
            launches[q]++;
         }
      }
      pending = 0;
      return 0;
   };
   const int lim = X->maxsteps;
   for (int k = 0; k < n && X->step < lim; k++) {
      int step = X->step + 1, itime = step;
      int smpt = (B->ismpt2g && itime > 5 && B->incsm != 1);
      int filt = ((itime + 1) % B->incsm == 0 || B->h.ebsn1 == '+');
      int rc = launch_one(X, smpt, filt, &ev[(size_t) pending * NE], 0, k == 0);
      if (rc < 0) {
         seterr(err, errlen, "kernel launch failed (%s%d)", "", -rc);
         rc_all = SLG_EHIP;
         break;
      }
      ran[pending] = {smpt, filt};
      pending++;
      X->hbs_buf = !filt ? rc : (X->path == SLG_PATH_STAGED ? X->parity : -1);
      X->parity = rc;
      X->step = step;
      note_step(X, step, rc);
      if (pending == RING && drain()) {
         seterr(err, errlen, "%s", "event timing failed");
         rc_all = SLG_EHIP;
         break;
      }
   }
   if (!rc_all && drain()) {
      seterr(err, errlen, "%s", "event timing failed");
      rc_all = SLG_EHIP;
   }
   join_streams(X);
   (void) hipStreamSynchronize(X->st);
   for (auto &e : ev) (void) hipEventDestroy(e);
   return rc_all;
}

extern "C" int slg_batch_sync(slg_batch *X, char *err, int errlen)
{
   HIPOK(hipStreamSynchronize(X->st));
   return 0;
}

extern "C" int slg_batch_current_step(const slg_batch *X)
{
   return X->step;
}

// Static basin field as double [jmxb][imxb] ("ZB", "ZBM", "YLT").
extern "C" int slg_basin_field(const slg_basin *B, const char *name, double *dst)
{
   const sl_basin *b = &B->h;
   const double *src = !strcmp(name, "ZB") ? b->zb : !strcmp(name, "ZBM") ? b->zbm
                     : !strcmp(name, "YLT") ? b->ylt : nullptr;
   if (!src) return SLG_EINPUT;
   memcpy(dst, src, sizeof(double) * (size_t) b->imxb * b->jmxb);
   return 0;
}

// Values TMSTPM hands back to the C driver after step n (tmstpm.f:59-70):
// ITIME, MHALT, DELT, storm latitude and longitude (XY2LLX of (C1+AX, C2+AY)),
// DELP (YDELP), size (YC24), speed (C21), direction (AZMTH - C22 in degrees).
// Step 0 gives the state after initialisation.  Kept for batches of <= 8 runs.
// The BDRYHT squares (DEPSFC's ISH/JSH list, depsfc.f:376-390), 1-based, up to
// cap of them; returns their number.
extern "C" int slg_basin_boundary(const slg_basin *B, int *ish, int *jsh, int cap)
{
   const sl_basin *b = &B->h;
   for (int n = 0; n < b->nbcpt && n < cap; n++) {
      ish[n] = b->ish[n];
      jsh[n] = b->jsh[n];
   }
   return b->nbcpt;
}

// ISSUBGRID (issubgrid.f:14-33): 1 when square (i,j) holds a 1-D flow side
// whose cut width (DELCUT) or cut interior/exterior length (CUTLI, CUTLE) is
// below one cell, else 0.
extern "C" int slg_basin_is_subgrid(const slg_basin *B, int i, int j)
{
   const sl_basin *b = &B->h;
   for (int l = 1; l <= b->nsqrwc; l++) {
      if (b->isqr[l - 1] != i || b->jsqr[l - 1] != j) continue;
      if (l <= b->nsqrw) {
         if (b->delcut[l - 1] < 1.) return 1;
      } else {
         const int ll = l - b->nsqrw;
         if (b->cutli[ll - 1] < 1. || b->cutle[ll - 1] < 1.) return 1;
      }
   }
   return 0;
}

// Overwrite a storm's HB (double [jmxb][imxb]) where its next step reads it:
// for a driver that changes the water level between steps (the tide coupling,
// slosh2.c:813-1027, adds the tide to every wet square).
template <class R>
static int batch_put_hb(slg_batch *X, int storm, const double *src, char *err, int errlen)
{
   const sl_basin *b = &X->b->h;
   SlRunDevT<R> &D = run(X, R());
   const int P = X->b->pitch;
   std::vector<R> f(D.fstride, R(0));
   for (int j = 1; j <= b->jmxb; j++)
      for (int i = 1; i <= b->imxb; i++) f[(size_t) (j - 1) * P + (i - 1)] = (R) src[PIX(b, i, j)];
   R *dst = (X->last_par[storm] ? D.hb1 : D.hb0) + (size_t) storm * D.fstride;
   HIPOK(hipStreamSynchronize(X->st));
   HIPOK(hipMemcpy(dst, f.data(), D.fstride * sizeof(R), hipMemcpyHostToDevice));
   return 0;
}

extern "C" int slg_batch_put(slg_batch *X, int storm, const char *name, const double *src, char *err, int errlen)
{
   if (!X->dmem || storm < 0 || storm >= X->nstorm || strcmp(name, "HB")) {
      seterr(err, errlen, "%s", "HB of an uploaded storm only");
      return SLG_EINPUT;
   }
   if (X->last_par[storm] != X->parity && X->sc[storm].nsteps > X->step) {
      seterr(err, errlen, "%s", "storm not at the batch's step");
      return SLG_EINPUT;
   }
   return X->prec == 64 ? batch_put_hb<double>(X, storm, src, err, errlen)
                        : batch_put_hb<float>(X, storm, src, err, errlen);
}

// The tide height BDRYHT adds at each boundary square (COMMON /BCPTS/ TIDESH,
// bdryht.f:85), n = the basin's number of BDRYHT squares, from the storm's next
// step on.  The first call gives every storm of the batch a row of zeros.
template <class R>
static int batch_tide(slg_batch *X, int storm, const double *tide, int n, char *err, int errlen)
{
   SlRunDevT<R> &D = run(X, R());
   const size_t nb = (size_t) std::max(n, 1);
   HIPOK(hipStreamSynchronize(X->st));
   if (!D.tidesh) {
      R *t = nullptr;
      HIPOK(hipMalloc((void **) &t, nb * X->nstorm * sizeof(R)));
      HIPOK(hipMemset(t, 0, nb * X->nstorm * sizeof(R)));
      D.tidesh = t;
      drop_graph(X);   // captured launches hold the old pointer
   }
   std::vector<R> v(nb, R(0));
   for (int k = 0; k < n; k++) v[k] = (R) tide[k];
   HIPOK(hipMemcpy(D.tidesh + (size_t) storm * n, v.data(), (size_t) n * sizeof(R), hipMemcpyHostToDevice));
   return 0;
}

extern "C" int slg_batch_set_boundary_tide(slg_batch *X, int storm, const double *tide, int n, char *err,
                                           int errlen)
{
   if (!X->dmem || storm < 0 || storm >= X->nstorm || n != X->b->h.nbcpt) {
      seterr(err, errlen, "%s", "an uploaded storm and one value per BDRYHT square");
      return SLG_EINPUT;
   }
   return X->prec == 64 ? batch_tide<double>(X, storm, tide, n, err, errlen)
                        : batch_tide<float>(X, storm, tide, n, err, errlen);
}

// ---------------------------------------------------------------- gridded tide
// DELT after each STMVAL call n = 1 .. nsteps of a run (what TMSTPM hands the
// driver, which moves its model clock by it; slosh2.c:1553), from the storm
// program: the state after SETCMP, changed by the events.
static void run_delts(const storm_sched &S, std::vector<double> &dt)
{
   const std::vector<double> &g = S.prog;
   const int n = (int) g[0], ic19 = (int) g[1], nev = (int) g[2];
   const double *ev = g.data() + SQ_BT + (ic19 > 1 ? ic19 - 1 : 0);
   double d = g[SQ_HDR + SQ_DELT];
   dt.assign((size_t) n + 1, 0.);
   for (int k = 1, e = 0; k <= n; k++) {
      for (; e < nev && (int) ev[(size_t) e * SQ_EV] <= k; e++) d = ev[(size_t) e * SQ_EV + 1 + SQ_DELT];
      dt[k] = d;
   }
}

// The gridded tide of the reference's -f_tide V2 modes, applied on the device
// between steps as its driver applies it (RunInit, slosh2.c:1789-1869;
// SetTideUsableFlag, :717-744; CalcTideGrid2, :813-1027; RunLoopStep,
// :1553-1582).  Each run's model clock starts at its start date (INITAL's
// date, Clock_ScanDate + hour + minute) and moves by DELT every step; the
// tide tables are the run's year's.  The cos / sin of every constituent's
// angle at every step are taken here on the host (libm, as the driver takes
// them), once per group of runs with the same year and clock.
extern "C" int slg_batch_set_tide(slg_batch *X, const char *ft03, const char *bhc, const char *adj, int mode,
                                  int thresh, char *err, int errlen)
{
   if (X->dmem || X->tide_mode) {
      seterr(err, errlen, "%s", "the tide is set once, before the upload");
      return SLG_EINPUT;
   }
   if (mode != 2 && mode != 21 && mode != 22 && mode != 24) {
      seterr(err, errlen, "tide mode %s%d: 2, 21, 22 or 24 (-f_tide V2, V2.1, V2.2, V2.4)", "", mode);
      return SLG_EINPUT;
   }
   static_assert(SL_NTIDE == SL_NUMT, "constituent count");
   const sl_basin *b = &X->b->h;
   const int S = X->nstorm;
   char e[256];
   sl_tide t;
   if (sl_tide_open(&t, bhc, adj, b->imxb, b->jmxb, e)) {
      seterr(err, errlen, "%s", e);
      return SLG_EINPUT;
   }
   // the hsf basins: RunLoopStep calls CalcTideGrid2 only once the clock has
   // reached tideClock, which PerformRun then moves 6 minutes on
   // (slosh2.c:1565-1577, 2388, 2583-2586)
   bool six = false;
   for (const char *h : {"hsf1", "hsfa", "hsfb", "hsfc", "hsfd", "hsfe", "hsff"})
      six = six || (!strncmp(t.abbrev, h, 4) && (t.abbrev[4] == 0 || t.abbrev[4] == ' '));
   const int ni = t.ni, nc = t.ncell;
   // clock groups: key = year, then the clock at every step
   std::map<std::vector<double>, int> gidx;
   std::vector<int> sgrp(S), years, gyear, gmax;
   std::vector<std::vector<double>> gclk;
   std::vector<double> dt, key;
   for (int s = 0; s < S; s++) {
      int d[5];
      slg_batch_start_date(X, s, d);
      double c = sl_clock_scan_date(d[0], d[1], d[2]);
      c = c + (d[3] * 3600.);
      c = c + (d[4] * 60);
      run_delts(X->sc[s], dt);
      key.assign(1, (double) d[0]);
      key.push_back(c);
      for (size_t k = 1; k < dt.size(); k++) {
         c = c + dt[k];
         key.push_back(c);
      }
      auto it = gidx.find(key);
      if (it == gidx.end()) {
         const int g = (int) gclk.size();
         it = gidx.emplace(key, g).first;
         gclk.emplace_back(key.begin() + 1, key.end());
         int y = 0;
         while (y < (int) years.size() && years[y] != d[0]) y++;
         if (y == (int) years.size()) years.push_back(d[0]);
         gyear.push_back(y);
         gmax.push_back((int) gclk.back().size() - 1);
      }
      sgrp[s] = it->second;
   }
   const int G = (int) gclk.size(), Y = (int) years.size();
   // cos / sin of each constituent's angle at each step of each group
   double ang[SL_NUMT];
   for (int k = 0; k < SL_NUMT; k++) ang[k] = sl_tide_speed(k);
   std::vector<long> goff(G);
   size_t ncs = 0;
   for (int g = 0; g < G; g++) {
      goff[g] = (long) ncs;
      ncs += gclk[g].size() * 2 * SL_NUMT;
   }
   // (the reference's build computes each pair with one sincos call, and
   // glibc's sincos differs from separate sin and cos in about 0.1 % of
   // arguments, so it is taken the same way here)
   // the number of the tide call after each step of each group (-1: none)
   std::vector<int> calls(ncs / (2 * SL_NUMT));
   for (int g = 0; g < G; g++) {
      int *c = &calls[goff[g] / (2 * SL_NUMT)], k = 0;
      double tclk = gclk[g][0];
      c[0] = 0;
      for (size_t n = 1; n < gclk[g].size(); n++) {
         if (!six) c[n] = (int) n;
         else if (gclk[g][n] >= tclk) {
            c[n] = ++k;
            tclk += 6 * 60;
         } else
            c[n] = -1;
      }
   }
   std::vector<double> casa(ncs);
   for (int g = 0; g < G; g++) {
      const double ac = sl_clock_scan_date(years[gyear[g]], 1, 1);
      double *o = &casa[goff[g]];
      for (double clk : gclk[g]) {
         const double now = (clk - ac) / 3600.;
         for (int k = 0; k < SL_NUMT; k++) sincos(ang[k] * now, &o[SL_NUMT + k], &o[k]);
         o += 2 * SL_NUMT;
      }
   }
   // squares: where the tide is added, and where it replaces HB over the
   // static height; ZB in double
   std::vector<unsigned char> sub((size_t) b->imxb * b->jmxb, 0), cfl(nc);
   for (int l = 1; l <= b->nsqrwc; l++) {   // ISSUBGRID (issubgrid.f:14-33)
      const bool narrow = l <= b->nsqrw ? b->delcut[l - 1] < 1.
                                        : (b->cutli[l - b->nsqrw - 1] < 1. || b->cutle[l - b->nsqrw - 1] < 1.);
      const int i = b->isqr[l - 1], j = b->jsqr[l - 1];
      if (narrow && i >= 1 && i <= b->imxb && j >= 1 && j <= b->jmxb) sub[PIX(b, i, j)] = 1;
   }
   std::vector<double> zb(nc);
   for (int c = 0; c < nc; c++) {
      const int i = c % ni + 1, j = c / ni + 1;
      const double z = b->zb[PIX(b, i, j)];
      bool use = true;
      if (mode != 2) {
         if (z <= thresh) use = false;
         else if (mode != 21 && sub[PIX(b, i, j)]) use = false;
      }
      cfl[c] = ((mode == 21 || mode == 22) && z >= 290) ? 2 : use ? 1 : 0;
      zb[c] = z;
   }
   std::vector<int> lst;   // the squares a step changes
   for (int c = 0; c < nc; c++)
      if (cfl[c]) lst.push_back(c);
   const int nl = (int) lst.size();
   if (lst.empty()) lst.push_back(0);
   const int nb = b->nbcpt;
   std::vector<int> bcc((size_t) std::max(nb, 1), 0);
   for (int n = 0; n < nb; n++) bcc[n] = (b->jsh[n] - 1) * ni + (b->ish[n] - 1);
   // one allocation: year tables, per-square values, groups, fields
   const size_t tab = (size_t) SL_NUMT * nc;
   size_t off[14], tot = 0;
   int n = 0;
   auto add = [&](size_t bytes) {
      off[n++] = tot;
      tot += (bytes + 255) & ~(size_t) 255;
   };
   add(Y * tab * 8); add(Y * tab * 8); add(Y * tab * 8);    // amp cv sv
   add((size_t) nc * 8); add((size_t) nc * 8); add(nc);     // mtl zb cfl
   add(bcc.size() * 4);                                     // bccell
   add(ncs * 8); add((size_t) G * 8);                        // casa goff
   add((size_t) G * 4); add((size_t) G * 4); add((size_t) S * 4);   // gyear gmax sgrp
   add(lst.size() * 4);                                     // list
   add(calls.size() * 4);                                   // call
   const size_t foff = tot;
   tot += (size_t) SL_MAXSPLIT * G * 2 * nc * 8;           // field, one per storm part
   char *d = nullptr;
   std::vector<double> amp(tab), cv(tab), sv(tab);
   hipError_t he = hipMalloc((void **) &d, tot);
   for (int y = 0; y < Y && he == hipSuccess; y++) {
      if (sl_tide_year(&t, ft03, years[y], amp.data(), cv.data(), sv.data(), e)) {
         sl_tide_free(&t);
         (void) hipFree(d);
         seterr(err, errlen, "%s", e);
         return SLG_EINPUT;
      }
      he = hipMemcpy(d + off[0] + y * tab * 8, amp.data(), tab * 8, hipMemcpyHostToDevice);
      if (he == hipSuccess) he = hipMemcpy(d + off[1] + y * tab * 8, cv.data(), tab * 8, hipMemcpyHostToDevice);
      if (he == hipSuccess) he = hipMemcpy(d + off[2] + y * tab * 8, sv.data(), tab * 8, hipMemcpyHostToDevice);
   }
   const void *src[14] = {nullptr, nullptr, nullptr, t.mtl, zb.data(), cfl.data(), bcc.data(), casa.data(),
                          goff.data(), gyear.data(), gmax.data(), sgrp.data(), lst.data(), calls.data()};
   const size_t len[14] = {0, 0, 0, (size_t) nc * 8, (size_t) nc * 8, (size_t) nc, bcc.size() * 4, ncs * 8,
                           (size_t) G * 8, (size_t) G * 4, (size_t) G * 4, (size_t) S * 4, lst.size() * 4,
                           calls.size() * 4};
   for (int k = 3; k < 14 && he == hipSuccess; k++) he = hipMemcpy(d + off[k], src[k], len[k], hipMemcpyHostToDevice);
   sl_tide_free(&t);
   if (he != hipSuccess) {
      if (d) (void) hipFree(d);
      seterr(err, errlen, "HIP error %s (%d)", hipGetErrorString(he), (int) he);
      return SLG_EHIP;
   }
   SlTideDev &T = X->tdev;
   T.ni = ni;
   T.ncell = nc;
   T.nbcpt = nb;
   T.ngrp = G;
   T.amp = (const double *) (d + off[0]);
   T.cv = (const double *) (d + off[1]);
   T.sv = (const double *) (d + off[2]);
   T.mtl = (const double *) (d + off[3]);
   T.zb = (const double *) (d + off[4]);
   T.cfl = (const uint8_t *) (d + off[5]);
   T.bccell = (const int *) (d + off[6]);
   T.casa = (const double *) (d + off[7]);
   T.goff = (const long *) (d + off[8]);
   T.gyear = (const int *) (d + off[9]);
   T.gmax = (const int *) (d + off[10]);
   T.sgrp = (const int *) (d + off[11]);
   T.list = (const int *) (d + off[12]);
   T.call = (const int *) (d + off[13]);
   T.nlist = nl;
   T.field = (double *) (d + foff);
   X->tide_mem = d;
   X->tide_mode = mode;
   return 0;
}

// STATICHT (staticht.f:14-41): the static height at square (i,j) from the
// storm values and radial table of the storm's next step, in the batch's
// precision, as BDRYHT computes it without the tide term.
template <class R>
static double static_height(const slg_batch *X, int storm, int i, int j)
{
   const storm_sched &S = X->sc[storm];
   const BasinImg<R> &I = img(X->b, R());
   const int k0 = std::min(X->step, S.nsteps - 1);
   const R *P = &sched_params(S, R())[(size_t) k0 * SL_NPAR];
   const R *T = &sched_tables(S, R())[(size_t) (int) P[P_TVER] * SL_NTABV * 8];
   const R xr = I.vi[2][i - 1] * I.vj[3][j - 1], yr = I.vi[3][i - 1] * I.vj[2][j - 1];
   const R x = xr - P[P_C1] - P[P_AX], y = yr - P[P_C2] - P[P_AY];
   const R rsq = x * x + y * y, r1 = std::sqrt(rsq) / R(5280.) + R(1.);
   int k = (int) r1;
   const R r2 = (R) k, dr = r1 - r2;
   if (k > 790) k = 790;
   const R *Rk = T + (size_t) k * 8, *Rk1 = Rk + 8;
   return (double) (Rk[5] + dr * (Rk1[5] - Rk[5]) + P[P_SEA]);   // column 5: DELP
}

extern "C" int slg_batch_static_height(const slg_batch *X, int storm, int i, int j, double *ht)
{
   const sl_basin *b = &X->b->h;
   if (storm < 0 || storm >= X->nstorm || i < 1 || i > b->imxb || j < 1 || j > b->jmxb ||
       !(X->prec == 64 ? X->b->i64.built : X->b->i32.built))
      return SLG_EINPUT;
   if (!X->sc[storm].hostp) return SLG_EUNSUP;   // the values live on the device (open with SLG_SCHED_HOST)
   *ht = X->prec == 64 ? static_height<double>(X, storm, i, j) : static_height<float>(X, storm, i, j);
   return 0;
}

extern "C" int slg_batch_track(const slg_batch *X, int storm, int n, double *out9)
{
   if (storm < 0 || storm >= X->nstorm) return SLG_EINPUT;
   const storm_sched &S = X->sc[storm];
   if (n < 0 || n > S.nsteps || S.rows.size() < (size_t) (n + 1) * 90) return SLG_EINPUT;
   const double *r = &S.rows[(size_t) n * 90];
   sl_storm t = S.s0;
   t.c1 = r[3];
   t.c2 = r[4];
   t.ax = r[5];
   t.ay = r[6];
   double lat, lon;
   sl_storm_latlon(&t, &X->b->h, &lat, &lon);
   out9[0] = r[0];
   out9[1] = r[23];
   out9[2] = r[2];
   out9[3] = lat;
   out9[4] = lon;
   out9[5] = r[9];
   out9[6] = r[12];
   out9[7] = r[7];
   out9[8] = X->b->h.azmth - r[8] * 180. / 3.14159265358979323846;
   return 0;
}

// Start of the run as INITAL hands it to the C driver (inital.f:204-249): the
// trk DateTime moved back by JHR - IBGNT hours.  out5 = year, month, day, hour,
// minute.
extern "C" int slg_batch_start_date(const slg_batch *X, int storm, int *out5)
{
   if (storm < 0 || storm >= X->nstorm) return SLG_EINPUT;
   const sl_storm &s = X->sc[storm].s0;
   int ndy[13] = {0, 31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
   int myr = s.date_yr, mbnt = s.date_mo, mday = s.date_dy, mhr = s.date_hr;
   if (myr % 4 == 0) ndy[2] = 29;
   if (myr % 100 == 0) ndy[2] = 28;
   if (myr % 400 == 0) ndy[2] = 29;
   const int ntmlng = s.jhr - s.ibgnt;
   int mbhr = mhr - ntmlng;
   int mbdy = mday + mbhr / 24;   // Fortran integer division truncates toward zero
   mbhr = mbhr % 24;
   if (mbhr < 0) {
      mbdy = mbdy - 1;
      mbhr = 24 + mbhr;
   }
   if (mbdy < 1) {
      mbnt = mbnt - 1;
      if (mbnt < 1) {
         myr = myr - 1;
         mbnt = mbnt + 12;
      }
      mbdy = ndy[mbnt] + mbdy;
   }
   if (mbdy > ndy[mbnt]) {
      mbdy = mbdy - ndy[mbnt];
      mbnt = mbnt + 1;
      if (mbnt > 12) {
         myr = myr + 1;
         mbnt = mbnt - 12;
      }
   }
   out5[0] = myr;
   out5[1] = mbnt;
   out5[2] = mbdy;
   out5[3] = mbhr;
   out5[4] = s.date_mn;
   return 0;
}

template <class R>
static int batch_get(slg_batch *X, int storm, const char *name, double *dst, char *err, int errlen)
{
   const sl_basin *b = &X->b->h;
   const SlRunDevT<R> &D = run(X, R());
   const R *src;
   if (!strcmp(name, "FLWSQR")) {   // transport across each 1-D flow side (nsqrwc values)
      if (!X->dmem) {
         seterr(err, errlen, "%s", "batch not uploaded");
         return SLG_EINPUT;
      }
      const int nl = b->nsqrwc;
      std::vector<R> f((size_t) std::max(nl, 1));
      HIPOK(hipStreamSynchronize(X->st));
      if (nl) HIPOK(hipMemcpy(f.data(), D.flwsqr + (size_t) storm * nl, (size_t) nl * sizeof(R), hipMemcpyDeviceToHost));
      for (int l = 0; l < nl; l++) dst[l] = (double) f[l];
      return 0;
   }
   if (!strcmp(name, "HBMAX")) {   // windowed maxima, held in double
      if (!X->hbmax) {
         seterr(err, errlen, "%s", "windowed maxima are off");
         return SLG_EINPUT;
      }
      std::vector<double> f(D.fstride);
      HIPOK(hipStreamSynchronize(X->st));
      HIPOK(hipMemcpy(f.data(), X->hbmax + (size_t) storm * D.fstride, D.fstride * 8, hipMemcpyDeviceToHost));
      const int P = X->b->pitch;
      for (int j = 1; j <= b->jmxb; j++)
         for (int i = 1; i <= b->imxb; i++)
            dst[PIX(b, i, j)] = (i < b->imxb && j < b->jmxb) ? f[(size_t) (j - 1) * P + (i - 1)] : 0.;
      return 0;
   }
   if (!strcmp(name, "HB")) src = X->last_par[storm] ? D.hb1 : D.hb0;
   else if (!strcmp(name, "HBS")) {
      if (X->hbs_buf < 0 || X->sc[storm].nsteps < X->step) {
         seterr(err, errlen, "%s", "HB before FILTER is kept only by the staged path");
         return SLG_EINPUT;
      }
      src = X->hbs_buf ? D.hb1 : D.hb0;
   }
   else if (!strcmp(name, "UB")) src = D.ub;
   else if (!strcmp(name, "VB")) src = D.vb;
   else if (!strcmp(name, "HMX")) src = D.hmx;
   else {
      seterr(err, errlen, "unknown field %s", name);
      return SLG_EINPUT;
   }
   if (!X->dmem) {
      seterr(err, errlen, "%s", "batch not uploaded");
      return SLG_EINPUT;
   }
   std::vector<R> f(D.fstride);
   HIPOK(hipStreamSynchronize(X->st));
   HIPOK(hipMemcpy(f.data(), src + (size_t) storm * D.fstride, D.fstride * sizeof(R), hipMemcpyDeviceToHost));
   const int P = X->b->pitch;
   for (int j = 1; j <= b->jmxb; j++)
      for (int i = 1; i <= b->imxb; i++) dst[PIX(b, i, j)] = (double) f[(size_t) (j - 1) * P + (i - 1)];
   return 0;
}

// download a field of one storm as double [jmxb][imxb]
extern "C" int slg_batch_get(slg_batch *X, int storm, const char *name, double *dst, char *err,
                             int errlen)
{
   const sl_basin *b = &X->b->h;
   if (storm < 0 || storm >= X->nstorm) {
      seterr(err, errlen, "storm %s%d out of range", "", storm);
      return SLG_EINPUT;
   }
   if (!strcmp(name, "HB0")) {   // initial water (INTLHT), double, on the host
      const size_t np = (size_t) b->imxb * b->jmxb;
      if (X->hb0.empty()) sl_initial_water(b, &X->sc[storm].s0, dst);
      else memcpy(dst, &X->hb0[(size_t) storm * np], np * sizeof(double));
      return 0;
   }
   if (!X->dmem) {
      seterr(err, errlen, "%s", "batch not uploaded");
      return SLG_EINPUT;
   }
   const int rc = check_runs_any(X, storm, 1, err, errlen);
   if (rc) return rc;
   return X->prec == 64 ? batch_get<double>(X, storm, name, dst, err, errlen)
                        : batch_get<float>(X, storm, name, dst, err, errlen);
}

extern "C" int slg_batch_stats(slg_batch *X, int storm, int *out8, char *err, int errlen)
{
   if (!X->dmem || storm < 0 || storm >= X->nstorm) {
      seterr(err, errlen, "%s", "no such storm on the device");
      return SLG_EINPUT;
   }
   const int *stat = X->prec == 64 ? X->d64.stat : X->d32.stat;
   HIPOK(hipStreamSynchronize(X->st));
   HIPOK(hipMemcpy(out8, stat + storm * 8, 32, hipMemcpyDeviceToHost));
   return 0;
}

template <class R>
static int batch_envelopes(slg_batch *X, short *env, short *mom, char *err, int errlen)
{
   slg_basin *B = X->b;
   const sl_basin *b = &B->h;
   const size_t n = (size_t) b->imxb * b->jmxb;
   short *denv = nullptr, *dmom = nullptr;
   if (env) HIPOK(hipMalloc((void **) &denv, n * X->nstorm * 2));
   if (mom) HIPOK(hipMalloc((void **) &dmom, n * 2));
   int rc = dev_envelopes(&img(B, R()).dev, &run(X, R()), img(B, R()).zbd, 0, X->nstorm, denv, dmom, X->st);
   if (rc == 0) {
      if (env) HIPOK(hipMemcpyAsync(env, denv, n * X->nstorm * 2, hipMemcpyDeviceToHost, X->st));
      if (mom) HIPOK(hipMemcpyAsync(mom, dmom, n * 2, hipMemcpyDeviceToHost, X->st));
      HIPOK(hipStreamSynchronize(X->st));
   }
   if (denv) (void) hipFree(denv);
   if (dmom) (void) hipFree(dmom);
   if (rc) {
      seterr(err, errlen, "envelope kernel failed (%s%d)", "", -rc);
      return SLG_EHIP;
   }
   return 0;
}

// Envelopes of every run ([nstorm][jmxb][imxb]) and their MOM ([jmxb][imxb]),
// computed on the device; either output may be NULL.
extern "C" int slg_batch_envelopes(slg_batch *X, short *env, short *mom, char *err, int errlen)
{
   if (!X->dmem) {
      seterr(err, errlen, "%s", "batch not uploaded");
      return SLG_EINPUT;
   }
   const int rc = check_runs_any(X, 0, X->nstorm, err, errlen);
   if (rc) return rc;
   return X->prec == 64 ? batch_envelopes<double>(X, env, mom, err, errlen)
                        : batch_envelopes<float>(X, env, mom, err, errlen);
}

// ---- envelopes without allocation ----------------------------------------
// No counterpart in the reference.  For callers that cannot allocate device
// memory (code under HIP graph capture, serving loops that forbid hipMalloc),
// slg_batch_envelopes_noalloc stages the device envelopes and the MOM in this
// fixed array instead of the buffers slg_batch_envelopes allocates: per GPU
// one copy at its own address, held until the process ends.  Its device
// address is taken once per device (hipGetSymbolAddress, with that device
// current); host code never passes the array's own name to a kernel or a copy.
#define SLG_NOALLOC_BYTES ((size_t) 32 << 20)
static __device__ __align__(256) unsigned char scratch[SLG_NOALLOC_BYTES];
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[64];

static int scratch_of_device(void **out, char *err, int errlen)
{
   int dev = -1;
   HIPOK(hipGetDevice(&dev));
   if (dev < 0 || dev >= 64) {
      seterr(err, errlen, "device %s%d out of range for the scratch table", "", dev);
      return SLG_EHIP;
   }
   std::lock_guard<std::mutex> lk(g_scratch_mu);
   if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      HIPOK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = p;
   }
   *out = g_scratch_ptr[dev];
   return 0;
}

template <class R>
static int batch_envelopes_noalloc(slg_batch *X, short *env, short *mom, char *err, int errlen)
{
   slg_basin *B = X->b;
   const sl_basin *b = &B->h;
   const size_t n = (size_t) b->imxb * b->jmxb, nb = (n * 2 + 255) & ~(size_t) 255;
   if (2 * nb > SLG_NOALLOC_BYTES) {
      seterr(err, errlen, "basin of %s%d points: its envelope and MOM do not fit the 32 MiB scratch", "", (int) n);
      return SLG_ELIMIT;
   }
   void *scr = nullptr;
   int rc = scratch_of_device(&scr, err, errlen);
   if (rc) return rc;
   short *dmom = (short *) scr, *denv = (short *) ((char *) scr + nb);
   // storms per chunk: the envelopes of a chunk share the scratch after the MOM
   const int per = (int) std::min<size_t>((size_t) X->nstorm, (SLG_NOALLOC_BYTES - nb) / (n * 2));
   for (int s0 = 0; s0 < X->nstorm; s0 += per) {
      const int ns = std::min(per, X->nstorm - s0);
      rc = dev_envelopes(&img(B, R()).dev, &run(X, R()), img(B, R()).zbd, s0, ns, env ? denv : nullptr,
                         mom ? dmom : nullptr, X->st, nullptr, s0 > 0);
      if (rc) {
         seterr(err, errlen, "envelope kernel failed (%s%d)", "", -rc);
         return SLG_EHIP;
      }
      // the next chunk's kernel follows this copy on the stream
      if (env) HIPOK(hipMemcpyAsync(env + (size_t) s0 * n, denv, (size_t) ns * n * 2, hipMemcpyDeviceToHost, X->st));
   }
   if (mom) HIPOK(hipMemcpyAsync(mom, dmom, n * 2, hipMemcpyDeviceToHost, X->st));
   HIPOK(hipStreamSynchronize(X->st));
   return 0;
}

// As slg_batch_envelopes, allocating nothing (the staging is the fixed scratch
// above).  Calls that overlap on one GPU share that scratch and race.
extern "C" int slg_batch_envelopes_noalloc(slg_batch *X, short *env, short *mom, char *err, int errlen)
{
   if (!X->dmem) {
      seterr(err, errlen, "%s", "batch not uploaded");
      return SLG_EINPUT;
   }
   const int rc = check_runs_any(X, 0, X->nstorm, err, errlen);
   if (rc) return rc;
   return X->prec == 64 ? batch_envelopes_noalloc<double>(X, env, mom, err, errlen)
                        : batch_envelopes_noalloc<float>(X, env, mom, err, errlen);
}

static void rex_free(slg_batch *X)
{
   for (void *p : {X->rex_scr, X->rex_out, X->rex_pack, X->rex_nb, X->rex_off})
      if (p) (void) hipFree(p);
   if (X->rex_host) (void) hipHostFree(X->rex_host);
   X->rex_scr = X->rex_out = X->rex_pack = X->rex_nb = X->rex_off = nullptr;
   X->rex_host = nullptr;
   X->rex_chunk = 0;
   X->rex_cap = 0;
}

// the batch's rex buffers for chunk storms of cap bytes (per: scratch bytes a
// storm), allocated on first use and when a call needs more
static bool rex_buffers(slg_batch *X, int chunk, long cap, size_t per)
{
   if (X->rex_chunk >= chunk && X->rex_cap >= cap) return true;
   (void) hipStreamSynchronize(X->st);
   rex_free(X);
   const size_t fr = (size_t) cap * chunk;
   if (hipMalloc(&X->rex_scr, per * chunk) != hipSuccess || hipMalloc(&X->rex_out, fr) != hipSuccess ||
       hipMalloc(&X->rex_pack, fr) != hipSuccess || hipMalloc(&X->rex_nb, sizeof(int) * chunk) != hipSuccess ||
       hipMalloc(&X->rex_off, sizeof(long) * chunk) != hipSuccess ||
       hipHostMalloc((void **) &X->rex_host, fr, 0) != hipSuccess) {
      rex_free(X);
      return false;
   }
   X->rex_chunk = chunk;
   X->rex_cap = cap;
   return true;
}

// rex frames of storms s0 .. s0+ns-1, in chunks whose scratch (2N+1 words a
// storm for run starts and records) stays near 256 MB
template <class R>
static int batch_rex(slg_batch *X, int src, int s0, int ns, unsigned char *buf, long cap, int *nbytes, char *err,
                     int errlen)
{
   slg_basin *B = X->b;
   const SlRunDevT<R> &D = run(X, R());
   const long N = (long) (B->h.imxb - 1) * (B->h.jmxb - 1);
   const R *hbs = nullptr;
   if (src == 0) hbs = X->hbs_buf ? D.hb1 : D.hb0;
   const double *hmax = src == 2 ? X->hbmax : nullptr;
   const size_t per = (size_t) (2 * N + 1) * 4;
   int chunk = (int) std::max<size_t>(1, ((size_t) 256 << 20) / per);
   if (chunk > ns) chunk = ns;
   if (!rex_buffers(X, chunk, cap, per)) {
      seterr(err, errlen, "%s", "no device memory for the rex frame buffers");
      return SLG_EHIP;
   }
   unsigned *scr = (unsigned *) X->rex_scr;
   unsigned char *dout = (unsigned char *) X->rex_out, *dpack = (unsigned char *) X->rex_pack;
   int *dnb = (int *) X->rex_nb;
   long *doff = (long *) X->rex_off;
   int rc = 0;
   HIPOK(hipStreamSynchronize(X->st));
   std::vector<long> off(chunk);
   for (int c = 0; c < ns && !rc; c += chunk) {
      const int n = std::min(chunk, ns - c);
      int r = dev_rex(&img(B, R()).dev, &D, hbs, hmax, img(B, R()).zbd, s0 + c, n, dout, cap, dnb, scr, X->st);
      if (r || hipMemcpyAsync(nbytes + c, dnb, sizeof(int) * n, hipMemcpyDeviceToHost, X->st) != hipSuccess ||
          hipStreamSynchronize(X->st) != hipSuccess) {
         rc = SLG_EHIP;
         break;
      }
      long tot = 0;
      for (int k = 0; k < n; k++) {
         if (nbytes[c + k] < 0) rc = SLG_ELIMIT;
         off[k] = tot;
         tot += nbytes[c + k] > 0 ? nbytes[c + k] : 0;
      }
      if (rc) break;
      // the chunk's frames back to back on the device, then one copy
      if (hipMemcpyAsync(doff, off.data(), sizeof(long) * n, hipMemcpyHostToDevice, X->st) != hipSuccess ||
          sl_dev_rex_pack(dout, cap, dnb, doff, n, dpack, X->st) ||
          hipMemcpyAsync(X->rex_host, dpack, (size_t) tot, hipMemcpyDeviceToHost, X->st) != hipSuccess ||
          hipStreamSynchronize(X->st) != hipSuccess) {
         rc = SLG_EHIP;
         break;
      }
      for (int k = 0; k < n; k++)
         memcpy(buf + (size_t) (c + k) * cap, X->rex_host + off[k], (size_t) nbytes[c + k]);
   }
   if (rc == SLG_ELIMIT) seterr(err, errlen, "a rex frame is longer than cap = %s%d bytes", "", (int) cap);
   else if (rc) seterr(err, errlen, "%s", "rex frame kernel or copy failed");
   return rc;
}

// rex v2 frames of storms s0 .. s0+ns-1 from HHB of the last step or HMX
// the source of rex values: 0 HHB of the last step, 1 HMX, 2 the windowed
// maxima; negative: an SLG_E* code, err set
static int rex_source(slg_batch *X, const char *field, int s0, int ns, char *err, int errlen)
{
   if (!X->dmem) {
      seterr(err, errlen, "%s", "batch not uploaded");
      return SLG_EINPUT;
   }
   if (!field || s0 < 0 || ns < 0 || s0 + ns > X->nstorm) {
      seterr(err, errlen, "%s", "bad storm range");
      return SLG_EINPUT;
   }
   const bool hhb = !strcmp(field, "HHB");
   const int src = hhb ? 0 : !strcmp(field, "HMX") ? 1 : !strcmp(field, "HBMAX") ? 2 : -1;
   if (src < 0) {
      seterr(err, errlen, "unknown rex field %s (HHB, HMX or HBMAX)", field);
      return SLG_EINPUT;
   }
   if (src == 2 && !X->hbmax) {
      seterr(err, errlen, "%s", "windowed maxima are off (slg_batch_set_maxstat)");
      return SLG_EINPUT;
   }
   if (hhb) {
      bool ok = X->step > 0 && X->hbs_buf >= 0;
      for (int s = s0; s < s0 + ns && ok; s++) ok = X->sc[s].nsteps >= X->step;
      if (!ok) {
         seterr(err, errlen, "%s", "HHB needs HB before FILTER of the last step of every storm: step it on the "
                                   "staged path, or a step that did not filter");
         return SLG_EINPUT;
      }
   }
   const int rc = ns ? check_runs_any(X, s0, ns, err, errlen) : 0;
   return rc ? rc : src;
}

extern "C" int slg_batch_rex_frames(slg_batch *X, const char *field, int s0, int ns, unsigned char *buf, long cap,
                                    int *nbytes, char *err, int errlen)
{
   if (!buf || !nbytes || cap < 1) {
      seterr(err, errlen, "%s", "bad buffer");
      return SLG_EINPUT;
   }
   const int src = rex_source(X, field, s0, ns, err, errlen);
   if (src < 0) return src;
   if (ns == 0) return 0;
   return X->prec == 64 ? batch_rex<double>(X, src, s0, ns, buf, cap, nbytes, err, errlen)
                        : batch_rex<float>(X, src, s0, ns, buf, cap, nbytes, err, errlen);
}

template <class R>
static int batch_rex_points(slg_batch *X, int src, int s0, int ns, int npts, const int *k, short *out)
{
   slg_basin *B = X->b;
   const SlRunDevT<R> &D = run(X, R());
   const R *hbs = src == 0 ? (X->hbs_buf ? D.hb1 : D.hb0) : nullptr;
   const size_t need = sizeof(int) * (size_t) npts + sizeof(int16_t) * (size_t) npts * ns;
   if (X->pt_cap < need) {
      if (hipStreamSynchronize(X->st) != hipSuccess) return SLG_EHIP;
      if (X->pt_buf) (void) hipFree(X->pt_buf);
      X->pt_buf = nullptr;
      X->pt_cap = 0;
      if (hipMalloc(&X->pt_buf, need) != hipSuccess) return SLG_EHIP;
      X->pt_cap = need;
   }
   int *dk = (int *) X->pt_buf;
   int16_t *dout = (int16_t *) (dk + npts);
   if (hipMemcpyAsync(dk, k, sizeof(int) * (size_t) npts, hipMemcpyHostToDevice, X->st) != hipSuccess)
      return SLG_EHIP;
   // storms go on the grid's y dimension: at most 65 535 a launch
   for (int c = 0; c < ns; c += 65535) {
      const int n = std::min(65535, ns - c);
      if (dev_rex_points(&img(B, R()).dev, &D, hbs, src == 2 ? X->hbmax : nullptr, img(B, R()).zbd, s0 + c, n, dk,
                         npts, dout + (size_t) c * npts, X->st))
         return SLG_EHIP;
   }
   if (hipMemcpyAsync(out, dout, sizeof(int16_t) * (size_t) npts * ns, hipMemcpyDeviceToHost, X->st) != hipSuccess ||
       hipStreamSynchronize(X->st) != hipSuccess)
      return SLG_EHIP;
   return 0;
}

// rex values at listed squares of storms s0 .. s0+ns-1 (rexout's series)
extern "C" int slg_batch_rex_points(slg_batch *X, const char *field, int s0, int ns, int npts, const int *xy,
                                    short *out, char *err, int errlen)
{
   if (npts < 0 || npts > SLG_MAX_POINTS || (npts && (!xy || !out))) {
      seterr(err, errlen, "%s", "bad point list or output buffer");
      return SLG_EINPUT;
   }
   const int src = rex_source(X, field, s0, ns, err, errlen);
   if (src < 0) return src;
   if (ns == 0 || npts == 0) return 0;
   const int I1 = X->b->h.imxb - 1, J1 = X->b->h.jmxb - 1;
   std::vector<int> k(npts);
   for (int p = 0; p < npts; p++) {
      const int x = xy[2 * p], y = xy[2 * p + 1];
      if (x < 1 || x > I1 || y < 1 || y > J1) {
         seterr(err, errlen, "point %s%d is outside the rex frame's squares", "", p);
         return SLG_EINPUT;
      }
      k[p] = (x - 1) * J1 + (y - 1);
   }
   const int r = X->prec == 64 ? batch_rex_points<double>(X, src, s0, ns, npts, k.data(), out)
                               : batch_rex_points<float>(X, src, s0, ns, npts, k.data(), out);
   if (r) seterr(err, errlen, "%s", "rex point kernel or copy failed");
   return r;
}

// ---- P-Surge windowed maxima (slosh2.c:1227-1316) --------------------------
template <class R>
static int maxstat_init(slg_batch *X, double *h)
{
   int r = dev_maxstat(&img(X->b, R()).dev, &run(X, R()), (const R *) nullptr, img(X->b, R()).zbd, h, nullptr, 1,
                       X->st);
   return (r || hipStreamSynchronize(X->st) != hipSuccess) ? SLG_EHIP : 0;
}

static int maxstat_reset(slg_batch *X, double *h)
{
   if (!X->dmem || !h) return SLG_EINPUT;
   return X->prec == 64 ? maxstat_init<double>(X, h) : maxstat_init<float>(X, h);
}

extern "C" int slg_batch_maxstat_reset(slg_batch *X)
{
   return maxstat_reset(X, X->hbmax);
}

extern "C" int slg_batch_maxstat2_reset(slg_batch *X)
{
   return maxstat_reset(X, X->hbmax2);
}

// one window's buffer on (allocated and started) or off (freed)
static int maxstat_switch(slg_batch *X, double **h, int on)
{
   if (!X->dmem) return SLG_EINPUT;
   if (hipStreamSynchronize(X->st) != hipSuccess) return SLG_EHIP;
   drop_graph(X);   // captured steps hold the update or not
   if (!on) {
      if (*h) (void) hipFree(*h);
      *h = nullptr;
      return 0;
   }
   const size_t fs = X->prec == 64 ? X->d64.fstride : X->d32.fstride;
   if (!*h && hipMalloc((void **) h, sizeof(double) * fs * X->nstorm) != hipSuccess) {
      *h = nullptr;
      return SLG_EHIP;
   }
   return maxstat_reset(X, *h);
}

extern "C" int slg_batch_set_maxstat(slg_batch *X, int on)
{
   return maxstat_switch(X, &X->hbmax, on);
}

extern "C" int slg_batch_set_maxstat2(slg_batch *X, int on)
{
   return maxstat_switch(X, &X->hbmax2, on);
}

template <class R>
static int maxstat_env(slg_batch *X, const double *h, short *env, char *err, int errlen)
{
   const sl_basin *b = &X->b->h;
   const size_t n = (size_t) b->imxb * b->jmxb;
   short *denv = nullptr;
   HIPOK(hipMalloc((void **) &denv, n * X->nstorm * 2));
   int rc = dev_envelopes(&img(X->b, R()).dev, &run(X, R()), img(X->b, R()).zbd, 0, X->nstorm, denv, nullptr,
                          X->st, h);
   if (rc == 0 && (hipMemcpyAsync(env, denv, n * X->nstorm * 2, hipMemcpyDeviceToHost, X->st) != hipSuccess ||
                   hipStreamSynchronize(X->st) != hipSuccess))
      rc = -1;
   (void) hipFree(denv);
   if (rc) {
      seterr(err, errlen, "%s", "windowed-maxima env kernel or copy failed");
      return SLG_EHIP;
   }
   return 0;
}

static int maxstat_env_any(slg_batch *X, const double *h, short *env, char *err, int errlen)
{
   if (!X->dmem || !h || !env) {
      seterr(err, errlen, "%s", "windowed maxima are off (slg_batch_set_maxstat / _maxstat2) or no buffer");
      return SLG_EINPUT;
   }
   const int rc = check_runs_any(X, 0, X->nstorm, err, errlen);
   if (rc) return rc;
   return X->prec == 64 ? maxstat_env<double>(X, h, env, err, errlen) : maxstat_env<float>(X, h, env, err, errlen);
}

extern "C" int slg_batch_maxstat_env(slg_batch *X, short *env, char *err, int errlen)
{
   return maxstat_env_any(X, X->hbmax, env, err, errlen);
}

extern "C" int slg_batch_maxstat2_env(slg_batch *X, short *env, char *err, int errlen)
{
   return maxstat_env_any(X, X->hbmax2, env, err, errlen);
}

// Datums of a run's track for the env trailer (ReadTrkFile, slosh2.c:1885-1975):
// OceanDatum and LakeDatum; with a CanalDatum line, LakeDatum and CanalDatum.
static int read_datums(const char *trk_path, float *ht1, float *ht2)
{
   FILE *fp = fopen(trk_path, "r");
   char line[512];
   float oc = 0.f, lk = 0.f, cn = -999.f;
   if (!fp) return SLG_EINPUT;
   if (fgets(line, sizeof line, fp) && strncmp(line, "FileVersion=", 12)) {
      // the 1992 format: the datum line after the titles, 100 rows, IBGNT
      // ITEND JHR and the landfall time (ReadTrkFile, slosh2.c:1940-1990)
      int k = 1;
      while (k < 105 && fgets(line, sizeof line, fp)) k++;
      fclose(fp);
      if (k < 105) return SLG_EINPUT;
      if (line[10] == 'x' || line[10] == 'X') {
         line[16] = 0;
         *ht2 = (float) atof(line + 11);
         line[10] = 0;
         *ht1 = (float) atof(line + 5);
      } else {
         line[10] = 0;
         *ht2 = (float) atof(line + 6);
         line[5] = 0;
         *ht1 = (float) atof(line);
      }
      return 0;
   }
   rewind(fp);
   while (fgets(line, sizeof line, fp)) {
      if (!strncmp(line, "Hour,", 5)) break;
      if (line[0] == '#') continue;
      if (!strncmp(line, "OceanDatum=", 11)) oc = (float) atof(line + 11);
      else if (!strncmp(line, "LakeDatum=", 10)) lk = (float) atof(line + 10);
      else if (!strncmp(line, "CanalDatum=", 11)) cn = (float) atof(line + 11);
   }
   fclose(fp);
   const bool canal = cn > -100;   // Okeechobee-style datums: lake and canal
   *ht1 = canal ? lk : oc;
   *ht2 = canal ? cn : lk;
   return 0;
}

// The comment EnvSave writes for a run (ReadTrkFile, slosh2.c:1884-1975):
// FileVersion=2013 files give their ExtraInfo= value (or "FileVersion=2013"),
// 1992 files their two title lines, each cut or blank-padded to 80 columns.
extern "C" int slg_track_env_comment(const char *trk_path, char *out, int cap)
{
   FILE *fp = fopen(trk_path, "r");
   char h1[512], h2[512], b[512];
   if (!fp || cap < 161) {
      if (fp) fclose(fp);
      return SLG_EINPUT;
   }
   memset(out, 0, (size_t) cap);
   if (!fgets(h1, 200, fp)) h1[0] = 0;
   if (!strncmp(h1, "FileVersion=", 12)) {
      snprintf(out, (size_t) cap, "%s", "FileVersion=2013");
      while (fgets(b, 200, fp)) {
         if (!strncmp(b, "Hour,", 5)) break;
         if (b[0] == '#') continue;
         if (!strncmp(b, "ExtraInfo=", 10)) snprintf(out, (size_t) cap, "%.160s", b + 10);
      }
   } else {
      if (!fgets(h2, 200, fp)) h2[0] = 0;
      h1[strcspn(h1, "\n")] = 0;
      h2[strcspn(h2, "\n")] = 0;
      snprintf(out, (size_t) cap, "%-80.80s%-80.80s", h1, h2);
   }
   fclose(fp);
   return 0;
}

extern "C" int slg_track_datums(const char *trk_path, float *ht1, float *ht2)
{
   if (read_datums(trk_path, ht1, ht2)) return SLG_EINPUT;
   sl_datum_override(ht1, ht2);    // slosh2.c:2238, a surge-only run
   return 0;
}

// The same for a run with the gridded tide (InitWater_TideModeOverride,
// slosh2.c:408-482, f_tide = mode).
extern "C" int slg_track_datums_tide(const char *trk_path, int mode, float *ht1, float *ht2)
{
   if ((mode != 2 && mode != 21 && mode != 22 && mode != 24) || read_datums(trk_path, ht1, ht2)) return SLG_EINPUT;
   sl_datum_override_tide(ht1, ht2, mode);
   return 0;
}

// env file in the layout EnvSave writes (slosh2.c:1611-1671): int32 imxb,
// jmxb, a 160-byte comment (NUL padded), imxb*jmxb int16 (I fastest), two
// float datums; all little-endian.
extern "C" int slg_write_env(const char *path, int imxb, int jmxb, const char *comment, const short *vals,
                             float ht1, float ht2)
{
   FILE *fp = fopen(path, "wb");
   if (!fp) return SLG_EINPUT;
   char c[160];
   memset(c, 0, sizeof c);
   if (comment) strncpy(c, comment, sizeof c);
   const int32_t d[2] = {imxb, jmxb};
   int ok = fwrite(d, 4, 2, fp) == 2 && fwrite(c, 1, 160, fp) == 160 &&
            fwrite(vals, 2, (size_t) imxb * jmxb, fp) == (size_t) imxb * jmxb && fwrite(&ht1, 4, 1, fp) == 1 &&
            fwrite(&ht2, 4, 1, fp) == 1;
   fclose(fp);
   return ok ? 0 : SLG_EINPUT;
}

// Online autotuner: time each launch configuration (fused with 64 x 8, 32 x 16
// and 16 x 32 tiles, and staged) on consecutive chunks of the run's own steps -- every
// configuration computes identical fields, so the trial steps are real
// progress -- and keep the fastest.  ms_out (4 entries, may be NULL) receives
// the time per step of each candidate.
extern "C" int slg_batch_autotune(slg_batch *X, int steps_per_trial, int *rows_out, int *path_out,
                                  double *ms_out, char *err, int errlen)
{
   const int cand[4][2] = {{SLG_PATH_FUSED, 8}, {SLG_PATH_FUSED, 16}, {SLG_PATH_FUSED, 32}, {SLG_PATH_STAGED, 16}};
   double best = 1e300;
   int bi = 1;
   const int keep_graph = X->graph_steps;
   X->graph_steps = 0;
   drop_graph(X);
   hipEvent_t e0, e1;
   HIPOK(hipEventCreate(&e0));
   HIPOK(hipEventCreate(&e1));
   for (int c = 0; c < 4; c++) {
      if (X->step + steps_per_trial > X->maxsteps) break;
      X->path = cand[c][0];
      X->tile_rows = cand[c][1];
      HIPOK(hipEventRecord(e0, X->st));
      int rc = slg_batch_step(X, steps_per_trial, err, errlen);
      if (rc) return rc;
      HIPOK(hipEventRecord(e1, X->st));
      HIPOK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPOK(hipEventElapsedTime(&ms, e0, e1));
      const double per = ms / steps_per_trial;
      if (ms_out) ms_out[c] = per;
      if (per < best) {
         best = per;
         bi = c;
      }
   }
   X->path = cand[bi][0];
   X->tile_rows = cand[bi][1];
   // then the storm group of the fused tile (the trials above ran SL_SGROUP)
   if (X->path == SLG_PATH_FUSED && X->nstorm > 1) {
      const int groups[3] = {1, 8, 16};
      int bg = slg_batch_storm_group(X);
      for (int c = 0; c < 3; c++) {
         if (X->step + steps_per_trial > X->maxsteps) break;
         X->d32.sgroup = X->d64.sgroup = groups[c];
         HIPOK(hipEventRecord(e0, X->st));
         int rc = slg_batch_step(X, steps_per_trial, err, errlen);
         if (rc) return rc;
         HIPOK(hipEventRecord(e1, X->st));
         HIPOK(hipEventSynchronize(e1));
         float ms = 0.f;
         HIPOK(hipEventElapsedTime(&ms, e0, e1));
         if (ms / steps_per_trial < best) {
            best = ms / steps_per_trial;
            bg = groups[c];
         }
      }
      X->d32.sgroup = X->d64.sgroup = bg;
   }
   (void) hipEventDestroy(e0);
   (void) hipEventDestroy(e1);
   X->graph_steps = keep_graph;
   if (rows_out) *rows_out = X->tile_rows;
   if (path_out) *path_out = X->path;
   return 0;
}

extern "C" void slg_batch_close(slg_batch *X)
{
   if (!X) return;
   drop_graph(X);
   rex_free(X);
   if (X->hbmax) (void) hipFree(X->hbmax);
   if (X->hbmax2) (void) hipFree(X->hbmax2);
   if (X->pt_buf) (void) hipFree(X->pt_buf);
   for (int p = 0; p < SL_MAXSPLIT - 1; p++)
      if (X->stx[p]) {
         (void) hipStreamSynchronize(X->stx[p]);
         (void) hipStreamDestroy(X->stx[p]);
         (void) hipEventDestroy(X->ev_join[p]);
      }
   for (auto &a : X->ev_phase)
      for (auto &e : a)
         if (e) (void) hipEventDestroy(e);
   if (X->dmem) {
      (void) hipStreamSynchronize(X->st);
      (void) hipStreamDestroy(X->st);
      (void) hipFree(X->dmem);
   }
   if (X->d32.tidesh) (void) hipFree(X->d32.tidesh);
   if (X->d64.tidesh) (void) hipFree(X->d64.tidesh);
   if (X->tide_mem) (void) hipFree(X->tide_mem);
   delete X;
}

// --------------------------------------------------------------------------
// exports in the tagged-record format of the reference state dumps, so the
// host initialisation can be compared with the reference's COMMON blocks and
// the oracle can start from exactly the product's inputs.

static void rec(FILE *fp, const char *name, char type, long n, const void *data)
{
   char hdr[20];
   int esz = (type == 'd') ? 8 : (type == 'i') ? 4 : (type == 'h') ? 2 : 1;
   memset(hdr, 0, sizeof hdr);
   strncpy(hdr, name, 16);
   hdr[16] = type;
   fwrite(hdr, 1, 20, fp);
   fwrite(&n, 8, 1, fp);
   fwrite(data, (size_t) esz, (size_t) n, fp);
}

extern "C" int slg_export_init(const slg_batch *X, int storm, const char *path)
{
   const sl_basin *b = &X->b->h;
   const sl_storm *s = &X->sc[storm].s0;
   FILE *fp = fopen(path, "wb");
   if (!fp) return SLG_EINPUT;
   size_t np = (size_t) b->imxb * b->jmxb;
   int d3[6] = {b->imxb, b->jmxb, b->imxb1, b->jmxb1, b->imxb2, b->jmxb2};
   rec(fp, "DUMB3", 'i', 6, d3);
   rec(fp, "ZB", 'd', (long) np, b->zb);
   rec(fp, "ZBM", 'd', (long) np, b->zbm);
   rec(fp, "ITREE", 'c', (long) np, b->itree);
   rec(fp, "IMANN", 'i', (long) np, b->imann);
   rec(fp, "KSKP", 'c', (long) np, b->kskp);
   rec(fp, "YLT", 'd', (long) np, b->ylt);
   std::vector<double> z(np, 0.), w;
   if (X->hb0.empty()) {
      w.resize(np);
      sl_initial_water(b, &X->sc[storm].s0, w.data());
   }
   const double *hb = X->hb0.empty() ? w.data() : &X->hb0[(size_t) storm * np];
   rec(fp, "HB", 'd', (long) np, hb);
   rec(fp, "UB", 'd', (long) np, z.data());
   rec(fp, "VB", 'd', (long) np, z.data());
   std::vector<double> hmx(np, 0.);   // HMX = HB on squares only (intlht.f:172-177)
   for (int j = 1; j <= b->jmxb1; j++)
      for (int i = 1; i <= b->imxb1; i++) hmx[PIX(b, i, j)] = hb[PIX(b, i, j)];
   rec(fp, "HMX", 'd', (long) np, hmx.data());
   rec(fp, "ELPCT", 'd', b->imxb, b->elpct);
   rec(fp, "ELPDT", 'd', b->imxb, b->elpdt);
   rec(fp, "ELPCL", 'd', b->imxb, b->elpcl);
   rec(fp, "ELPDL", 'd', b->imxb, b->elpdl);
   rec(fp, "ELPCL2", 'd', b->imxb, b->elpcl2);
   rec(fp, "ELPDL2", 'd', b->imxb, b->elpdl2);
   rec(fp, "SINL2", 'd', b->jmxb, b->sinl2);
   rec(fp, "SINT2", 'd', b->jmxb, b->sint2);
   rec(fp, "SINT", 'd', b->jmxb, b->sint);
   rec(fp, "COST", 'd', b->jmxb, b->cost);
   rec(fp, "SINL", 'd', b->jmxb, b->sinl);
   rec(fp, "COSL", 'd', b->jmxb, b->cosl);
   rec(fp, "IS", 'i', b->jmxb, b->is);
   rec(fp, "MS", 'i', b->jmxb, b->ms);
   rec(fp, "MF", 'i', b->jmxb, b->mf);
   rec(fp, "IE", 'i', b->jmxb, b->ie);
   rec(fp, "ME", 'i', b->jmxb, b->me);
   rec(fp, "X1B", 'd', 10, b->x1b);
   rec(fp, "CORL", 'd', 90, b->corl);
   rec(fp, "AR2", 'd', 600 * 245, b->ar2);
   rec(fp, "AI2", 'd', 600 * 245, b->ai2);
   rec(fp, "BR2", 'd', 600 * 245, b->br2);
   rec(fp, "BI2", 'd', 600 * 245, b->bi2);
   rec(fp, "CR2", 'd', 600 * 245, b->cr2);
   rec(fp, "CI2", 'd', 600 * 245, b->ci2);
   double gam[6] = {b->gama, b->gama1, b->gamfp[0], b->gamfp[1], b->gamfp[2], b->gamfp[3]};
   rec(fp, "GAMAF", 'd', 6, gam);
   double egth[4] = {5280., b->delt, 32.2, b->cor};
   rec(fp, "EGTH", 'd', 4, egth);
   double bsn[7] = {b->phi, b->alto, b->alno, b->phi1, b->alt1, b->aln1, b->alt1c};
   rec(fp, "BSN", 'd', 7, bsn);
   double pol[4] = {b->degree, b->rmouth, b->azmth, b->dela};
   rec(fp, "POLAR", 'd', 4, pol);
   double ell[3] = {b->aaxis, b->baxis, b->abqab};
   rec(fp, "ELLIP", 'd', 3, ell);
   double org[2] = {b->xmouth, b->ymouth};
   rec(fp, "ORIGN", 'd', 2, org);
   double dat[2] = {s->seadtm, s->dtmlak};
   rec(fp, "DATUM", 'd', 2, dat);
   double dlt[4] = {b->dltin[0], b->dltin[1], b->dltin[2], s->dltb};
   rec(fp, "DLTDTA", 'd', 4, dlt);
   rec(fp, "HTER", 'd', 1, &b->hter);
   rec(fp, "NBCPT", 'i', 1, &b->nbcpt);
   rec(fp, "ISH", 'i', b->nbcpt, b->ish);
   rec(fp, "JSH", 'i', b->nbcpt, b->jsh);
   rec(fp, "TIDESH", 'd', b->nbcpt, b->tidesh);
   int ints[21] = {s->ic12, s->ic19, s->itime, s->mhalt, b->iver, b->iprj, b->ismth, s->key1,
                   0, s->jhr, s->itmadv, s->nhrad, s->ibgnt, s->itend, s->ndlthr, s->ndlth2,
                   b->nsqrs, b->nsqrw, b->nsqrwc, b->npss, b->ncut};
   rec(fp, "INTS", 'i', 21, ints);
   double reals[5] = {s->bt, s->dlthr, s->dlth2, s->treal, s->zlato};
   rec(fp, "REALS", 'd', 5, reals);
   char flags[8] = {b->dollar[0], b->dollar[1], b->ebsn, b->ebsn1, b->ebsn2, b->nof1d, ' ', ' '};
   rec(fp, "FLAGS", 'c', 8, flags);
   const int nl = b->nsqrwc, nc = b->ncut;
   rec(fp, "EDGE_I", 'i', nl, b->isqr);
   rec(fp, "EDGE_J", 'i', nl, b->jsqr);
   rec(fp, "EDGE_SIDE", 'i', nl, b->iside);
   rec(fp, "KTREE", 'c', nl, b->ktree);
   rec(fp, "HWEIR", 'd', nl, b->hweir);
   rec(fp, "ZBMIN", 'd', nl, b->zbmin);
   rec(fp, "BANK1", 'd', nl, b->bank1);
   rec(fp, "BANK2", 'd', nl, b->bank2);
   rec(fp, "DELCUT", 'd', nl, b->delcut);
   rec(fp, "COS1D", 'd', nl, b->cos1d);
   rec(fp, "SIN1D", 'd', nl, b->sin1d);
   rec(fp, "CUTL", 'd', nc, b->cutl);
   rec(fp, "CUTLI", 'd', nc, b->cutli);
   rec(fp, "CUTLE", 'd', nc, b->cutle);
   {
      int pts[40];   // IPT0(2,5) JPT0 IPTL JPTL (COMMON /CHNL/)
      for (int k = 0; k < 5; k++)
         for (int e = 0; e < 2; e++) {
            pts[0 + k * 2 + e] = b->ipt0[k][e];
            pts[10 + k * 2 + e] = b->jpt0[k][e];
            pts[20 + k * 2 + e] = b->iptl[k][e];
            pts[30 + k * 2 + e] = b->jptl[k][e];
         }
      rec(fp, "CHNL_PTS", 'i', 40, pts);
      double cd[15], ca[10], cn[5] = {b->aman, b->bman, b->gb, b->agb, b->agb13};
      for (int k = 0; k < 5; k++) {
         cd[k] = b->chlnh[k];
         cd[5 + k] = b->chwth[k];
         cd[10 + k] = b->chdph[k];
         ca[k] = b->aresq0[k];
         ca[5 + k] = b->aresql[k];
      }
      rec(fp, "CHNLD", 'd', 15, cd);
      rec(fp, "CHNARA", 'd', 10, ca);
      rec(fp, "CHNLNO", 'd', 5, cn);
   }
   rec(fp, "STA", 'c', 16, b->sta);
   double strmsb[8] = {s->c1, s->c2, s->c21, s->c22, s->ax, s->ay, s->ptency, s->rtency};
   rec(fp, "STRMSB", 'd', 8, strmsb);
   double wtemp[5] = {s->ydelp, s->zdelp, s->pnn, s->yc24, s->zc24};
   rec(fp, "WTEMP", 'd', 5, wtemp);
   rec(fp, "TRK_X", 'd', 999, s->X + 1);
   rec(fp, "TRK_Y", 'd', 999, s->Y + 1);
   rec(fp, "TRK_PT", 'd', 999, s->PT + 1);
   rec(fp, "TRK_R", 'd', 999, s->R + 1);
   rec(fp, "TRK_DIR", 'd', 999, s->DIR + 1);
   rec(fp, "TRK_SP", 'd', 999, s->SP + 1);
   rec(fp, "S", 'd', 800, s->S + 1);
   rec(fp, "C", 'd', 800, s->C + 1);
   rec(fp, "P", 'd', 800, s->P + 1);
   rec(fp, "DELP", 'd', 800, s->DELP + 1);
   rec(fp, "SW", 'd', 800, s->SW + 1);
   rec(fp, "CW", 'd', 800, s->CW + 1);
   rec(fp, "W1218", 'd', 2, s->w1218);
   rec(fp, "WMAX", 'd', 2, s->wmax);
   rec(fp, "X12", 'd', 50, s->x12 + 1);
   fclose(fp);
   return 0;
}

extern "C" int slg_export_storm(const slg_batch *X, int storm, const char *rows_path,
                                const char *tables_path)
{
   const storm_sched &S = X->sc[storm];
   FILE *fp = fopen(rows_path, "wb");
   if (!fp) return SLG_EINPUT;
   fwrite(S.rows.data(), sizeof(double), S.rows.size(), fp);
   fclose(fp);
   fp = fopen(tables_path, "wb");
   if (!fp) return SLG_EINPUT;
   // the table sets in double, replayed from the schedule
   sl_storm s = S.s0;
   int gen = s.table_gen;
   for (int n = 1; n <= S.nsteps; n++) {
      sl_storm_ttimer(&s);
      sl_storm_stmval(&s, &X->b->h);
      if (s.table_gen != gen) {
         gen = s.table_gen;
         rec(fp, "STEP", 'i', 1, &n);
         rec(fp, "S", 'd', 800, s.S + 1);
         rec(fp, "C", 'd', 800, s.C + 1);
         rec(fp, "P", 'd', 800, s.P + 1);
         rec(fp, "DELP", 'd', 800, s.DELP + 1);
         rec(fp, "SW", 'd', 800, s.SW + 1);
         rec(fp, "CW", 'd', 800, s.CW + 1);
         rec(fp, "W1218", 'd', 2, s.w1218);
         rec(fp, "WMAX", 'd', 2, s.wmax);
         rec(fp, "X12", 'd', 50, s.x12 + 1);
      }
   }
   fclose(fp);
   return 0;
}
