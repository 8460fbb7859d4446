/* orc_rex.c -- CPU restatement of the rex v2 frame encoding.
 *
 * Test infrastructure only: the checker of the device encoder
 * (slg_batch_rex_frames).  It is pinned byte for byte against the
 * reference's own writer (oracle/rex_ref.c) by tests/test_rex.py.
 *
 * One frame is the basin loop of RexSaveStep / RexSaveEnv
 * (sorc/slosh.fd/rex.c:453-499, 554-620): cells I-outer, J-inner over
 * (imxb-1) x (jmxb-1); the value (int)(HB*10+.5), 999 where HB+ZB == 0,
 * clipped to [-320, 700] (999 kept); run-length records of Stuff2_xxx
 * (pack.c:166-212): a run of up to 256 equal values as 8 bits (run-1) and
 * 10 bits (value+320, 999 as 701), written most significant bit first as
 * fileBitWrite does (tendian.c:848-940); the last byte padded with zeros.
 */
#include <stddef.h>

struct orc_bits {
   unsigned char *out;
   long n;          /* whole bytes written */
   unsigned acc;    /* bits of the byte being filled */
   int nacc;
};

static void put_bits(struct orc_bits *w, unsigned v, int nbits)
{
   for (int b = nbits - 1; b >= 0; b--) {
      w->acc = (w->acc << 1) | ((v >> b) & 1u);
      if (++w->nacc == 8) {
         w->out[w->n++] = (unsigned char) w->acc;
         w->acc = 0;
         w->nacc = 0;
      }
   }
}

/* rex.c:466-477: the value stored for one cell */
static int cell_value(double hb, double zb)
{
   int v = (hb + zb == 0.0) ? 999 : (int) ((hb * 10) + .5);
   if (v < -320) v = -320;
   else if (v > 700 && v != 999) v = 700;
   return v;
}

/* hb, zb: [jmxb][imxb], I fastest.  Returns the frame's length in bytes. */
long orc_rex_frame(const double *hb, const double *zb, int imxb, int jmxb, unsigned char *out)
{
   struct orc_bits w = {out, 0, 0u, 0};
   int prev = -1, run = 0;
   for (int i = 0; i < imxb - 1; i++)
      for (int j = 0; j < jmxb - 1; j++) {
         const size_t p = (size_t) j * (size_t) imxb + (size_t) i;
         const int v = cell_value(hb[p], zb[p]);
         const int c = (v == 999 ? 701 : v) + 320;
         if (c == prev) {
            if (++run == 256) {
               put_bits(&w, 255u, 8);
               put_bits(&w, (unsigned) prev, 10);
               run = 0;
            }
         } else {
            if (run) {
               put_bits(&w, (unsigned) (run - 1), 8);
               put_bits(&w, (unsigned) prev, 10);
            }
            prev = c;
            run = 1;
         }
      }
   if (run) {
      put_bits(&w, (unsigned) (run - 1), 8);
      put_bits(&w, (unsigned) prev, 10);
   }
   if (w.nacc) put_bits(&w, 0u, 8 - w.nacc);
   return w.n;
}
