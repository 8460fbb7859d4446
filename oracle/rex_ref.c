/* rex_ref.c -- drives the reference's own rex writer and datum override.
 *
 * Test infrastructure only: it links the reference's C objects as they are
 * built from the reference sources (oracle/Makefile target _ref/rexref64) and
 * calls
 *   RexOpen / RexSaveHeader / RexSaveStep / RexSaveEnv / RexClose
 *     (sorc/slosh.fd/rex.c:150-180, 264-379, 387-499, 505-650), and
 *   InitWater_TideModeOverride (sorc/slosh.fd/slosh2.c:408-470)
 * with inputs from a spec file, so that the oracle's frame encoder and the
 * product's rex writer can be checked byte for byte against the reference.
 *
 *   rexref64 rex <spec> <out.rex>
 *     spec, little-endian: int32 imxb, jmxb, nframes; char comment[201];
 *     char abbrev[5]; double zb[jmxb][imxb]; then per frame double track[6]
 *     (lat, lon, wind speed, wind direction, delp, size), double clock,
 *     double hb[jmxb][imxb]; then char trk[257] and double hb_env[jmxb][imxb]
 *     (the envelope frame).
 *   rexref64 datum <in> <out>
 *     float pairs (ht1, ht2) in, the pairs after the surge-only override
 *     (f_tide 0) out.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define BAS_X 999
#define BAS_Y 1655

/* rexType as rex.h declares it (the reference's default build) */
typedef struct {
   FILE *fp;
   int32_t Offset;
   int32_t trkOffset;
   char version;
   char comment[201];
   char bsnAbrev[5];
} rexType;

int RexOpen(rexType *rex, const char *rexName, char version);
void RexSaveHeader(rexType *rex, int imxb, int jmxb, const char comment[201], const char bsnAbrev[5]);
void RexClose(rexType *rex);
int RexSaveStep(rexType *rex, double stormLat, double stormLon, double wspeed, double wdirect, double delp,
                double size2, double hb[BAS_Y][BAS_X], double zb[BAS_Y][BAS_X], int imxb, int jmxb,
                double clock);
int RexSaveEnv(rexType *rex, const char *trkName, double hb[BAS_Y][BAS_X], double zb[BAS_Y][BAS_X], int imxb,
               int jmxb, int f_tide);
int InitWater_TideModeOverride(float *ht1, int f_tide, float *ht2);

static double HB[BAS_Y][BAS_X], ZB[BAS_Y][BAS_X];

static int rd(void *p, size_t n, FILE *fp)
{
   return fread(p, 1, n, fp) == n ? 0 : -1;
}

static int rd_field(double f[BAS_Y][BAS_X], int imxb, int jmxb, FILE *fp)
{
   for (int j = 0; j < jmxb; j++)
      if (rd(f[j], sizeof(double) * (size_t) imxb, fp)) return -1;
   return 0;
}

static int do_rex(const char *spec, const char *out)
{
   FILE *fp = fopen(spec, "rb");
   if (!fp) return 1;
   int32_t hdr[3];
   char comment[201], abbrev[5], trk[257];
   if (rd(hdr, sizeof hdr, fp) || rd(comment, 201, fp) || rd(abbrev, 5, fp)) return 1;
   const int imxb = hdr[0], jmxb = hdr[1], nf = hdr[2];
   if (imxb < 2 || imxb > BAS_X || jmxb < 2 || jmxb > BAS_Y) return 1;
   if (rd_field(ZB, imxb, jmxb, fp)) return 1;
   rexType rex;
   memset(&rex, 0, sizeof rex);
   if (RexOpen(&rex, out, 2)) return 1;
   RexSaveHeader(&rex, imxb, jmxb, comment, abbrev);
   for (int k = 0; k < nf; k++) {
      double t[7];
      if (rd(t, sizeof t, fp) || rd_field(HB, imxb, jmxb, fp)) return 1;
      if (RexSaveStep(&rex, t[0], t[1], t[2], t[3], t[4], t[5], HB, ZB, imxb, jmxb, t[6])) return 1;
   }
   if (rd(trk, 257, fp) || rd_field(HB, imxb, jmxb, fp)) return 1;
   fclose(fp);
   if (RexSaveEnv(&rex, trk, HB, ZB, imxb, jmxb, 0)) return 1;
   RexClose(&rex);
   return 0;
}

static int do_datum(const char *in, const char *out)
{
   FILE *fi = fopen(in, "rb"), *fo = fopen(out, "wb");
   if (!fi || !fo) return 1;
   float v[2];
   while (fread(v, sizeof(float), 2, fi) == 2) {
      InitWater_TideModeOverride(&v[0], 0, &v[1]);
      if (fwrite(v, sizeof(float), 2, fo) != 2) return 1;
   }
   fclose(fi);
   return fclose(fo) ? 1 : 0;
}

int main(int argc, char **argv)
{
   if (argc == 4 && !strcmp(argv[1], "rex")) return do_rex(argv[2], argv[3]);
   if (argc == 4 && !strcmp(argv[1], "datum")) return do_datum(argv[2], argv[3]);
   fprintf(stderr, "usage: %s rex <spec> <out.rex> | datum <in> <out>\n", argv[0]);
   return 2;
}
