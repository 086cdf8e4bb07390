/*
 * archive.c -- gauge configurations for the ms10 program: openQCD's exported file
 * format (import_cnfg / export_cnfg, as called at main/ms10.c:712-717) and a
 * seeded synthetic generator on the GPU.
 *
 * The exported format is written from openQCD-1.4's documentation of
 * archive.c (not part of this repository): header of 4 int32 lattice sizes
 * and the average plaquette, then the 8 links stored at each odd point, odd
 * points in lexicographic order, little endian.
 *
 * The imported (per-process) format of write_cnfg / read_cnfg, used by ms10
 * -noexp (main/ms10.c:706-713), is the memory image of one process's field:
 * 16 native ints (the process grid NPROC0..3, the local sizes L0..3, the
 * process block sizes and the process coordinates cpr[0..3]), the local
 * plaquette sum (double) and the 4*VOLUME links of udfld() in its own layout,
 * native byte order.  Also written from openQCD-1.4's documentation; with one
 * process owning the lattice the grid is 1x1x1x1.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "utils.h"
#include "internal.h"

double avg_plaq(void)
{
   double s = lnk_plaq_sum(udfld());
   double np = (bc_type() == 0) ? (double)(6 * L0 - 3) * L1 * L2 * L3 : 6.0 * VOLUME;
   return s / (3.0 * np);
}

void cnfg_sizes(const char *in, int n[4])
{
   FILE *f = fopen(in, "rb");
   error_root(f == NULL, 1, "cnfg_sizes [archive.c]", "Unable to open configuration file %s", in);
   stdint_t h[4];
   error_root(fread(h, sizeof(stdint_t), 4, f) != 4, 1, "cnfg_sizes [archive.c]",
              "Incorrect read count (%s)", in);
   fclose(f);
   if (endianness() == BIG_ENDIAN) bswap_int(4, h);
   for (int m = 0; m < 4; m++) n[m] = (int)h[m];
}

/* odd points in lexicographic order, each with its 8 stored links */
static void odd_block(int l, su3_dble *dst, su3_dble *ub, int to_field)
{
   int ix = ipt[l];
   su3_dble *p = ub + 8 * (ix - VOLUME / 2);
   if (to_field)
      memcpy(p, dst, 8 * sizeof(su3_dble));
   else
      memcpy(dst, p, 8 * sizeof(su3_dble));
}

void export_cnfg(const char *out)
{
   su3_dble *ud = udfld();
   double plaq = avg_plaq();
   ud = udfld();
   FILE *f = fopen(out, "wb");
   error_root(f == NULL, 1, "export_cnfg [archive.c]", "Unable to open output file %s", out);
   stdint_t h[4] = {L0, L1, L2, L3};
   double pl = plaq;
   int be = (endianness() == BIG_ENDIAN);
   if (be) {
      bswap_int(4, h);
      bswap_double(1, &pl);
   }
   size_t iw = fwrite(h, sizeof(stdint_t), 4, f);
   iw += fwrite(&pl, sizeof(double), 1, f);
   su3_dble blk[8];
   for (int l = 0; l < VOLUME; l++) {
      if (ipt[l] < VOLUME / 2) continue;
      odd_block(l, blk, ud, 0);
      if (be) bswap_double(8 * 18, blk);
      iw += fwrite(blk, sizeof(su3_dble), 8, f);
   }
   fclose(f);
   error_root(iw != 5 + 8 * (size_t)(VOLUME / 2), 1, "export_cnfg [archive.c]", "Incorrect write count");
}

void import_cnfg(const char *in)
{
   int n[4];
   cnfg_sizes(in, n);
   error_root(n[0] != L0 || n[1] != L1 || n[2] != L2 || n[3] != L3, 1, "import_cnfg [archive.c]",
              "Lattice sizes of %s (%dx%dx%dx%d) do not match", in, n[0], n[1], n[2], n[3]);
   FILE *f = fopen(in, "rb");
   error_root(f == NULL, 1, "import_cnfg [archive.c]", "Unable to open configuration file %s", in);
   stdint_t h[4];
   double plaq0;
   size_t ir = fread(h, sizeof(stdint_t), 4, f);
   ir += fread(&plaq0, sizeof(double), 1, f);
   int be = (endianness() == BIG_ENDIAN);
   if (be) bswap_double(1, &plaq0);
   su3_dble *ud = udfld();
   su3_dble blk[8];
   for (int l = 0; l < VOLUME; l++) {
      if (ipt[l] < VOLUME / 2) continue;
      ir += fread(blk, sizeof(su3_dble), 8, f);
      if (be) bswap_double(8 * 18, blk);
      odd_block(l, blk, ud, 1);
   }
   fclose(f);
   error_root(ir != 5 + 8 * (size_t)(VOLUME / 2), 1, "import_cnfg [archive.c]",
              "Incorrect read count (%s)", in);
   set_flags(UPDATED_UD);
   set_bc();
   double plaq1 = avg_plaq();
   double eps = sqrt(6.0 * (double)VOLUME) * DBL_EPSILON;
   error_root(fabs(plaq1 - plaq0) > eps, 1, "import_cnfg [archive.c]",
              "Incorrect average plaquette in %s", in);
}

void random_cnfg(unsigned long long seed, double eps)
{
   wlg_ctx *c = wl_gauge_ctx();
   wl_check(wlg_generate(c, (uint64_t)seed, eps), "random_cnfg [archive.c]");
   wl_device_changed();
}

/* header of an imported (per-process) configuration file */
static void imported_head(int ldat[16])
{
   int h[16] = {NPROC0, NPROC1, NPROC2, NPROC3, L0, L1, L2, L3,
                NPROC0_BLK, NPROC1_BLK, NPROC2_BLK, NPROC3_BLK, cpr[0], cpr[1], cpr[2], cpr[3]};
   memcpy(ldat, h, sizeof(h));
}

void write_cnfg(const char *out)
{
   su3_dble *ud = udfld();
   double plaq = lnk_plaq_sum(ud);
   ud = udfld();
   int ldat[16];
   imported_head(ldat);
   FILE *f = fopen(out, "wb");
   error_root(f == NULL, 1, "write_cnfg [archive.c]", "Unable to open output file %s", out);
   size_t iw = fwrite(ldat, sizeof(int), 16, f);
   iw += fwrite(&plaq, sizeof(double), 1, f);
   iw += fwrite(ud, sizeof(su3_dble), 4 * (size_t)VOLUME, f);
   fclose(f);
   error_root(iw != 17 + 4 * (size_t)VOLUME, 1, "write_cnfg [archive.c]", "Incorrect write count");
}

void cnfg_sizes_imported(const char *in, int n[4])
{
   FILE *f = fopen(in, "rb");
   error_root(f == NULL, 1, "cnfg_sizes_imported [archive.c]", "Unable to open configuration file %s",
              in);
   int ldat[16];
   size_t ir = fread(ldat, sizeof(int), 16, f);
   fclose(f);
   error_root(ir != 16, 1, "cnfg_sizes_imported [archive.c]", "Incorrect read count (%s)", in);
   error_root(ldat[0] != 1 || ldat[1] != 1 || ldat[2] != 1 || ldat[3] != 1, 1,
              "cnfg_sizes_imported [archive.c]",
              "%s was written on a %dx%dx%dx%d process grid; one process owns the lattice here", in,
              ldat[0], ldat[1], ldat[2], ldat[3]);
   for (int m = 0; m < 4; m++) n[m] = ldat[4 + m];
}

void read_cnfg(const char *in)
{
   FILE *f = fopen(in, "rb");
   error_root(f == NULL, 1, "read_cnfg [archive.c]", "Unable to open configuration file %s", in);
   int ldat[16], want[16];
   imported_head(want);
   size_t ir = fread(ldat, sizeof(int), 16, f);
   error_root(ir != 16 || memcmp(ldat, want, sizeof(want)) != 0, 1, "read_cnfg [archive.c]",
              "Lattice or process grid of %s does not match", in);
   double plaq0;
   ir += fread(&plaq0, sizeof(double), 1, f);
   su3_dble *ud = udfld();
   ir += fread(ud, sizeof(su3_dble), 4 * (size_t)VOLUME, f);
   fclose(f);
   error_root(ir != 17 + 4 * (size_t)VOLUME, 1, "read_cnfg [archive.c]", "Incorrect read count (%s)",
              in);
   set_flags(UPDATED_UD);
   double plaq1 = lnk_plaq_sum(udfld());
   double eps = sqrt(6.0 * (double)VOLUME) * DBL_EPSILON * fabs(plaq0);
   error_root(fabs(plaq1 - plaq0) > eps, 1, "read_cnfg [archive.c]", "Incorrect plaquette sum in %s",
              in);
}

