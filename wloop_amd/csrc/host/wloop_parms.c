/*
 * wloop_parms.c -- Wilson-loop parameter data base (wloop.h), following
 * modules/wloop/wloop_parms.c: validation (:156-177), proj = (proj_iter > 0)
 * (:171), wls = msl^2 mwlt mwlr N0 (:182), wl_mat pre-filled with 1.0
 * (:184-191), the [Wloop] section (:203-285), the binary .par record
 * msl+4 int32 then 6 doubles, little endian (:343-382), wl_idx (:431-434).
 */
#include <stdlib.h>
#include <string.h>

#include "global.h"
#include "lattice.h"
#include "utils.h"
#include "wloop.h"

static wloop_parms_t wl = {1, {0}, 0, 0, 0, OP_SMEARS, {0.0}, 0.0, {0.0}, 0, 0, NULL};
static double *wlm = NULL;

wloop_parms_t set_wloop_parms(int msl, int nss[], int mwlt, int mwlr, op_smear_t op_smear,
                              double alpha_action[3], double alpha_ape, double alpha_hyp[2],
                              int proj_iter)
{
   error_root((op_smear != APE) && (op_smear != HYP), 1, "set_wloop_parms [wloop_parms.c]",
              "Unknown type of operator smearing");
   error_root((msl < 1) || (msl > MAX_SM_LEVELS), 1, "set_wloop_parms [wloop_parms.c]",
              "max smearing level out of range");
   error_root(VOLUME == 0, 1, "set_wloop_parms [wloop_parms.c]", "Lattice sizes are not set");
   for (int k = 0; k < msl; k++)
      error_root(nss[k] < 0, 1, "set_wloop_parms [wloop_parms.c]",
                 "Number of smearing steps must be non-negative");

   wl.msl = msl;
   for (int k = 0; k < MAX_SM_LEVELS; k++) wl.nss[k] = (k < msl) ? nss[k] : 0;
   wl.alpha_ape = alpha_ape;
   wl.alpha_hyp[0] = alpha_hyp[0];
   wl.alpha_hyp[1] = alpha_hyp[1];
   for (int k = 0; k < 3; k++) wl.alpha_action[k] = alpha_action[k];
   wl.proj = (proj_iter > 0);
   wl.proj_iter = proj_iter;

   error_root((mwlt < 1) || (mwlt > NPROC0 * L0 / 2 + 1), 1, "set_wloop_parms [wloop_parms.c]",
              "max wloop time extension out of range");
   error_root((mwlr < 1) || (mwlr > NPROC1 * L1 / 2 + 1) || (mwlr > NPROC2 * L2 / 2 + 1) ||
              (mwlr > NPROC3 * L3 / 2 + 1), 1, "set_wloop_parms [wloop_parms.c]",
              "max wloop spatial extension out of range");

   wl.mwlt = mwlt;
   wl.mwlr = mwlr;
   wl.op_smear = op_smear;
   wl.wls = wl.msl * wl.msl * wl.mwlt * wl.mwlr * L0 * NPROC0;

   free(wlm);
   wlm = malloc(sizeof(double) * (size_t)wl.wls);
   error_root(wlm == NULL, 1, "set_wloop_parms [wloop_parms.c]", "Could not allocate wloop matrix");
   for (int i = 0; i < wl.wls; i++) wlm[i] = 1.0;
   wl.wl_mat = wlm;
   return wl;
}

wloop_parms_t wloop_parms(void) { return wl; }

void read_wloop_parms(void)
{
   int msl, mwlt, mwlr, proj_iter, idr;
   int nss[MAX_SM_LEVELS] = {0};
   double alpha_ape = 0.0, alpha_hyp[2] = {0.0, 0.0}, alpha_action[3];
   char line[NAME_SIZE];
   const op_smear_t kinds[] = {APE, HYP};

   find_section("Wloop");
   read_line("op_smear", "%127s", line);
   if (strcmp(line, "APE") == 0)
      idr = 0;
   else if (strcmp(line, "HYP") == 0)
      idr = 1;
   else {
      idr = 2;
      error_root(1, 1, "read_wloop_parms [wloop_parms.c]", "Unknown operator smearing %s", line);
   }
   read_line("mwlt", "%d", &mwlt);
   read_line("mwlr", "%d", &mwlr);
   read_line("proj_iter", "%d", &proj_iter);
   read_dprms("alpha_action", 3, alpha_action);
   read_line("msl", "%d", &msl);
   error_root((msl < 1) || (msl > MAX_SM_LEVELS), 1, "read_wloop_parms [wloop_parms.c]",
              "max smearing level out of range");
   read_iprms("nss", msl, nss);
   if (idr == 0)
      read_line("alpha", "%lf", &alpha_ape);
   else
      read_dprms("alpha", 2, alpha_hyp);

   set_wloop_parms(msl, nss, mwlt, mwlr, kinds[idr], alpha_action, alpha_ape, alpha_hyp, proj_iter);
}

static void print_double(double x)
{
   int n = fdigits(x);
   printf(" %.*f", IMAX(n, 1), x);
}

void print_wloop_parms(void)
{
   printf("Wloop parameters:\n");
   printf("msl = %d\n", wl.msl);
   printf("nss =");
   for (int i = 0; i < wl.msl; i++) printf(" %d", wl.nss[i]);
   printf("\n");
   printf("mwlt = %d\n", wl.mwlt);
   printf("mwlr = %d\n", wl.mwlr);
   printf("alpha_action =");
   for (int i = 0; i < 3; i++) print_double(wl.alpha_action[i]);
   printf("\n");
   printf("proj_iter = %d\n", wl.proj_iter);
   printf("operator smearing = ");
   if (wl.op_smear == APE) {
      printf("APE\n");
      printf("alpha =");
      print_double(wl.alpha_ape);
      printf("\n");
   } else if (wl.op_smear == HYP) {
      printf("HYP\n");
      printf("alpha =");
      for (int i = 0; i < 2; i++) print_double(wl.alpha_hyp[i]);
      printf("\n");
   }
   printf("\n");
}

static void parms_record(stdint_t *istd, double *dstd)
{
   for (int i = 0; i < wl.msl; i++) istd[i] = (stdint_t)wl.nss[i];
   istd[wl.msl + 0] = (stdint_t)wl.mwlt;
   istd[wl.msl + 1] = (stdint_t)wl.mwlr;
   istd[wl.msl + 2] = (stdint_t)wl.op_smear;
   istd[wl.msl + 3] = (stdint_t)wl.proj_iter;
   dstd[0] = wl.alpha_ape;
   dstd[1] = wl.alpha_hyp[0];
   dstd[2] = wl.alpha_hyp[1];
   dstd[3] = wl.alpha_action[0];
   dstd[4] = wl.alpha_action[1];
   dstd[5] = wl.alpha_action[2];
}

void write_wloop_parms(FILE *fdat)
{
   stdint_t istd[MAX_SM_LEVELS + 4];
   double dstd[6];
   parms_record(istd, dstd);
   if (endianness() == BIG_ENDIAN) {
      bswap_int(wl.msl + 4, istd);
      bswap_double(6, dstd);
   }
   size_t iw = fwrite(istd, sizeof(stdint_t), (size_t)wl.msl + 4, fdat);
   iw += fwrite(dstd, sizeof(double), 6, fdat);
   error_root(iw != (size_t)wl.msl + 10, 1, "write_wloop_parms [wloop_parms.c]", "Incorrect write count");
}

void check_wloop_parms(FILE *fdat)
{
   stdint_t istd[MAX_SM_LEVELS + 4], ref_i[MAX_SM_LEVELS + 4];
   double dstd[6], ref_d[6];
   size_t ir = fread(istd, sizeof(stdint_t), (size_t)wl.msl + 4, fdat);
   ir += fread(dstd, sizeof(double), 6, fdat);
   error_root(ir != (size_t)wl.msl + 10, 1, "check_wloop_parms [wloop_parms.c]", "Incorrect read count");
   if (endianness() == BIG_ENDIAN) {
      bswap_int(wl.msl + 4, istd);
      bswap_double(6, dstd);
   }
   parms_record(ref_i, ref_d);
   int ie = 0;
   for (int i = 0; i < wl.msl + 4; i++) ie |= (istd[i] != ref_i[i]);
   for (int i = 0; i < 6; i++) ie |= (dstd[i] != ref_d[i]);
   error_root(ie != 0, 1, "check_wloop_parms [wloop_parms.c]", "Parameters do not match");
}

int wl_idx(int sl1, int sl2, int t, int r, int x0)
{
   return sl1 + wl.msl * (sl2 + wl.msl * (t + wl.mwlt * (r + wl.mwlr * (cpr[0] * L0 + x0))));
}
