/*
 * wloop.h -- Wilson-loop module interface (reference include/wloop.h:13-66),
 * served by libwloop_amd.so; the measurement itself runs on the GPU.
 *
 *   wloop_sum        wloop.c:611-688   W_ij(R,T;x0) of the current udfld() into wl.wl_mat
 *   free_wloop       wloop.c:511-521
 *   smear            smear.c:99-121    operator levels (cumulative nss)
 *   get_pus          smear.c:124-127   level pointer table, pindx order (host mirror)
 *   free_smear       smear.c:130-141
 *   set_wloop_parms  wloop_parms.c:96-194  parameter data base
 *   wloop_parms      wloop_parms.c:197-200
 *   read_wloop_parms wloop_parms.c:203-285 [Wloop] section of stdin
 *   print_wloop_parms wloop_parms.c:288-340
 *   write_wloop_parms wloop_parms.c:343-382 binary .par record
 *   check_wloop_parms wloop_parms.c:385-429
 *   wl_idx           wloop_parms.c:431-434 i + msl(j + msl(t + mwlt(r + mwlr x0)))
 */
#ifndef WLOOP_H
#define WLOOP_H

#include <stdio.h>
#include "global.h"
#include "su3.h"

#define MAX_SM_LEVELS (5)
#define SLICE0 (L1 * L2 * L3)

#define wloop_RELEASE "wloop-1.4"

typedef enum
{
   APE, HYP,
   OP_SMEARS
} op_smear_t;

typedef struct
{
   int msl, nss[MAX_SM_LEVELS];
   int mwlt, mwlr, wls;
   op_smear_t op_smear;
   double alpha_action[3];
   double alpha_ape;
   double alpha_hyp[2];
   int proj;
   int proj_iter;
   double *wl_mat;
} wloop_parms_t;

#define pindx(sl, ix, k) ((k) + 3 * ((ix) + VOLUME * (sl)))
#define pindxs(sl, ixs, x0, k) ((k) + 3 * ((ipt[(ixs) + SLICE0 * (x0)]) + VOLUME * (sl)))

#ifdef __cplusplus
extern "C" {
#endif

extern void free_wloop(void);
extern void wloop_sum(int keep);

extern wloop_parms_t set_wloop_parms(int msl, int nss[], int mwlt, int mwlr, op_smear_t op_smear,
                                     double alpha_action[3], double alpha_ape, double alpha_hyp[2],
                                     int proj_iter);
extern wloop_parms_t wloop_parms(void);
extern void read_wloop_parms(void);
extern void print_wloop_parms(void);
extern void write_wloop_parms(FILE *fdat);
extern void check_wloop_parms(FILE *fdat);
extern int wl_idx(int sl1, int sl2, int t, int r, int x0);

extern void smear(void);
extern su3_dble **get_pus(void);
extern void free_smear(void);

/* No reference counterpart: arithmetic of wloop_sum and smear on the GPU,
 * 0 = plain FP64 (default, the reference's precision), 1 = FP32 staples and
 * transporter products with FP64 projection and FP64 sums. */
extern void set_wloop_precision(int precision);

#ifdef __cplusplus
}
#endif
#endif
