/*
 * wl_oracle.h -- CPU FP64 restatement of the wloop measurement path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the FP64 reference the HIP path
 * is tested against: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product library (libwloop_amd.so) never links or calls it.
 *
 * Field layout used throughout the oracle: lexicographic sites with x3
 * fastest, s = x3 + N3*(x2 + N2*(x1 + N1*x0)); a link field is
 * double u[V][4][18] (mu = 0 is time), each 3x3 complex matrix stored row
 * major as (re,im) pairs c11..c33 -- the memory image of openQCD's su3_dble.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each function is listed in the header of wl_oracle.c.
 */
#ifndef WL_ORACLE_H
#define WL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dims[4] = {N0,N1,N2,N3}; bc = 0 (open in time) or 3 (periodic). */

/* ---- synthetic configuration generator (no reference counterpart) ---- */
void wo_generate(const int *dims, int bc, uint64_t seed, double eps, double *u);
void wo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void wo_random_su3(uint64_t seed, uint32_t site, uint32_t stream, double *g);

/* ---- single-matrix and staple primitives ---- */
void wo_project(double *m, int iter);                 /* projection.c:120-156 */
void wo_staple_sum(const int *dims, const double *a, const double *b,
                   int mu, int nu, double *v);         /* staple_sum.c:53-87 */

/* ---- smearing ---- */
/* hyp_links.c:398-472: time,space select which directions are replaced. */
void wo_hyp(const int *dims, int bc, double *u, double a1, double a2, double a3,
            int proj, int piter, int time, int space);
/* hyp_spatial_links.c:289-344: in place on the spatial links of u. */
void wo_hyp_spatial(const int *dims, double *u, double a1, double a2,
                    int proj, int piter);
/* ape_links.c:217-284: one step; time=0 spatial only, time=1 all four. */
void wo_ape(const int *dims, double *u, double alpha, int proj, int piter, int time);

/* ---- operator levels + transporters + contraction ---- */
typedef struct {
   int msl, nss[5], mwlt, mwlr;
   int op_smear;            /* 0 = APE, 1 = HYP (wloop.h:18-22) */
   double alpha_ape, alpha_hyp[2];
   int proj, proj_iter;
} wo_parms;

/* wloop.c:524-608 + smear.c:99-121 on an already HYP-smeared field u.
 * mode 0: the reference loop structure (products recomputed per level pair,
 *         time transporters recomputed per R);
 * mode 1: products shared between level pairs (same products, same sums:
 *         results are bit-identical to mode 0).
 * nthreads: OpenMP threads (<=0: library default).  wl has
 * msl*msl*mwlt*mwlr*N0 entries indexed by wl_idx (wloop_parms.c:431-434). */
void wo_wloop(const int *dims, const double *u, const wo_parms *p, int mode,
              int nthreads, double *wl);

/* ms10.c:633-646: hyp_links(alpha_action) on u in place, then wo_wloop. */
void wo_ms10(const int *dims, int bc, double *u, const double *alpha_action,
             const wo_parms *p, int mode, int nthreads, double *wl);

/* Spatial operator level sl (3V links, [V][3][18]) as built by smear(). */
void wo_levels(const int *dims, const double *u, const wo_parms *p, double *lev);

/* Wall seconds of one instance of each stage, reference loop structure:
 * t[0] HYP, t[1] one APE (or 3-d HYP) step, t[2] one time-transporter update,
 * t[3] one spatial-transporter update of all levels, t[4] one (R,T)
 * contraction over all level pairs (mean of (1,1) and (mwlr,mwlt)). */
void wo_time_stages(const int *dims, int bc, const double *u, const double *alpha_action,
                    const wo_parms *p, int nthreads, double *t);

/* ---- test helpers (openQCD random_g/transform_ud/shift_ud analogues) ---- */
void wo_gauge_transform(const int *dims, double *u, uint64_t seed);
void wo_shift(const int *dims, double *u, const int *s);
double wo_plaquette(const int *dims, const double *u);
int wo_wl_idx(const wo_parms *p, int sl1, int sl2, int t, int r, int x0);
void wo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
