/*
 * smear.h -- smearing entry points of the wloop module interface
 * (reference include/smear.h:49-72), served by libwloop_amd.so on the GPU.
 *
 *   hyp_links        hyp_links.c:468-472   HYP of all four directions of udfld()
 *   hyp_time_links   hyp_links.c:454-458   ... time links only
 *   hyp_space_links  hyp_links.c:461-465   ... spatial links only
 *   free_hyp         hyp_links.c:376-395
 *   hyp_spatial_links hyp_spatial_links.c:289-344  3-d HYP of a link field
 *   free_hyp_spatial hyp_spatial_links.c:269-286
 *   ape_links        ape_links.c:275-278   APE step of all four directions
 *   ape_spatial_links ape_links.c:281-284  APE step of the spatial links
 *   staple_sum       staple_sum.c:53-87    v(x,mu) += staples from u1 (nu) and u2 (mu)
 *   approx_project_to_su3_dble projection.c:120-156 (single host matrix)
 *
 * Fields passed by pointer are either udfld() or fields from alloc_lnk().
 * Errors (invalid arguments, device failures) abort through error().
 */
#ifndef SMEAR_H
#define SMEAR_H

#include "su3.h"

/* z = a*u + b*v and u = v + w, element-wise on su3_dble (smear.h:8-47) */
#define _su3_lcomb(z, a, u, b, v)                                   \
   do {                                                             \
      const double *pu_ = (const double *)&(u);                     \
      const double *pv_ = (const double *)&(v);                     \
      double *pz_ = (double *)&(z);                                 \
      int e_;                                                       \
      for (e_ = 0; e_ < 18; e_++) pz_[e_] = (a) * pu_[e_] + (b) * pv_[e_]; \
   } while (0)

#define _su3_add(u, v, w)                                           \
   do {                                                             \
      const double *pv_ = (const double *)&(v);                     \
      const double *pw_ = (const double *)&(w);                     \
      double *pu_ = (double *)&(u);                                 \
      int e_;                                                       \
      for (e_ = 0; e_ < 18; e_++) pu_[e_] = pv_[e_] + pw_[e_];      \
   } while (0)

#ifdef __cplusplus
extern "C" {
#endif

extern void free_hyp(void);
extern void hyp_time_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep);
extern void hyp_space_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep);
extern void hyp_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep);

extern void free_hyp_spatial(void);
extern void hyp_spatial_links(su3_dble *ub, double alpha1, double alpha2, int proj_su3, int piter, int keep);

extern void ape_links(su3_dble *ub, double a, int proj_su3, int piter, int keep);
extern void ape_spatial_links(su3_dble *ub, double a, int proj_su3, int piter, int keep);

extern void staple_sum(su3_dble *u1, su3_dble *u2, int mu, int nu, su3_dble *vb, su3_dble *cbuf);

extern void approx_project_to_su3_dble(su3_dble *u, int iter);

#ifdef __cplusplus
}
#endif
#endif
