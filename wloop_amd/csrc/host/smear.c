/*
 * smear.c -- smearing entry points of smear.h, run on the GPU.
 *
 * udfld() is smeared in place on the device copy of the gauge field (the
 * host copy is refreshed lazily by udfld()); a field from alloc_lnk() is
 * uploaded to a second device context, smeared there and copied back.
 * Reference: modules/smear/hyp_links.c, ape_links.c, hyp_spatial_links.c,
 * staple_sum.c, projection.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "smear.h"
#include "internal.h"

static void hyp_dirs(double a1, double a2, double a3, int proj, int piter, int time, int space)
{
   if (a1 == 0.0) return;                               /* hyp_links.c:405-406 */
   double a[3] = {a1, a2, a3};
   wlg_ctx *c = wl_gauge_ctx();
   wl_check(wlg_hyp(c, a, proj != 0, piter, time, space, wl_precision()), "hyp_links [smear.c]");
   wl_device_changed();
}

void hyp_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep)
{
   (void)keep;
   hyp_dirs(alpha1, alpha2, alpha3, proj_su3, piter, 1, 1);
}

void hyp_time_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep)
{
   (void)keep;
   hyp_dirs(alpha1, alpha2, alpha3, proj_su3, piter, 1, 0);
}

void hyp_space_links(double alpha1, double alpha2, double alpha3, int proj_su3, int piter, int keep)
{
   (void)keep;
   hyp_dirs(alpha1, alpha2, alpha3, proj_su3, piter, 0, 1);
}

/* HYP scratch lives in the device context's arena and is reused by the next
 * call; free_wloop() releases it (hyp_links.c:376-395 frees host buffers) */
void free_hyp(void) {}

void free_hyp_spatial(void) {}

/* one smearing step on ub: the device gauge field if ub is udfld(), else via
 * the auxiliary context; which = 0 APE 4-d, 1 APE spatial, 2 HYP 3-d */
static void on_field(su3_dble *ub, int which, const double *arg, const char *where)
{
   int prec = wl_precision();
   wlg_ctx *c;
   int own = (ub == wl_host_field());
   double *aos = NULL;
   if (own) {
      c = wl_gauge_ctx();
   } else {
      c = wl_aux_ctx();
      aos = malloc(sizeof(double) * 72 * (size_t)VOLUME);
      error(aos == NULL, 1, where, "Unable to allocate transfer buffer");
      wl_pack_lex(ub, aos);
      wl_check(wlg_upload(c, aos), where);
   }
   int proj = (int)arg[2], piter = (int)arg[3];
   if (which == 0)
      wl_check(wlg_ape_step(c, arg[0], proj, piter, 1, prec), where);
   else if (which == 1)
      wl_check(wlg_ape_step(c, arg[0], proj, piter, 0, prec), where);
   else
      wl_check(wlg_hyp_spatial_step(c, arg[0], arg[1], proj, piter, prec), where);
   if (own) {
      wl_device_changed();
      set_bc();                                         /* ape_links.c:264-268 */
   } else {
      wl_check(wlg_download(c, aos), where);
      wl_unpack_lex(aos, ub);
      free(aos);
   }
}

void ape_links(su3_dble *ub, double a, int proj_su3, int piter, int keep)
{
   (void)keep;
   double arg[4] = {a, 0.0, proj_su3 != 0, piter};
   on_field(ub, 0, arg, "ape_links [smear.c]");
}

void ape_spatial_links(su3_dble *ub, double a, int proj_su3, int piter, int keep)
{
   (void)keep;
   double arg[4] = {a, 0.0, proj_su3 != 0, piter};
   on_field(ub, 1, arg, "ape_spatial_links [smear.c]");
}

void hyp_spatial_links(su3_dble *ub, double alpha1, double alpha2, int proj_su3, int piter, int keep)
{
   (void)keep;
   double arg[4] = {alpha1, alpha2, proj_su3 != 0, piter};
   on_field(ub, 2, arg, "hyp_spatial_links [smear.c]");
}

/* vb(x,mu) += staples of (u1: nu-links, u2: mu-links); cbuf unused (no
 * boundary with one process) */
void staple_sum(su3_dble *u1, su3_dble *u2, int mu, int nu, su3_dble *vb, su3_dble *cbuf)
{
   (void)cbuf;
   size_t n = 72 * (size_t)VOLUME;
   double *a = malloc(sizeof(double) * 3 * n);
   error(a == NULL, 1, "staple_sum [smear.c]", "Unable to allocate transfer buffers");
   wl_pack_lex(u1, a);
   wl_pack_lex(u2, a + n);
   wl_pack_lex(vb, a + 2 * n);
   wl_check(wlg_staple_sum(wl_aux_ctx(), a, a + n, mu, nu, a + 2 * n), "staple_sum [smear.c]");
   wl_unpack_lex(a + 2 * n, vb);
   free(a);
}

/* approximate projection of one host matrix (projection.c:120-156,
 * hep-lat/0506008): scale to sum |u|^2 = 3, then iter times
 * X = u (3 - u^dag u)/2, u = (1 - i Im det X / 3) X */
void approx_project_to_su3_dble(su3_dble *u, int iter)
{
   double *p = (double *)u;
   double d = 0.0;
   for (int e = 0; e < 18; e++) d += p[e] * p[e];
   if (d < sqrt(18.0) * DBL_EPSILON) return;
   d = 1.0 / sqrt(d / 3.0);
   for (int e = 0; e < 18; e++) p[e] *= d;
   for (int it = 0; it < iter; it++) {
      const complex_dble *c = &u->c11;
      complex_dble h[9], x[9];
      for (int a = 0; a < 3; a++)
         for (int b = 0; b < 3; b++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
               re += c[3 * k + a].re * c[3 * k + b].re + c[3 * k + a].im * c[3 * k + b].im;
               im += c[3 * k + a].re * c[3 * k + b].im - c[3 * k + a].im * c[3 * k + b].re;
            }
            h[3 * a + b].re = (a == b ? 1.5 : 0.0) - 0.5 * re;
            h[3 * a + b].im = -0.5 * im;
         }
      for (int a = 0; a < 3; a++)
         for (int b = 0; b < 3; b++) {
            double re = 0.0, im = 0.0;
            for (int k = 0; k < 3; k++) {
               re += c[3 * a + k].re * h[3 * k + b].re - c[3 * a + k].im * h[3 * k + b].im;
               im += c[3 * a + k].re * h[3 * k + b].im + c[3 * a + k].im * h[3 * k + b].re;
            }
            x[3 * a + b].re = re;
            x[3 * a + b].im = im;
         }
      /* Im det x by expansion along the first row */
      double m1r = (x[4].re * x[8].re - x[4].im * x[8].im) - (x[5].re * x[7].re - x[5].im * x[7].im);
      double m1i = (x[4].re * x[8].im + x[4].im * x[8].re) - (x[5].re * x[7].im + x[5].im * x[7].re);
      double m2r = (x[3].re * x[8].re - x[3].im * x[8].im) - (x[5].re * x[6].re - x[5].im * x[6].im);
      double m2i = (x[3].re * x[8].im + x[3].im * x[8].re) - (x[5].re * x[6].im + x[5].im * x[6].re);
      double m3r = (x[3].re * x[7].re - x[3].im * x[7].im) - (x[4].re * x[6].re - x[4].im * x[6].im);
      double m3i = (x[3].re * x[7].im + x[3].im * x[7].re) - (x[4].re * x[6].im + x[4].im * x[6].re);
      double di = (x[0].re * m1i + x[0].im * m1r) - (x[1].re * m2i + x[1].im * m2r)
                + (x[2].re * m3i + x[2].im * m3r);
      double zi = -di / 3.0;
      complex_dble *o = &u->c11;
      for (int e = 0; e < 9; e++) {
         o[e].re = x[e].re - zi * x[e].im;
         o[e].im = x[e].im + zi * x[e].re;
      }
   }
}
