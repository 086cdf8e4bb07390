/*
 * lnk.c -- link-field container (lnk.h): addressing of modules/lnk/lnkcom.c
 * with one process owning the whole lattice, host staple utilities of
 * modules/lnk/lnkstaple.c, and the plaquette sum of modules/lnk/lnkplaq.c
 * (computed on the GPU).
 *
 * Link (ix,mu) of a field sits at the odd end point: for odd ix at
 * [8*(ix-VOLUME/2)+2*mu], for even ix at the odd neighbour iup[ix][mu],
 * slot 2*mu+1 (lnkcom.c:840-859).  There are no boundary links (BNDRY = 0):
 * the halo routines have nothing to exchange.
 */
#include <stdlib.h>
#include <string.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "internal.h"

static int idx_empty[1] = {0};

static int ofs_lnk(int ix, int mu)
{
   if (ix < VOLUME / 2) return 8 * (iup[ix][mu] - VOLUME / 2) + 2 * mu + 1;
   return 8 * (ix - VOLUME / 2) + 2 * mu;
}

su3_dble *lnk(su3_dble *ub, int ix, int mu) { return ub + ofs_lnk(ix, mu); }

su3_dble *lnkf(su3_dble *ub, int ix, int mu, int face)
{
   (void)face;
   error((ix < 0) || (ix >= VOLUME), 1, "lnkf [lnk.c]",
         "Point %d is outside the lattice (no boundary links with one process)", ix);
   return ub + ofs_lnk(ix, mu);
}

su3_dble *lnk_buf(su3_dble *ub, int ix, int mu, int face)
{
   (void)ub; (void)ix; (void)mu; (void)face;
   error(1, 1, "lnk_buf [lnk.c]", "There are no boundary buffers with one process");
   return NULL;
}

su3_dble *alloc_lnk(void)
{
   geometry();
   su3_dble *ub = calloc((size_t)LOCAL + BNDRY_K, sizeof(su3_dble));
   error(ub == NULL, 1, "alloc_lnk [lnk.c]", "Unable to allocate memory for link field");
   return ub;
}

su3_dble *alloc_buf_uk(void) { return NULL; }       /* lnkcom.c:785-803 with NPROC = 1 */

void free_lnk(su3_dble *ub, int cb)
{
   (void)cb;
   free(ub);
}

void free_buf_uk(su3_dble *ub) { free(ub); }

void copy_bnd(su3_dble *ub) { (void)ub; }
void copy_u0(su3_dble *ub) { (void)ub; }
void copyback_u0(su3_dble *ub) { (void)ub; }
void copy_u0n(su3_dble *ub) { (void)ub; }
void copyback_u0n(su3_dble *ub) { (void)ub; }
void copy_uk(su3_dble *ub) { (void)ub; }
void copyback_uk(su3_dble *ub) { (void)ub; }
void send_uk(su3_dble *ub, su3_dble *sbuf) { (void)ub; (void)sbuf; }

int *get_lnk_idx_u0(void) { return idx_empty; }
int *get_lnk_idx_uk(void) { return idx_empty; }

void lnk_plaq(su3_dble *ub, int ix, int mu, int nu, su3_dble **u)
{
   u[0] = lnk(ub, ix, mu);
   u[1] = lnkf(ub, iup[ix][mu], nu, mu);
   u[2] = lnk(ub, ix, nu);
   u[3] = lnkf(ub, iup[ix][nu], mu, nu);
}

void assign_lnk2lnk(su3_dble *ub1, su3_dble *ub2)
{
   memcpy(ub2, ub1, sizeof(su3_dble) * (4 * (size_t)VOLUME + BNDRY / 4));
}

void assign_ud2lnk(su3_dble *ub)
{
   memcpy(ub, udfld(), sizeof(su3_dble) * (4 * (size_t)VOLUME + BNDRY / 4));
}

void assign_lnk2ud(su3_dble *ub)
{
   memcpy(udfld(), ub, sizeof(su3_dble) * (4 * (size_t)VOLUME + BNDRY / 4));
   set_flags(UPDATED_UD);
}

/* ---- single-staple host utilities (lnkstaple.c:184-410) ----------------- */

static void mul(const su3_dble *a, const su3_dble *b, su3_dble *w, int da, int db)
{
   /* w = op(a) op(b), op = identity or dagger */
   const complex_dble *pa = &a->c11, *pb = &b->c11;
   su3_dble t;
   complex_dble *pt = &t.c11;
   for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) {
         double re = 0.0, im = 0.0;
         for (int j = 0; j < 3; j++) {
            complex_dble x = da ? pa[3 * j + i] : pa[3 * i + j];
            complex_dble y = db ? pb[3 * k + j] : pb[3 * j + k];
            if (da) x.im = -x.im;
            if (db) y.im = -y.im;
            re += x.re * y.re - x.im * y.im;
            im += x.re * y.im + x.im * y.re;
         }
         pt[3 * i + k].re = re;
         pt[3 * i + k].im = im;
      }
   *w = t;
}

static void add_to(su3_dble *v, const su3_dble *t)
{
   double *pv = (double *)v;
   const double *pt = (const double *)t;
   for (int e = 0; e < 18; e++) pv[e] += pt[e];
}

void add_up_staple_lnks_save(su3_dble *u1, su3_dble *u2, su3_dble *u3, su3_dble *v)
{
   su3_dble t;
   mul(u2, u3, &t, 0, 1);     /* u2 u3^dag */
   mul(u1, &t, &t, 0, 0);
   add_to(v, &t);
}

void add_dn_staple_lnks_save(su3_dble *u1, su3_dble *u2, su3_dble *u3, su3_dble *v)
{
   su3_dble t;
   mul(u2, u3, &t, 0, 0);     /* u2 u3 */
   mul(u1, &t, &t, 1, 0);     /* u1^dag (u2 u3) */
   add_to(v, &t);
}

void add_up_staple_save(su3_dble *u1o, su3_dble *u2o, int ix, int mu, int nu, su3_dble *v)
{
   add_up_staple_lnks_save(lnk(u1o, ix, nu), lnkf(u2o, iup[ix][nu], mu, nu),
                           lnkf(u1o, iup[ix][mu], nu, mu), v);
}

void add_dn_staple_save(su3_dble *u1o, su3_dble *u2o, int ix, int mu, int nu, su3_dble *v)
{
   int iy = idn[ix][nu];
   add_dn_staple_lnks_save(lnk(u1o, iy, nu), lnk(u2o, iy, mu), lnkf(u1o, iup[iy][mu], nu, mu), v);
}

/* the SU(3)-assuming variants give the same matrix for SU(3) links */
void add_up_staple(su3_dble *u1o, su3_dble *u2o, int ix, int mu, int nu, su3_dble *v)
{
   add_up_staple_save(u1o, u2o, ix, mu, nu, v);
}

void add_dn_staple(su3_dble *u1o, su3_dble *u2o, int ix, int mu, int nu, su3_dble *v)
{
   add_dn_staple_save(u1o, u2o, ix, mu, nu, v);
}

void add_up_staple_lnks(su3_dble *u1, su3_dble *u2, su3_dble *u3, su3_dble *v)
{
   add_up_staple_lnks_save(u1, u2, u3, v);
}

void add_dn_staple_lnks(su3_dble *u1, su3_dble *u2, su3_dble *u3, su3_dble *v)
{
   add_dn_staple_lnks_save(u1, u2, u3, v);
}

/* ---- plaquette sum (lnkplaq.c:117-128), on the GPU ------------------------ */

double lnk_plaq_sum(su3_dble *ub)
{
   double s = 0.0;
   if (ub == wl_host_field()) {
      wl_check(wlg_plaquette(wl_gauge_ctx(), &s), "lnk_plaq_sum [lnk.c]");
   } else {
      wlg_ctx *a = wl_aux_ctx();
      double *aos = malloc(sizeof(double) * 72 * (size_t)VOLUME);
      error(aos == NULL, 1, "lnk_plaq_sum [lnk.c]", "Unable to allocate transfer buffer");
      wl_pack_lex(ub, aos);
      wl_check(wlg_upload(a, aos), "lnk_plaq_sum [lnk.c]");
      free(aos);
      wl_check(wlg_plaquette(a, &s), "lnk_plaq_sum [lnk.c]");
   }
   return s;
}
