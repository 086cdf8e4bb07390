/*
 * wl_oracle.c -- CPU FP64 restatement of the wloop measurement path.
 *
 * TEST INFRASTRUCTURE ONLY (the reference of the tests, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * The reference (bjoern-leder/wloop, wloop-1.4) cannot be built here: it
 * needs openQCD-1.4, which is not vendored.  This file restates its algorithm
 * from the sources, one function per reference routine, in plain C99.
 *
 * Parity pins (what checks each function; see tests/test_oracle.py):
 *   wo_project        fixed points (SU(3), c*SU(3), 0), polar factor of
 *                     V*H and det-phase removal of V*exp(i phi) at p=30,
 *                     unitarity/det after p=30 (devel/smear/check2.c)
 *   wo_staple_sum     gauge covariance, unit field, brute-force numpy staple
 *   wo_hyp            textbook Hasenfratz-Knechtli HYP written independently
 *                     in numpy (hep-lat/0103029), HYP(a,0,0)=(1-a)APE4d
 *                     (devel/smear/check3.c:95-121), gauge covariance
 *   wo_hyp_spatial    HYP3d(a,0)=(1-a)APE3d (devel/smear/check3.c:136-182);
 *                     alpha2!=0: closed form on an abelian constant-flux
 *                     field with a distinct flux per plane, which holds for
 *                     the wiring of hyp_spatial_links.c:187-234 (level-1
 *                     links decorated in the staple's own plane) and fails
 *                     for the textbook one
 *   wo_ape            unit field -> (1+4a)1 unprojected, gauge covariance,
 *                     brute-force numpy APE step
 *   wo_wloop          unit field -> W=1 (devel/wloop/check1.c), gauge
 *                     invariance (check2.c), translation invariance with
 *                     x0 shift (check3.c), brute-force path walk on small
 *                     lattices, abelian closed form cos(Phi R T), open-bc
 *                     exact zeros, mode0 == mode1 bitwise
 *   wo_generate       unitarity/det of every link, Haar moments of g,
 *                     exact zeros at x0=N0-1 (open bc); no reference
 *                     counterpart (openQCD random_ud is absent) -- its
 *                     output is pinned only by these properties
 */
#include "wl_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SRT18 4.242640687119285

typedef struct { double re, im; } cplx;
typedef struct { cplx c[3][3]; } su3;   /* same memory image as su3_dble */

/* ------------------------------------------------------------------ */
/* 3x3 complex algebra (textbook; summation order as openQCD's macros) */
/* ------------------------------------------------------------------ */

static void su3_mul(const su3 *a, const su3 *b, su3 *w)        /* w = a*b */
{
   su3 t;
   for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) {
         double re = a->c[i][0].re * b->c[0][k].re - a->c[i][0].im * b->c[0][k].im
                   + a->c[i][1].re * b->c[1][k].re - a->c[i][1].im * b->c[1][k].im
                   + a->c[i][2].re * b->c[2][k].re - a->c[i][2].im * b->c[2][k].im;
         double im = a->c[i][0].re * b->c[0][k].im + a->c[i][0].im * b->c[0][k].re
                   + a->c[i][1].re * b->c[1][k].im + a->c[i][1].im * b->c[1][k].re
                   + a->c[i][2].re * b->c[2][k].im + a->c[i][2].im * b->c[2][k].re;
         t.c[i][k].re = re;
         t.c[i][k].im = im;
      }
   *w = t;
}

static void su3_mul_dag(const su3 *a, const su3 *b, su3 *w)    /* w = a*b^dag */
{
   su3 t;
   for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) {
         double re = a->c[i][0].re * b->c[k][0].re + a->c[i][0].im * b->c[k][0].im
                   + a->c[i][1].re * b->c[k][1].re + a->c[i][1].im * b->c[k][1].im
                   + a->c[i][2].re * b->c[k][2].re + a->c[i][2].im * b->c[k][2].im;
         double im = a->c[i][0].im * b->c[k][0].re - a->c[i][0].re * b->c[k][0].im
                   + a->c[i][1].im * b->c[k][1].re - a->c[i][1].re * b->c[k][1].im
                   + a->c[i][2].im * b->c[k][2].re - a->c[i][2].re * b->c[k][2].im;
         t.c[i][k].re = re;
         t.c[i][k].im = im;
      }
   *w = t;
}

static void su3_dag_mul(const su3 *a, const su3 *b, su3 *w)    /* w = a^dag*b */
{
   su3 t;
   for (int i = 0; i < 3; i++)
      for (int k = 0; k < 3; k++) {
         double re = a->c[0][i].re * b->c[0][k].re + a->c[0][i].im * b->c[0][k].im
                   + a->c[1][i].re * b->c[1][k].re + a->c[1][i].im * b->c[1][k].im
                   + a->c[2][i].re * b->c[2][k].re + a->c[2][i].im * b->c[2][k].im;
         double im = a->c[0][i].re * b->c[0][k].im - a->c[0][i].im * b->c[0][k].re
                   + a->c[1][i].re * b->c[1][k].im - a->c[1][i].im * b->c[1][k].re
                   + a->c[2][i].re * b->c[2][k].im - a->c[2][i].im * b->c[2][k].re;
         t.c[i][k].re = re;
         t.c[i][k].im = im;
      }
   *w = t;
}

static void su3_acc(su3 *v, const su3 *a)                      /* v += a */
{
   double *pv = (double *)v;
   const double *pa = (const double *)a;
   for (int e = 0; e < 18; e++) pv[e] = pv[e] + pa[e];
}

/* Re tr(a^dag b), summed row by row as three vector products
 * (wloop.c:478-481, _vector_prod_re). */
static double su3_retr_dag_mul(const su3 *a, const su3 *b)
{
   double r[3];
   for (int i = 0; i < 3; i++)
      r[i] = a->c[i][0].re * b->c[i][0].re + a->c[i][0].im * b->c[i][0].im
           + a->c[i][1].re * b->c[i][1].re + a->c[i][1].im * b->c[i][1].im
           + a->c[i][2].re * b->c[i][2].re + a->c[i][2].im * b->c[i][2].im;
   return r[0] + r[1] + r[2];
}

/* ------------------------------------------------------------------ */
/* approximate SU(3) projection, hep-lat/0506008 (projection.c:87-156) */
/* ------------------------------------------------------------------ */

static double det_im(const su3 *u)                        /* projection.c:87-117 */
{
   const cplx (*c)[3] = u->c;
   double d1r = (c[1][1].re * c[2][2].re - c[1][1].im * c[2][2].im)
              - (c[1][2].re * c[2][1].re - c[1][2].im * c[2][1].im);
   double d1i = (c[1][1].re * c[2][2].im + c[1][1].im * c[2][2].re)
              - (c[1][2].re * c[2][1].im + c[1][2].im * c[2][1].re);
   double d2r = (c[1][0].re * c[2][2].re - c[1][0].im * c[2][2].im)
              - (c[1][2].re * c[2][0].re - c[1][2].im * c[2][0].im);
   double d2i = (c[1][0].re * c[2][2].im + c[1][0].im * c[2][2].re)
              - (c[1][2].re * c[2][0].im + c[1][2].im * c[2][0].re);
   double d3r = (c[1][0].re * c[2][1].re - c[1][0].im * c[2][1].im)
              - (c[1][1].re * c[2][0].re - c[1][1].im * c[2][0].im);
   double d3i = (c[1][0].re * c[2][1].im + c[1][0].im * c[2][1].re)
              - (c[1][1].re * c[2][0].im + c[1][1].im * c[2][0].re);
   return (c[0][0].re * d1i + c[0][0].im * d1r)
        - (c[0][1].re * d2i + c[0][1].im * d2r)
        + (c[0][2].re * d3i + c[0][2].im * d3r);
}

static void project(su3 *u, int iter)                    /* projection.c:120-156 */
{
   double *p = (double *)u;
   double d = 0.0;
   for (int e = 0; e < 18; e++) d += p[e] * p[e];
   if (d < SRT18 * DBL_EPSILON)
      return;
   d = 1.0 / sqrt(d / 3.0);
   for (int e = 0; e < 18; e++) p[e] = d * p[e];
   for (int it = 0; it < iter; it++) {
      su3 x;
      su3_dag_mul(u, u, &x);
      double *px = (double *)&x;
      for (int e = 0; e < 18; e++) px[e] = -0.5 * px[e];
      x.c[0][0].re += 1.5;
      x.c[1][1].re += 1.5;
      x.c[2][2].re += 1.5;
      su3_mul(u, &x, &x);
      double zi = det_im(&x) / -3.0;
      for (int a = 0; a < 3; a++)
         for (int b = 0; b < 3; b++) {
            cplx v = x.c[a][b];
            u->c[a][b].re = v.re - zi * v.im;        /* (1 + i zi) * v */
            u->c[a][b].im = v.im + zi * v.re;
         }
   }
}

void wo_project(double *m, int iter) { project((su3 *)m, iter); }

/* ------------------------------------------------------------------ */
/* lattice geometry: lexicographic, x3 fastest, periodic index wrap     */
/* ------------------------------------------------------------------ */

typedef struct {
   int n[4];
   long vol;
   int *up, *dn;   /* [vol][4] */
} geom;

static void geom_init(geom *g, const int *dims)
{
   for (int m = 0; m < 4; m++) g->n[m] = dims[m];
   g->vol = (long)dims[0] * dims[1] * dims[2] * dims[3];
   g->up = malloc(sizeof(int) * 4 * g->vol);
   g->dn = malloc(sizeof(int) * 4 * g->vol);
   for (long s = 0; s < g->vol; s++) {
      int x[4];
      long r = s;
      x[3] = r % dims[3]; r /= dims[3];
      x[2] = r % dims[2]; r /= dims[2];
      x[1] = r % dims[1]; r /= dims[1];
      x[0] = (int)r;
      for (int m = 0; m < 4; m++) {
         int y[4], z[4];
         for (int q = 0; q < 4; q++) { y[q] = x[q]; z[q] = x[q]; }
         y[m] = (x[m] + 1) % dims[m];
         z[m] = (x[m] + dims[m] - 1) % dims[m];
         g->up[4 * s + m] = y[3] + dims[3] * (y[2] + dims[2] * (y[1] + dims[1] * y[0]));
         g->dn[4 * s + m] = z[3] + dims[3] * (z[2] + dims[2] * (z[1] + dims[1] * z[0]));
      }
   }
}

static void geom_free(geom *g) { free(g->up); free(g->dn); }

static inline su3 *L4(double *u, long s, int mu) { return (su3 *)(u + 18 * (4 * s + mu)); }
static inline const su3 *C4(const double *u, long s, int mu) { return (const su3 *)(u + 18 * (4 * s + mu)); }

/* ------------------------------------------------------------------ */
/* staples (lnkstaple.c:318-382, staple_sum.c:53-87)                    */
/* ------------------------------------------------------------------ */

/* v(x,mu) += A_nu(x) B_mu(x+nu) A_nu(x+mu)^dag
 *          + A_nu(x-nu)^dag B_mu(x-nu) A_nu(x-nu+mu)           */
static void staple_sum(const geom *g, const double *a, const double *b, int mu, int nu,
                       double *v)
{
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g->vol; x++) {
      su3 t;
      long xpn = g->up[4 * x + nu], xpm = g->up[4 * x + mu];
      su3_mul_dag(C4(b, xpn, mu), C4(a, xpm, nu), &t);     /* lnkstaple.c:323 */
      su3_mul(C4(a, x, nu), &t, &t);
      su3_acc(L4(v, x, mu), &t);
      long xmn = g->dn[4 * x + nu];
      long xmnpm = g->up[4 * xmn + mu];
      su3_mul(C4(b, xmn, mu), C4(a, xmnpm, nu), &t);      /* lnkstaple.c:334 */
      su3_dag_mul(C4(a, xmn, nu), &t, &t);
      su3_acc(L4(v, x, mu), &t);
   }
}

void wo_staple_sum(const int *dims, const double *a, const double *b, int mu, int nu,
                   double *v)
{
   geom g;
   geom_init(&g, dims);
   staple_sum(&g, a, b, mu, nu, v);
   geom_free(&g);
}

/* v = alpha*u + beta*v, then project (hyp_links.c:154-191, _su3_lcomb) */
static void decorate(const geom *g, const double *u, double *v, int mu0, int mu1,
                     double alpha, double beta, int proj, int piter)
{
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g->vol; x++)
      for (int mu = mu0; mu < mu1; mu++) {
         const double *pu = (const double *)C4(u, x, mu);
         double *pv = (double *)L4(v, x, mu);
         for (int e = 0; e < 18; e++) pv[e] = alpha * pu[e] + beta * pv[e];
         if (proj) project((su3 *)pv, piter);
      }
}

static void set_bc(const geom *g, int bc, double *u)    /* openQCD set_bc, type 0 */
{
   if (bc != 0) return;
   long slice = (long)g->n[1] * g->n[2] * g->n[3];
   for (long x = (g->n[0] - 1) * slice; x < g->vol; x++)
      memset(L4(u, x, 0), 0, sizeof(su3));
}

/* ------------------------------------------------------------------ */
/* HYP smearing (hyp_links.c:106-472)                                   */
/* ------------------------------------------------------------------ */

/* pair partition index, hyp_links.c:124-139 */
static const int part4[4][4] = {{0, 0, 1, 2}, {0, 0, 2, 1}, {1, 2, 0, 0}, {2, 1, 0, 0}};

void wo_hyp(const int *dims, int bc, double *u, double a1, double a2, double a3,
            int proj, int piter, int time, int space)
{
   if (a1 == 0.0) return;                               /* hyp_links.c:405-406 */
   geom g;
   geom_init(&g, dims);
   size_t fsz = sizeof(double) * 18 * 4 * g.vol;
   double *d[6];
   for (int i = 0; i < 6; i++) d[i] = calloc(1, fsz);
   double al[3] = {1.0 - a1, 1.0 - a2, 1.0 - a3};
   double be[3] = {a1 / 6.0, a2 / 4.0, a3 / 2.0};

   /* level 1 (hyp_links.c:320-373): V1[part{mu,eta}]_mu from U staples */
   for (int mu = 0; mu < 4; mu++)
      for (int i = 1; i < 4; i++) {
         int eta = (mu + i) & 3;
         staple_sum(&g, u, u, mu, eta, d[part4[mu][eta]]);
      }
   for (int i = 0; i < 3; i++) decorate(&g, u, d[i], 0, 4, al[2], be[2], proj, piter);

   /* level 2 (hyp_links.c:244-317) */
   for (int mu = 0; mu < 4; mu++)
      for (int i = 1; i < 4; i++) {
         int nu = (mu + i) & 3;
         int n = part4[mu][nu];
         for (int rho = 0; rho < 4; rho++)
            if (rho != mu && rho != nu)
               staple_sum(&g, d[n], d[part4[nu][rho]], mu, rho, d[3 + n]);
      }
   for (int i = 0; i < 3; i++) decorate(&g, u, d[3 + i], 0, 4, al[1], be[1], proj, piter);

   /* level 3 (hyp_links.c:194-241); reuse d[0] as output */
   memset(d[0], 0, fsz);
   int mu0 = 1 - (time > 0), mu1 = 1 + 3 * (space > 0);
   for (int mu = mu0; mu < mu1; mu++)
      for (int i = 1; i < 4; i++) {
         int nu = (mu + i) & 3;
         int n = part4[nu][mu];
         staple_sum(&g, d[3 + n], d[3 + n], mu, nu, d[0]);
      }
   decorate(&g, u, d[0], mu0, mu1, al[0], be[0], proj, piter);

   /* replace the smeared directions (hyp_links.c:432-444), then set_bc */
   for (long x = 0; x < g.vol; x++)
      for (int mu = mu0; mu < mu1; mu++) *L4(u, x, mu) = *L4(d[0], x, mu);
   set_bc(&g, bc, u);

   for (int i = 0; i < 6; i++) free(d[i]);
   geom_free(&g);
}

/* 3-d HYP on the spatial links (hyp_spatial_links.c:140-344) */
static const int part3[4][4] = {{0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 1}, {0, 0, 1, 0}};

void wo_hyp_spatial(const int *dims, double *u, double a1, double a2, int proj, int piter)
{
   geom g;
   geom_init(&g, dims);
   size_t fsz = sizeof(double) * 18 * 4 * g.vol;
   double *d[3];
   for (int i = 0; i < 3; i++) d[i] = calloc(1, fsz);
   double al[2] = {1.0 - a1, 1.0 - a2};
   double be[2] = {a1 / 4.0, a2 / 2.0};

   /* level 1 (hyp_spatial_links.c:237-283) */
   for (int mu = 1; mu < 4; mu++)
      for (int i = 0; i < 2; i++) {
         int eta = ((mu + i) % 3) + 1;
         staple_sum(&g, u, u, mu, eta, d[part3[mu][eta]]);
      }
   for (int i = 0; i < 2; i++) decorate(&g, u, d[i], 1, 4, al[1], be[1], proj, piter);

   /* level 2 (hyp_spatial_links.c:187-234): nu-links from slot [nu][mu],
    * mu-links from slot [mu][nu], exactly as the reference wires them */
   for (int mu = 1; mu < 4; mu++)
      for (int i = 0; i < 2; i++) {
         int nu = ((mu + i) % 3) + 1;
         staple_sum(&g, d[part3[nu][mu]], d[part3[mu][nu]], mu, nu, d[2]);
      }
   decorate(&g, u, d[2], 1, 4, al[0], be[0], proj, piter);

   for (long x = 0; x < g.vol; x++)
      for (int mu = 1; mu < 4; mu++) *L4(u, x, mu) = *L4(d[2], x, mu);
   for (int i = 0; i < 3; i++) free(d[i]);
   geom_free(&g);
}

/* ------------------------------------------------------------------ */
/* APE smearing, one step (ape_links.c:147-284)                         */
/* ------------------------------------------------------------------ */

static void ape_step(const geom *g, double *u, double *work, double alpha, int proj,
                     int piter, int time)
{
   size_t fsz = sizeof(double) * 18 * 4 * g->vol;
   double *cp = work, *vo = work + 18 * 4 * g->vol;
   memcpy(cp, u, fsz);                                   /* ape_links.c:230 */
   memset(vo, 0, fsz);
   int m0 = 1 - (time > 0);
   for (int mu = m0; mu < 4; mu++)                       /* ape_links.c:162-169 */
      for (int eta = m0; eta < 4; eta++)
         if (eta != mu) staple_sum(g, cp, cp, mu, eta, vo);
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g->vol; x++)                     /* ape_links.c:235-262 */
      for (int mu = m0; mu < 4; mu++) {
         double *pu = (double *)L4(u, x, mu);
         const double *pv = (const double *)L4(vo, x, mu);
         for (int e = 0; e < 18; e++) pu[e] = pu[e] + alpha * pv[e];
         if (proj) project((su3 *)pu, piter);
      }
}

void wo_ape(const int *dims, double *u, double alpha, int proj, int piter, int time)
{
   geom g;
   geom_init(&g, dims);
   double *work = malloc(sizeof(double) * 2 * 18 * 4 * g.vol);
   ape_step(&g, u, work, alpha, proj, piter, time);
   free(work);
   geom_free(&g);
}

/* ------------------------------------------------------------------ */
/* operator levels (smear.c:99-121) and Wilson loops (wloop.c:153-608)  */
/* ------------------------------------------------------------------ */

int wo_wl_idx(const wo_parms *p, int sl1, int sl2, int t, int r, int x0)
{                                                      /* wloop_parms.c:431-434 */
   return sl1 + p->msl * (sl2 + p->msl * (t + p->mwlt * (r + p->mwlr * x0)));
}

/* Progress of long runs on stderr (stage ends), when WO_PROGRESS is set. */
static void progress(const char *what, int n)
{
   if (getenv("WO_PROGRESS")) {
      fprintf(stderr, "[oracle] %s %d done\n", what, n);
      fflush(stderr);
   }
}

static void build_levels(const geom *g, const double *u, const wo_parms *p, double *lev)
{
   size_t fsz = sizeof(double) * 18 * 4 * g->vol;
   double *f = malloc(fsz);
   double *work = malloc(2 * fsz);
   memcpy(f, u, fsz);                                    /* smear.c:107 */
   for (int sl = 0; sl < p->msl; sl++) {
      for (int i = 0; i < p->nss[sl]; i++) {             /* cumulative, smear.c:111-118 */
         if (p->op_smear == 1) {
            int dims[4] = {g->n[0], g->n[1], g->n[2], g->n[3]};
            wo_hyp_spatial(dims, f, p->alpha_hyp[0], p->alpha_hyp[1], p->proj, p->proj_iter);
         } else
            ape_step(g, f, work, p->alpha_ape, p->proj, p->proj_iter, 0);
      }
      double *l = lev + (size_t)sl * 18 * 3 * g->vol;  /* copy_lnk2us, smear.c:79-96 */
      for (long x = 0; x < g->vol; x++)
         for (int k = 0; k < 3; k++)
            memcpy(l + 18 * (3 * x + k), L4(f, x, k + 1), sizeof(su3));
      progress("operator level", sl);
   }
   free(work);
   free(f);
}

void wo_levels(const int *dims, const double *u, const wo_parms *p, double *lev)
{
   geom g;
   geom_init(&g, dims);
   build_levels(&g, u, p, lev);
   geom_free(&g);
}

static inline const su3 *C3(const double *l, long s, int k) { return (const su3 *)(l + 18 * (3 * s + k)); }

/* time transporter update L_T(x) = U_0(x) L_{T-1}(x+0) (wloop.c:153-204) */
static void lt_update(const geom *g, const double *u, const double *lt, double *out)
{
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g->vol; x++)
      su3_mul(C4(u, x, 0), (const su3 *)(lt + 18 * g->up[4 * x + 0]), (su3 *)(out + 18 * x));
}

/* spatial transporter update S_R(x,k) = F_k(x) S_{R-1}(x+k,k) (wloop.c:207-258) */
static void sr_update(const geom *g, const double *f, const double *prev, double *out)
{
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g->vol; x++)
      for (int k = 0; k < 3; k++)
         su3_mul(C3(f, x, k), C3(prev, g->up[4 * x + k + 1], k), (su3 *)(out + 18 * (3 * x + k)));
}

/* One (R,T) block of the contraction into wl (wloop.c:436-508, :558-600).
 * mode 0 repeats the reference loop structure, mode 1 shares the products. */
static void contract_rt(const geom *g, const wo_parms *p, int r, int t, const double *lt,
                        const double *S, int mode, double *acc, double *wl)
{
   const int msl = p->msl, N0 = g->n[0], R = r + 1, T = t + 1;
   const long slice = g->vol / N0;
   const double norm = 9.0 * g->n[1] * g->n[2] * g->n[3];    /* wloop.c:535 */
   const size_t lstride = (size_t)18 * 3 * g->vol;
   if (mode == 0) {
      double *res = acc;
      for (int sl1 = 0; sl1 < msl; sl1++)
         for (int sl2 = sl1; sl2 < msl; sl2++) {
            const double *s1 = S + sl1 * lstride, *s2 = S + sl2 * lstride;
#pragma omp parallel for schedule(static)
            for (int x0 = 0; x0 < N0; x0++) {
               double r0 = 0.0, r1 = 0.0;
               long xo = (long)((x0 + T) % N0) * slice;
               for (long xs = 0; xs < slice; xs++) {
                  long x = (long)x0 * slice + xs;
                  const su3 *ub = (const su3 *)(lt + 18 * x);
                  for (int k = 0; k < 3; k++) {
                     long xr = x;
                     for (int q = 0; q < R; q++) xr = g->up[4 * xr + k + 1];
                     const su3 *ubr = (const su3 *)(lt + 18 * xr);
                     su3 m0, m1;
                     su3_mul(ub, C3(s2, xo + xs, k), &m0);
                     su3_mul(C3(s1, x, k), ubr, &m1);
                     r0 += su3_retr_dag_mul(&m0, &m1);
                     if (sl1 != sl2) {
                        su3_mul(ub, C3(s1, xo + xs, k), &m0);
                        su3_mul(C3(s2, x, k), ubr, &m1);
                        r1 += su3_retr_dag_mul(&m0, &m1);
                     }
                  }
               }
               res[x0] = r0;
               res[x0 + N0] = r1;
            }
            for (int x0 = 0; x0 < N0; x0++) {                  /* wloop.c:584-597 */
               wl[wo_wl_idx(p, sl1, sl2, t, r, x0)] = res[x0] / norm;
               if (sl1 != sl2) wl[wo_wl_idx(p, sl2, sl1, t, r, x0)] = res[x0 + N0] / norm;
            }
         }
   } else {
#pragma omp parallel for schedule(static)
      for (int x0 = 0; x0 < N0; x0++) {
         double *a = acc + (size_t)x0 * msl * msl;
         for (int q = 0; q < msl * msl; q++) a[q] = 0.0;
         long xo = (long)((x0 + T) % N0) * slice;
         su3 A[5], B[5];
         for (long xs = 0; xs < slice; xs++) {
            long x = (long)x0 * slice + xs;
            const su3 *ub = (const su3 *)(lt + 18 * x);
            for (int k = 0; k < 3; k++) {
               long xr = x;
               for (int q = 0; q < R; q++) xr = g->up[4 * xr + k + 1];
               const su3 *ubr = (const su3 *)(lt + 18 * xr);
               for (int j = 0; j < msl; j++) su3_mul(ub, C3(S + j * lstride, xo + xs, k), &A[j]);
               for (int i = 0; i < msl; i++) su3_mul(C3(S + i * lstride, x, k), ubr, &B[i]);
               for (int i = 0; i < msl; i++)
                  for (int j = 0; j < msl; j++) a[i * msl + j] += su3_retr_dag_mul(&A[j], &B[i]);
            }
         }
      }
      for (int x0 = 0; x0 < N0; x0++)
         for (int i = 0; i < msl; i++)
            for (int j = 0; j < msl; j++)
               wl[wo_wl_idx(p, i, j, t, r, x0)] = acc[(size_t)x0 * msl * msl + i * msl + j] / norm;
   }
}

void wo_wloop(const int *dims, const double *u, const wo_parms *p, int mode, int nthreads,
              double *wl)
{
#ifdef _OPENMP
   int saved = omp_get_max_threads();
   if (nthreads > 0) omp_set_num_threads(nthreads);
#else
   (void)nthreads;
#endif
   geom g;
   geom_init(&g, dims);
   const int msl = p->msl;
   const int N0 = g.n[0];
   size_t lsz = sizeof(double) * 18 * 3 * g.vol;

   double *lev = malloc(lsz * msl);                     /* F^(sl), the pus[sl] table */
   build_levels(&g, u, p, lev);
   double *slong = malloc(lsz * msl);                   /* pus[msl+sl] */
   double *stmp = malloc(lsz);
   double *lt = malloc(sizeof(su3) * g.vol);            /* ud_long time links */
   double *ltmp = malloc(sizeof(su3) * g.vol);
   double *acc = malloc(sizeof(double) * (N0 * msl * msl + 2 * N0));

   for (int r = 0; r < p->mwlr; r++) {
      const int R = r + 1;
      for (long x = 0; x < g.vol; x++)                  /* assign_u02lnk, wloop.c:550 */
         memcpy(lt + 18 * x, C4(u, x, 0), sizeof(su3));
      for (int t = 0; t < p->mwlt; t++) {
         const int T = t + 1;
         if (T > 1) {                                     /* recomputed per R, wloop.c:554-556 */
            lt_update(&g, u, lt, ltmp);
            double *sw = lt; lt = ltmp; ltmp = sw;
         }
         if (R > 1 && T == 1) {                           /* wloop.c:562-563 */
            for (int sl = 0; sl < msl; sl++) {
               const double *f = lev + (size_t)sl * 18 * 3 * g.vol;
               const double *prev = (R == 2) ? f : slong + (size_t)sl * 18 * 3 * g.vol;
               sr_update(&g, f, prev, stmp);
               memcpy(slong + (size_t)sl * 18 * 3 * g.vol, stmp, lsz);
            }
         }
         contract_rt(&g, p, r, t, lt, (R > 1) ? slong : lev, mode, acc, wl);
      }
      progress("R", R);
   }
   free(acc); free(ltmp); free(lt); free(stmp); free(slong); free(lev);
   geom_free(&g);
#ifdef _OPENMP
   omp_set_num_threads(saved);
#endif
}

static double now_s(void)
{
#ifdef _OPENMP
   return omp_get_wtime();
#else
   return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/* Wall seconds of one instance of each stage of the reference loop structure
 * on lattice dims (CPU baseline sampling):
 *   t[0] hyp_links on u; t[1] one APE step (or 3-d HYP step);
 *   t[2] one time-transporter update; t[3] one spatial-transporter update of
 *   all levels; t[4] one (R,T) contraction over all level pairs (mode 0),
 *   mean of (R,T) = (1,1) and (mwlr,mwlt). */
void wo_time_stages(const int *dims, int bc, const double *u, const double *aa,
                    const wo_parms *p, int nthreads, double *t)
{
#ifdef _OPENMP
   int saved = omp_get_max_threads();
   if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
   geom g;
   geom_init(&g, dims);
   const int msl = p->msl;
   size_t fsz = sizeof(double) * 18 * 4 * g.vol, lsz = sizeof(double) * 18 * 3 * g.vol;
   double *w = malloc(fsz), *work = malloc(2 * fsz);
   memcpy(w, u, fsz);
   double t0 = now_s();
   wo_hyp(dims, bc, w, aa[0], aa[1], aa[2], p->proj, p->proj_iter, 1, 1);
   t[0] = now_s() - t0;
   double *f = malloc(fsz);
   memcpy(f, w, fsz);
   t0 = now_s();
   if (p->op_smear == 1)
      wo_hyp_spatial(dims, f, p->alpha_hyp[0], p->alpha_hyp[1], p->proj, p->proj_iter);
   else
      ape_step(&g, f, work, p->alpha_ape, p->proj, p->proj_iter, 0);
   t[1] = now_s() - t0;
   double *lev = malloc(lsz * msl), *lt = malloc(sizeof(su3) * g.vol), *lt2 = malloc(sizeof(su3) * g.vol);
   for (int sl = 0; sl < msl; sl++)
      for (long x = 0; x < g.vol; x++)
         for (int k = 0; k < 3; k++)
            memcpy(lev + (size_t)sl * 18 * 3 * g.vol + 18 * (3 * x + k), L4(f, x, k + 1), sizeof(su3));
   for (long x = 0; x < g.vol; x++) memcpy(lt + 18 * x, C4(w, x, 0), sizeof(su3));
   t0 = now_s();
   lt_update(&g, w, lt, lt2);
   t[2] = now_s() - t0;
   double *sl2 = malloc(lsz * msl);
   t0 = now_s();
   for (int sl = 0; sl < msl; sl++)
      sr_update(&g, lev + (size_t)sl * 18 * 3 * g.vol, lev + (size_t)sl * 18 * 3 * g.vol,
                sl2 + (size_t)sl * 18 * 3 * g.vol);
   t[3] = now_s() - t0;
   double *acc = malloc(sizeof(double) * (g.n[0] * msl * msl + 2 * g.n[0]));
   double *wl = malloc(sizeof(double) * msl * msl * p->mwlt * p->mwlr * g.n[0]);
   t0 = now_s();
   contract_rt(&g, p, 0, 0, lt2, sl2, 0, acc, wl);
   contract_rt(&g, p, p->mwlr - 1, p->mwlt - 1, lt2, sl2, 0, acc, wl);
   t[4] = 0.5 * (now_s() - t0);
   free(wl); free(acc); free(sl2); free(lt2); free(lt); free(lev); free(f); free(work); free(w);
   geom_free(&g);
#ifdef _OPENMP
   omp_set_num_threads(saved);
#endif
}

void wo_ms10(const int *dims, int bc, double *u, const double *aa, const wo_parms *p,
             int mode, int nthreads, double *wl)
{
#ifdef _OPENMP
   int saved = omp_get_max_threads();
   if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
   wo_hyp(dims, bc, u, aa[0], aa[1], aa[2], p->proj, p->proj_iter, 1, 1);  /* ms10.c:640 */
   progress("hyp", 1);
   wo_wloop(dims, u, p, mode, nthreads, wl);                                 /* ms10.c:642 */
#ifdef _OPENMP
   omp_set_num_threads(saved);
#endif
}

void wo_set_threads(int n)
{
#ifdef _OPENMP
   if (n > 0) omp_set_num_threads(n);
#else
   (void)n;
#endif
}

/* ------------------------------------------------------------------ */
/* synthetic configuration generator                                    */
/* Philox4x32-10 (Salmon et al., SC'11), Box-Muller normals,            */
/* U = exp(i eps H) by scaled Taylor series + squaring, Gram-Schmidt    */
/* reunitarisation, Haar gauge transform g(x) U_mu(x) g(x+mu)^dag.      */
/* ------------------------------------------------------------------ */

void wo_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
   uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
   uint32_t k0 = key[0], k1 = key[1];
   for (int r = 0; r < 10; r++) {
      uint64_t p0 = (uint64_t)0xD2511F53u * c0;
      uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
      uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
      uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
      c0 = hi1 ^ c1 ^ k0;
      c1 = lo1;
      c2 = hi0 ^ c3 ^ k1;
      c3 = lo0;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
   }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Two N(0,1) numbers from draw `draw` of stream `stream` at `site`. */
static void gauss2(uint64_t seed, uint32_t site, uint32_t stream, uint32_t draw, double *g)
{
   uint32_t ctr[4] = {site, stream, draw, 0x574c4f4fu};
   uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
   uint32_t o[4];
   wo_philox4x32(ctr, key, o);
   const double two53 = 1.0 / 9007199254740992.0;
   double u1 = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * two53;
   double u2 = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * two53;
   double rad = sqrt(-2.0 * log(1.0 - u1));
   double th = 6.283185307179586 * u2;
   g[0] = rad * cos(th);
   g[1] = rad * sin(th);
}

/* Gram-Schmidt on rows 1,2; row 3 = conj(r1 x r2): an SU(3) matrix. */
static void reunitarize(su3 *m)
{
   double n = 0.0;
   for (int b = 0; b < 3; b++) n += m->c[0][b].re * m->c[0][b].re + m->c[0][b].im * m->c[0][b].im;
   n = 1.0 / sqrt(n);
   for (int b = 0; b < 3; b++) { m->c[0][b].re *= n; m->c[0][b].im *= n; }
   double pr = 0.0, pi = 0.0;                    /* p = r1^dag . row2 */
   for (int b = 0; b < 3; b++) {
      pr += m->c[0][b].re * m->c[1][b].re + m->c[0][b].im * m->c[1][b].im;
      pi += m->c[0][b].re * m->c[1][b].im - m->c[0][b].im * m->c[1][b].re;
   }
   for (int b = 0; b < 3; b++) {
      m->c[1][b].re -= pr * m->c[0][b].re - pi * m->c[0][b].im;
      m->c[1][b].im -= pr * m->c[0][b].im + pi * m->c[0][b].re;
   }
   n = 0.0;
   for (int b = 0; b < 3; b++) n += m->c[1][b].re * m->c[1][b].re + m->c[1][b].im * m->c[1][b].im;
   n = 1.0 / sqrt(n);
   for (int b = 0; b < 3; b++) { m->c[1][b].re *= n; m->c[1][b].im *= n; }
   for (int b = 0; b < 3; b++) {
      int b1 = (b + 1) % 3, b2 = (b + 2) % 3;
      cplx x = m->c[0][b1], y = m->c[1][b2], z = m->c[0][b2], w = m->c[1][b1];
      double cr = (x.re * y.re - x.im * y.im) - (z.re * w.re - z.im * w.im);
      double ci = (x.re * y.im + x.im * y.re) - (z.re * w.im + z.im * w.re);
      m->c[2][b].re = cr;
      m->c[2][b].im = -ci;
   }
}

/* exp(i eps H), H Hermitian traceless with N(0,1) real components */
static void random_link(uint64_t seed, uint32_t site, uint32_t mu, double eps, su3 *u)
{
   double g[10];
   for (int d = 0; d < 5; d++) gauss2(seed, site, mu, (uint32_t)d, g + 2 * d);
   double tr = (g[0] + g[1] + g[2]) / 3.0;
   su3 h;
   memset(&h, 0, sizeof(h));
   h.c[0][0].re = g[0] - tr;
   h.c[1][1].re = g[1] - tr;
   h.c[2][2].re = g[2] - tr;
   h.c[0][1].re = g[3]; h.c[0][1].im = g[4];
   h.c[0][2].re = g[5]; h.c[0][2].im = g[6];
   h.c[1][2].re = g[7]; h.c[1][2].im = g[8];
   h.c[1][0].re = g[3]; h.c[1][0].im = -g[4];
   h.c[2][0].re = g[5]; h.c[2][0].im = -g[6];
   h.c[2][1].re = g[7]; h.c[2][1].im = -g[8];
   /* Q = i eps H */
   su3 q;
   for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
         q.c[a][b].re = -eps * h.c[a][b].im;
         q.c[a][b].im = eps * h.c[a][b].re;
      }
   double nrm = 0.0;
   for (int e = 0; e < 18; e++) nrm += ((double *)&q)[e] * ((double *)&q)[e];
   nrm = sqrt(nrm);
   int s = 0;
   while (nrm > 0.25) { nrm *= 0.5; s++; }
   double sc = ldexp(1.0, -s);
   for (int e = 0; e < 18; e++) ((double *)&q)[e] *= sc;
   /* Horner: E = I + Q(I + Q/2(I + Q/3(... (I + Q/12)))) */
   su3 e;
   memset(&e, 0, sizeof(e));
   for (int a = 0; a < 3; a++) e.c[a][a].re = 1.0;
   for (int k = 12; k >= 1; k--) {
      su3 t;
      su3_mul(&q, &e, &t);
      double inv = 1.0 / (double)k;
      for (int x = 0; x < 18; x++) ((double *)&t)[x] *= inv;
      for (int a = 0; a < 3; a++) t.c[a][a].re += 1.0;
      e = t;
   }
   for (int k = 0; k < s; k++) su3_mul(&e, &e, &e);
   reunitarize(&e);
   *u = e;
}

/* Haar-distributed SU(3): Gram-Schmidt on two Gaussian rows. */
void wo_random_su3(uint64_t seed, uint32_t site, uint32_t stream, double *out)
{
   double g[12];
   for (int d = 0; d < 6; d++) gauss2(seed, site, stream, (uint32_t)d, g + 2 * d);
   su3 m;
   memset(&m, 0, sizeof(m));
   for (int b = 0; b < 3; b++) {
      m.c[0][b].re = g[2 * b]; m.c[0][b].im = g[2 * b + 1];
      m.c[1][b].re = g[6 + 2 * b]; m.c[1][b].im = g[7 + 2 * b];
   }
   reunitarize(&m);
   memcpy(out, &m, sizeof(su3));
}

void wo_generate(const int *dims, int bc, uint64_t seed, double eps, double *u)
{
   geom g;
   geom_init(&g, dims);
   double *gt = malloc(sizeof(su3) * g.vol);
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g.vol; x++) {
      for (int mu = 0; mu < 4; mu++) random_link(seed, (uint32_t)x, (uint32_t)mu, eps, L4(u, x, mu));
      wo_random_su3(seed, (uint32_t)x, 4u, gt + 18 * x);
   }
#pragma omp parallel for schedule(static)
   for (long x = 0; x < g.vol; x++)
      for (int mu = 0; mu < 4; mu++) {
         su3 t;
         su3_mul_dag(L4(u, x, mu), (const su3 *)(gt + 18 * g.up[4 * x + mu]), &t);
         su3_mul((const su3 *)(gt + 18 * x), &t, L4(u, x, mu));
      }
   set_bc(&g, bc, u);
   free(gt);
   geom_free(&g);
}

/* ------------------------------------------------------------------ */
/* test helpers                                                         */
/* ------------------------------------------------------------------ */

void wo_gauge_transform(const int *dims, double *u, uint64_t seed)
{
   geom g;
   geom_init(&g, dims);
   double *gt = malloc(sizeof(su3) * g.vol);
   for (long x = 0; x < g.vol; x++) wo_random_su3(seed, (uint32_t)x, 7u, gt + 18 * x);
   for (long x = 0; x < g.vol; x++)
      for (int mu = 0; mu < 4; mu++) {
         su3 t;
         su3_mul_dag(L4(u, x, mu), (const su3 *)(gt + 18 * g.up[4 * x + mu]), &t);
         su3_mul((const su3 *)(gt + 18 * x), &t, L4(u, x, mu));
      }
   free(gt);
   geom_free(&g);
}

/* u'(x) = u(x - s), periodic in every direction */
void wo_shift(const int *dims, double *u, const int *s)
{
   geom g;
   geom_init(&g, dims);
   size_t fsz = sizeof(double) * 18 * 4 * g.vol;
   double *c = malloc(fsz);
   memcpy(c, u, fsz);
   for (long x = 0; x < g.vol; x++) {
      long y = x;
      for (int m = 0; m < 4; m++) {
         int st = ((s[m] % g.n[m]) + g.n[m]) % g.n[m];
         for (int q = 0; q < st; q++) y = g.dn[4 * y + m];
      }
      memcpy(L4(u, x, 0), C4(c, y, 0), 4 * sizeof(su3));
   }
   free(c);
   geom_free(&g);
}

/* sum over x and mu<nu of Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag */
double wo_plaquette(const int *dims, const double *u)
{
   geom g;
   geom_init(&g, dims);
   double sum = 0.0;
   for (long x = 0; x < g.vol; x++)
      for (int mu = 0; mu < 4; mu++)
         for (int nu = mu + 1; nu < 4; nu++) {
            su3 a, b;
            su3_mul(C4(u, x, mu), C4(u, g.up[4 * x + mu], nu), &a);
            su3_mul(C4(u, x, nu), C4(u, g.up[4 * x + nu], mu), &b);
            sum += su3_retr_dag_mul(&b, &a);
         }
   geom_free(&g);
   return sum;
}
