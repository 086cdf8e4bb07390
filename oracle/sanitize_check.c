/*
 * sanitize_check.c -- runs every function of the oracle on small lattices, for
 * a build with AddressSanitizer and UndefinedBehaviorSanitizer (test
 * infrastructure: `make -C oracle sanitize`, tests/test_sanitizers.py).
 * Exits 0 when the results pass a few invariants; the sanitizers abort on any
 * out-of-bounds access, use after free or undefined operation.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "wl_oracle.h"

static int fail(const char *what)
{
   fprintf(stderr, "sanitize_check: %s\n", what);
   return 1;
}

int main(void)
{
   const int dims[4] = {6, 4, 4, 4};
   const long V = 6 * 4 * 4 * 4;
   size_t fsz = sizeof(double) * 18 * 4 * (size_t)V;
   double *u = malloc(fsz), *w = malloc(fsz), *a = malloc(fsz), *b = malloc(fsz);
   double *lev = malloc(sizeof(double) * 18 * 3 * (size_t)V * 3);
   if (!u || !w || !a || !b || !lev) return fail("allocation");
   wo_parms p;
   memset(&p, 0, sizeof(p));
   p.msl = 3;
   p.nss[0] = 0;
   p.nss[1] = 1;
   p.nss[2] = 2;
   p.mwlt = 3;
   p.mwlr = 2;
   p.alpha_ape = 0.5;
   p.proj = 1;
   p.proj_iter = 4;
   const int nw = p.msl * p.msl * p.mwlt * p.mwlr * dims[0];
   double *w0 = malloc(sizeof(double) * nw), *w1 = malloc(sizeof(double) * nw);
   if (!w0 || !w1) return fail("allocation");
   const double aa[3] = {1.0, 1.0, 0.5};
   for (int bc = 0; bc <= 3; bc += 3) {
      wo_generate(dims, bc, 11, 0.4, u);
      for (int t = 0; t <= 1; t++) {
         memcpy(w, u, fsz);
         wo_hyp(dims, bc, w, 1.0, 1.0, 0.5, 1, 4, t, 1 - t + t);
         memcpy(w, u, fsz);
         wo_ape(dims, w, 0.5, 1, 4, t);
      }
      memcpy(w, u, fsz);
      wo_hyp_spatial(dims, w, 0.75, 0.6, 1, 4);
      memcpy(a, u, fsz);
      memset(b, 0, fsz);
      wo_staple_sum(dims, u, a, 1, 2, b);
      wo_levels(dims, u, &p, lev);
      memcpy(w, u, fsz);
      wo_ms10(dims, bc, w, aa, &p, 0, 2, w0);
      memcpy(w, u, fsz);
      wo_ms10(dims, bc, w, aa, &p, 1, 2, w1);
      if (memcmp(w0, w1, sizeof(double) * nw) != 0) return fail("modes 0 and 1 differ");
      double t5[5];
      wo_time_stages(dims, bc, u, aa, &p, 2, t5);
   }
   memcpy(w, u, fsz);
   wo_gauge_transform(dims, w, 5);
   const int s[4] = {1, 2, 0, 3};
   wo_shift(dims, w, s);
   if (!(wo_plaquette(dims, u) > 0.0)) return fail("plaquette");
   double m[18];
   for (int e = 0; e < 18; e++) m[e] = 0.1 * (e + 1);
   wo_project(m, 30);
   if (wo_wl_idx(&p, 1, 2, 2, 1, 5) != 1 + 3 * (2 + 3 * (2 + 3 * (1 + 2 * 5))))
      return fail("wl_idx");
   free(u); free(w); free(a); free(b); free(lev); free(w0); free(w1);
   printf("sanitize_check: ok\n");
   return 0;
}
