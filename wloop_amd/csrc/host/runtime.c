/*
 * runtime.c -- the openQCD-1.4 pieces the path needs (lattice.h, global.h):
 * runtime lattice sizes, geometry, the double-precision gauge field udfld()
 * and its coherence with the copy resident on the GPU, set_bc, error
 * reporting.  Semantics follow openQCD as used by the reference
 * (hyp_links.c:408-447, wloop.c:139, ms10.c:712-727).
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "internal.h"

int wl_dims[4] = {0, 0, 0, 0};
int wl_volume = 0;
int *ipt = NULL;
int (*iup)[4] = NULL;
int (*idn)[4] = NULL;
int *map = NULL;
int cpr[4] = {0, 0, 0, 0};

static int bc = 3;
static int device = 0;
static su3_dble *ud = NULL;
static wlg_ctx *ctx = NULL, *aux = NULL;
static int host_newer = 1, device_newer = 0;
static int geometry_done = 0;

void error(int test, int no, const char *name, const char *format, ...)
{
   if (test == 0) return;
   va_list ap;
   va_start(ap, format);
   printf("\nError in %s:\n", name);
   vprintf(format, ap);
   va_end(ap);
   printf("\nProgram aborted\n\n");
   fflush(stdout);
   va_start(ap, format);
   fprintf(stderr, "Error in %s: ", name);
   vfprintf(stderr, format, ap);
   fprintf(stderr, "\n");
   va_end(ap);
   exit(no > 0 ? no : 1);
}

void error_root(int test, int no, const char *name, const char *format, ...)
{
   if (test == 0) return;
   va_list ap;
   char msg[1024];
   va_start(ap, format);
   vsnprintf(msg, sizeof(msg), format, ap);
   va_end(ap);
   error(1, no, name, "%s", msg);
}

void error_chk(void) {}

void message(const char *format, ...)
{
   va_list ap;
   va_start(ap, format);
   vprintf(format, ap);
   va_end(ap);
}

void wl_check(int rc, const char *where)
{
   if (rc != 0) error(1, 1, where, "%s", wlg_error());
}

static void free_geometry(void)
{
   free(ipt);
   free(iup);
   free(idn);
   ipt = NULL;
   iup = idn = NULL;
   geometry_done = 0;
}

void set_lattice_sizes(int n0, int n1, int n2, int n3)
{
   int n[4] = {n0, n1, n2, n3};
   for (int m = 0; m < 4; m++)
      error((n[m] < 2) || (n[m] % 2 != 0), 1, "set_lattice_sizes [runtime.c]",
            "Lattice extents must be even and >= 2 (got %dx%dx%dx%d)", n0, n1, n2, n3);
   if (n0 == wl_dims[0] && n1 == wl_dims[1] && n2 == wl_dims[2] && n3 == wl_dims[3]) return;
   if (ctx) wlg_destroy(ctx);
   if (aux) wlg_destroy(aux);
   ctx = aux = NULL;
   free(ud);
   ud = NULL;
   free_geometry();
   wl_release_levels_mirror();
   for (int m = 0; m < 4; m++) wl_dims[m] = n[m];
   wl_volume = n0 * n1 * n2 * n3;
   host_newer = 1;
   device_newer = 0;
}

/* even points first, each parity in lexicographic order (x3 fastest) */
void geometry(void)
{
   error(wl_volume == 0, 1, "geometry [runtime.c]", "Lattice sizes are not set");
   if (geometry_done) return;
   int V = wl_volume;
   ipt = malloc(sizeof(int) * V);
   iup = malloc(sizeof(int[4]) * V);
   idn = malloc(sizeof(int[4]) * V);
   error(!ipt || !iup || !idn, 1, "geometry [runtime.c]", "Unable to allocate index arrays");
   int ne = 0, no = V / 2;
   for (int l = 0; l < V; l++) {
      int r = l, x3 = r % L3; r /= L3;
      int x2 = r % L2; r /= L2;
      int x1 = r % L1;
      int x0 = r / L1;
      ipt[l] = ((x0 + x1 + x2 + x3) % 2 == 0) ? ne++ : no++;
   }
   for (int l = 0; l < V; l++) {
      int x[4], r = l;
      x[3] = r % L3; r /= L3;
      x[2] = r % L2; r /= L2;
      x[1] = r % L1;
      x[0] = r / L1;
      int ix = ipt[l];
      for (int mu = 0; mu < 4; mu++) {
         int y[4], z[4];
         for (int q = 0; q < 4; q++) { y[q] = x[q]; z[q] = x[q]; }
         y[mu] = (x[mu] + 1) % wl_dims[mu];
         z[mu] = (x[mu] + wl_dims[mu] - 1) % wl_dims[mu];
         iup[ix][mu] = ipt[y[3] + L3 * (y[2] + L2 * (y[1] + L1 * y[0]))];
         idn[ix][mu] = ipt[z[3] + L3 * (z[2] + L2 * (z[1] + L1 * z[0]))];
      }
   }
   geometry_done = 1;
}

void set_bc_type(int type)
{
   error((type != 0) && (type != 3), 1, "set_bc_type [runtime.c]",
         "Only open (0) and periodic (3) boundary conditions are supported");
   if (type != bc) {
      if (ctx) wl_levels_guard(ctx);
      if (ctx) wlg_destroy(ctx);
      if (aux) wlg_destroy(aux);
      ctx = aux = NULL;
      host_newer = 1;
      device_newer = 0;
   }
   bc = type;
}

int bc_type(void) { return bc; }

void set_wloop_device(int dev)
{
   if (dev != device) {
      su3_dble *u = udfld();   /* keep the field on the host while switching */
      (void)u;
      if (ctx) wl_levels_guard(ctx);
      if (ctx) wlg_destroy(ctx);
      if (aux) wlg_destroy(aux);
      ctx = aux = NULL;
      host_newer = 1;
   }
   device = dev;
}

static void unit_field(su3_dble *u, int n)
{
   memset(u, 0, sizeof(su3_dble) * n);
   for (int i = 0; i < n; i++) {
      u[i].c11.re = 1.0;
      u[i].c22.re = 1.0;
      u[i].c33.re = 1.0;
   }
}

su3_dble *wl_host_field(void) { return ud; }

static void download_if_needed(void)
{
   if (device_newer && ctx) {
      wl_levels_guard(ctx);
      double *aos = malloc(sizeof(double) * 72 * (size_t)VOLUME);
      error(aos == NULL, 1, "udfld [runtime.c]", "Unable to allocate transfer buffer");
      wl_check(wlg_download(ctx, aos), "udfld [runtime.c]");
      wl_unpack_lex(aos, ud);
      free(aos);
      device_newer = 0;
   }
}

/* the host field; the caller may modify it, so the device copy is refreshed
 * before the next GPU operation */
su3_dble *udfld(void)
{
   geometry();
   if (ud == NULL) {
      ud = malloc(sizeof(su3_dble) * 4 * (size_t)VOLUME);
      error(ud == NULL, 1, "udfld [runtime.c]", "Unable to allocate the gauge field");
      unit_field(ud, 4 * VOLUME);
      host_newer = 1;
   }
   download_if_needed();
   host_newer = 1;
   return ud;
}

void copy_bnd_ud(void) {}

void set_flags(int event)
{
   if (event == UPDATED_UD) host_newer = 1;
}

void set_bc(void)
{
   if (bc != 0) return;
   if (device_newer && ctx) {
      wl_check(wlg_set_bc(ctx), "set_bc [runtime.c]");
      return;
   }
   su3_dble *u = udfld();
   int slice = L1 * L2 * L3;
   for (int l = (L0 - 1) * slice; l < VOLUME; l++)
      memset(lnk(u, ipt[l], 0), 0, sizeof(su3_dble));
}

void wl_pack_lex(const su3_dble *ub, double *aos)
{
   for (int l = 0; l < VOLUME; l++)
      for (int mu = 0; mu < 4; mu++)
         memcpy(aos + 18 * (4 * (size_t)l + mu), lnk((su3_dble *)ub, ipt[l], mu), sizeof(su3_dble));
}

void wl_unpack_lex(const double *aos, su3_dble *ub)
{
   for (int l = 0; l < VOLUME; l++)
      for (int mu = 0; mu < 4; mu++)
         memcpy(lnk(ub, ipt[l], mu), aos + 18 * (4 * (size_t)l + mu), sizeof(su3_dble));
}

static wlg_ctx *new_ctx(void)
{
   wlg_ctx *c = NULL;
   int dims[4] = {L0, L1, L2, L3};
   wl_check(wlg_create(&c, dims, bc, device), "wl_gauge_ctx [runtime.c]");
   return c;
}

wlg_ctx *wl_gauge_ctx_raw(void) { return ctx; }

wlg_ctx *wl_gauge_ctx(void)
{
   geometry();
   if (ud == NULL) udfld();
   if (ctx == NULL) {
      ctx = new_ctx();
      host_newer = 1;
   }
   wl_levels_guard(ctx);
   if (host_newer) {
      double *aos = malloc(sizeof(double) * 72 * (size_t)VOLUME);
      error(aos == NULL, 1, "wl_gauge_ctx [runtime.c]", "Unable to allocate transfer buffer");
      wl_pack_lex(ud, aos);
      wl_check(wlg_upload(ctx, aos), "wl_gauge_ctx [runtime.c]");
      free(aos);
      host_newer = 0;
   }
   return ctx;
}

void wl_device_changed(void)
{
   device_newer = 1;
   host_newer = 0;
}

wlg_ctx *wl_aux_ctx(void)
{
   geometry();
   if (aux == NULL) aux = new_ctx();
   return aux;
}
