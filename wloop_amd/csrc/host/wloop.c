/*
 * wloop.c -- wloop_sum, smear and the level table (wloop.h), run on the GPU.
 *
 * wloop_sum (modules/wloop/wloop.c:524-688) measures the current udfld()
 * (already HYP-smeared by the caller, as ms10.c:640-642 does): operator
 * levels, time and spatial transporters and the W_ij(R,T;x0) contraction run
 * on the device, and the wl_idx-ordered result lands in wl.wl_mat.
 * smear() (smear.c:99-121) builds the operator levels on the device;
 * get_pus() (smear.c:124-127) returns the reference's pointer table over a
 * host mirror of the levels, downloaded on demand (a slow path for
 * inspection -- the measurement never needs it).  Levels of an explicit
 * smear() are mirrored before any later GPU operation on the gauge field
 * reuses their device memory; those of a wloop_sum() are mirrored by
 * get_pus() or free_wloop() while they are still on the device.
 */
#include <stdlib.h>
#include <string.h>

#include "global.h"
#include "lattice.h"
#include "lnk.h"
#include "wloop.h"
#include "internal.h"

static int precision = WLG_FP64;
static su3_dble *usb = NULL, **pus = NULL;
/* where the operator levels of the last smear()/wloop_sum() are:
 *   LEV_NONE   none (or released by free_smear)
 *   LEV_DEVICE only on the device (the host mirror is filled on demand)
 *   LEV_HOST   in the host mirror usb/pus
 * lev_keep: the levels came from an explicit smear(), so they are copied to
 * the host before any later operation on the gauge field reuses their device
 * memory (get_pus() then returns them as the reference does) */
enum { LEV_NONE = 0, LEV_DEVICE, LEV_HOST };
static int lev_state = LEV_NONE, lev_keep = 0;

void set_wloop_precision(int p)
{
   error((p != WLG_FP64) && (p != WLG_MIXED), 1, "set_wloop_precision [wloop.c]",
         "Unknown precision %d (0 = FP64, 1 = mixed)", p);
   precision = p;
}

int wl_precision(void) { return precision; }

static wlg_parms to_device(const wloop_parms_t *w)
{
   wlg_parms p;
   memset(&p, 0, sizeof(p));
   p.msl = w->msl;
   for (int i = 0; i < 5; i++) p.nss[i] = w->nss[i];
   p.mwlt = w->mwlt;
   p.mwlr = w->mwlr;
   p.op_smear = (w->op_smear == HYP) ? 1 : 0;
   p.alpha_ape = w->alpha_ape;
   p.alpha_hyp[0] = w->alpha_hyp[0];
   p.alpha_hyp[1] = w->alpha_hyp[1];
   p.proj_iter = w->proj ? w->proj_iter : 0;
   for (int i = 0; i < 3; i++) p.alpha_action[i] = w->alpha_action[i];
   return p;
}

void wl_release_levels_mirror(void)
{
   free(usb);
   free(pus);
   usb = NULL;
   pus = NULL;
   lev_state = LEV_NONE;
}

/* fill usb/pus from the device: levels 0..msl-1 and, when the last
 * measurement formed them (mwlr >= 2), the transporters of its last R in the
 * upper msl slots (pus[msl+sl], wloop.c:207-258); otherwise those stay zero */
static void mirror_levels(wlg_ctx *c, const char *where)
{
   wloop_parms_t w = wloop_parms();
   int msl = w.msl, V = VOLUME, slice = L1 * L2 * L3;
   size_t nmat = (size_t)3 * 2 * msl * V;
   free(usb);
   free(pus);
   usb = calloc(nmat, sizeof(su3_dble));
   pus = malloc(sizeof(su3_dble *) * nmat);
   double *lev = malloc(sizeof(double) * 54 * (size_t)V);
   error(!usb || !pus || !lev, 1, where, "Unable to allocate the level mirror");
   /* storage ordered (sl, x0, spatial site, k) as smear.c:45-76 */
   su3_dble *q = usb;
   for (int sl = 0; sl < 2 * msl; sl++)
      for (int x0 = 0; x0 < L0; x0++)
         for (int ixs = 0; ixs < slice; ixs++)
            for (int k = 0; k < 3; k++) pus[pindxs(sl, ixs, x0, k)] = q++;
   int nsl = (lev_keep || w.mwlr < 2) ? msl : 2 * msl;
   for (int sl = 0; sl < nsl; sl++) {
      wl_check(wlg_level_download(c, sl, lev), where);
      for (int l = 0; l < V; l++)
         for (int k = 0; k < 3; k++)
            memcpy(pus[pindx(sl, ipt[l], k)], lev + 18 * (3 * (size_t)l + k), sizeof(su3_dble));
   }
   free(lev);
   lev_state = LEV_HOST;
}

/* called before an operation reuses the device memory of the levels */
void wl_levels_guard(wlg_ctx *c)
{
   if (lev_state != LEV_DEVICE) return;
   if (lev_keep)
      mirror_levels(c, "smear [wloop.c]");
   else
      lev_state = LEV_NONE;
}

void smear(void)
{
   wloop_parms_t w = wloop_parms();
   error(w.wl_mat == NULL, 1, "smear [wloop.c]", "Wloop parameters are not set");
   wlg_parms p = to_device(&w);
   wlg_ctx *c = wl_gauge_ctx();
   wl_check(wlg_levels(c, &p, precision), "smear [wloop.c]");
   wl_release_levels_mirror();
   lev_state = LEV_DEVICE;
   lev_keep = 1;
}

su3_dble **get_pus(void)
{
   if (lev_state == LEV_HOST) return pus;
   error(lev_state != LEV_DEVICE, 1, "get_pus [wloop.c]",
         "No operator levels: call smear() first (the levels of a wloop_sum() are released by the "
         "next GPU operation on the gauge field)");
   mirror_levels(wl_gauge_ctx_raw(), "get_pus [wloop.c]");
   return pus;
}

void free_smear(void)
{
   wl_release_levels_mirror();
   lev_state = LEV_NONE;
}

/* releases the device scratch; the levels survive on the host as the
 * reference's do (free_wloop does not free them, wloop.c:511-521) */
void free_wloop(void)
{
   wlg_ctx *c = wl_gauge_ctx_raw();
   if (lev_state == LEV_DEVICE) mirror_levels(c, "free_wloop [wloop.c]");
   if (c) wl_check(wlg_trim(c), "free_wloop [wloop.c]");
}

void wloop_sum(int keep)
{
   wloop_parms_t w = wloop_parms();
   error(w.wl_mat == NULL, 1, "wloop_sum [wloop.c]", "Wloop parameters are not set");
   wlg_parms p = to_device(&w);
   wlg_ctx *c = wl_gauge_ctx();
   wl_check(wlg_wloop(c, &p, precision, w.wl_mat), "wloop_sum [wloop.c]");
   wl_release_levels_mirror();
   lev_state = LEV_DEVICE;
   lev_keep = 0;
   if (keep == 0) free_wloop();
}
