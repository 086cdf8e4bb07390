/* internal.h -- shared state of the C host layer (not part of the ABI). */
#ifndef WL_HOST_INTERNAL_H
#define WL_HOST_INTERNAL_H

#include "wloop_gpu.h"
#include "su3.h"

/* device context holding the current gauge field; uploads udfld() first if
 * the host copy is newer.  Every caller is about to run a device operation
 * that may reuse the memory of the operator levels (wl_levels_guard). */
wlg_ctx *wl_gauge_ctx(void);
/* the same context as it is (NULL if none), no upload and no level guard */
wlg_ctx *wl_gauge_ctx_raw(void);
/* keep the operator levels of smear() for get_pus() before c's scratch
 * memory is reused */
void wl_levels_guard(wlg_ctx *c);
/* the device copy of the gauge field has just been changed */
void wl_device_changed(void);
/* host storage of udfld() without coherence side effects (may be NULL) */
su3_dble *wl_host_field(void);
/* a second context of the same lattice for operations on alloc_lnk() fields */
wlg_ctx *wl_aux_ctx(void);
/* openQCD link field (udfld or alloc_lnk layout) <-> lexicographic AoS [V][4][18] */
void wl_pack_lex(const su3_dble *ub, double *aos);
void wl_unpack_lex(const double *aos, su3_dble *ub);
/* abort with the device layer's message if rc != 0 */
void wl_check(int rc, const char *where);
int wl_precision(void);
void wl_release_levels_mirror(void);

#endif
