/*
 * wloop_gpu.h -- device-level C ABI of libwloop_amd.so (MI355X / gfx950).
 *
 * This is the thin layer between the C host code (the module entry points of
 * smear.h / wloop.h / lnk.h and the ms10 driver) and the hand-written HIP
 * kernels.  Plain pointers and sizes only; every function returns 0 on
 * success and a negative code on failure, with the message available from
 * wlg_error().  The legacy entry points turn a failure into an abort through
 * error(), as openQCD's error() does in the reference.
 *
 * Host link fields handed to this layer use the lexicographic AoS image
 * double[V][nl][18] (x3 fastest, mu = 0 time), the memory image of
 * su3_dble per link.  Inside the device a field is stored as planes
 * (structure of arrays) for coalesced access.
 */
#ifndef WLOOP_GPU_H
#define WLOOP_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wlg_ctx wlg_ctx;

/* Measurement parameters; fields as wloop_parms_t (include/wloop.h) plus the
 * static-action HYP parameters.  nss[] is cumulative (smear.c:107-120). */
typedef struct {
   int msl;
   int nss[5];
   int mwlt, mwlr;
   int op_smear;            /* 0 = APE, 1 = HYP */
   double alpha_ape;
   double alpha_hyp[2];
   int proj_iter;           /* projection enabled iff proj_iter > 0 */
   double alpha_action[3];  /* HYP of the static action (ms10.c:640) */
   int hyp_time_only;       /* 0: hyp_links (ms10), 1: hyp_time_links */
} wlg_parms;

enum { WLG_FP64 = 0, WLG_MIXED = 1 };

/* stage timers, milliseconds accumulated since the last wlg_timers_reset */
enum {
   WLG_T_HYP = 0, WLG_T_LEVELS, WLG_T_TIME_TR, WLG_T_SPACE_TR, WLG_T_CONTRACT, WLG_T_FINISH,
   WLG_T_COUNT
};

/* Context for one lattice N0 x N1 x N2 x N3 (N0 = time) on one device;
 * bc = 0 open (U_0 at x0 = N0-1 is zero) or 3 periodic. */
int wlg_create(wlg_ctx **ctx, const int dims[4], int bc, int device);
void wlg_destroy(wlg_ctx *ctx);
const char *wlg_error(void);

/* gauge field transfer (host AoS image [V][4][18]) and synthetic generation */
int wlg_upload(wlg_ctx *ctx, const double *u);
int wlg_download(wlg_ctx *ctx, double *u);
int wlg_generate(wlg_ctx *ctx, uint64_t seed, double eps);

/* HYP smearing of the resident gauge field, in place (hyp_links.c:398-472);
 * proj != 0 applies approx_project_to_su3_dble(piter) after each level */
int wlg_hyp(wlg_ctx *ctx, const double alpha[3], int proj, int piter, int time, int space,
            int precision);
/* one APE step of the resident field in place: spatial (four_d = 0,
 * ape_spatial_links) or all four directions (ape_links), ape_links.c:217-284 */
int wlg_ape_step(wlg_ctx *ctx, double alpha, int proj, int piter, int four_d, int precision);
/* one 3-d HYP step of the resident field's spatial links (hyp_spatial_links.c) */
int wlg_hyp_spatial_step(wlg_ctx *ctx, double a1, double a2, int proj, int piter,
                         int precision);
/* U_0(x0 = N0-1) = 0 on the resident field if the context has open bc (set_bc) */
int wlg_set_bc(wlg_ctx *ctx);
/* sum over x and planes of Re tr U_mu(x)U_nu(x+mu)(U_nu(x)U_mu(x+nu))^dag, time
 * plaquettes at x0 = N0-1 left out for open bc (lnkplaq.c:68-128) */
int wlg_plaquette(wlg_ctx *ctx, double *sum);
/* wlg_plaquette that allocates no memory.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (the per-workgroup partial sums, at most 1024, and the total) is a
 * fixed static device array of 2048 doubles per GPU, held until the process
 * ends; a lattice whose plaquette would need more is rejected with an error.
 * Calls that overlap on one GPU share that array and race: serialise them.
 * Same result, bit for bit, as wlg_plaquette. */
int wlg_plaquette_noalloc(wlg_ctx *ctx, double *sum);
/* v_mu += S_{mu nu}[a,b] on host AoS fields [V][4][18] (staple_sum.c:53-87) */
int wlg_staple_sum(wlg_ctx *ctx, const double *a, const double *b, int mu, int nu, double *v);

/* Operator levels + transporters + contraction on the resident (already
 * HYP-smeared) field: wloop_sum (wloop.c:524-688).  wl receives
 * msl*msl*mwlt*mwlr*N0 doubles in wl_idx order. */
int wlg_wloop(wlg_ctx *ctx, const wlg_parms *p, int precision, double *wl);
/* ms10 set_data (ms10.c:633-646): HYP with alpha_action, then wlg_wloop */
int wlg_measure(wlg_ctx *ctx, const wlg_parms *p, int precision, double *wl);
/* operator levels only (smear.c:99-121) of the resident field */
int wlg_levels(wlg_ctx *ctx, const wlg_parms *p, int precision);
/* spatial links of operator level sl (0 <= sl < msl) of the last
 * wlg_levels/wlg_wloop/wlg_measure, or for msl <= sl < 2 msl the spatial
 * transporters S_R of level sl-msl for the last R (>= 2) of the last
 * wlg_wloop/wlg_measure (the reference's pus[msl+sl] after wloop_sum); host
 * AoS [V][3][18].  Fails once a later operation on the context (upload,
 * generation, smearing, plaquette, staple sum, trim) has reused their memory. */
int wlg_level_download(wlg_ctx *ctx, int sl, double *lev);

/* Batches (no reference counterpart): the full measurement (wlg_measure) of
 * n contexts -- lattices of any sizes, one device -- captured once into one
 * HIP graph whose per-context branches run concurrently.  wlg_batch_run
 * replays it on the fields currently resident in the contexts (refresh them
 * with wlg_generate/wlg_upload between runs) and copies W of context i to
 * wl[i].  Buffers are allocated at creation.  While a context belongs to a
 * batch its buffers are never reallocated or freed: an eager call that needs
 * more memory, wlg_trim and a change of wlg_tune fail with an error instead
 * (destroy the batch first).  Destroy a batch before its contexts. */
typedef struct wlg_batch wlg_batch;
int wlg_batch_create(wlg_batch **batch, int n, wlg_ctx *const *ctxs, const wlg_parms *parms,
                     int precision);
int wlg_batch_run(wlg_batch *batch, double *const *wl);
double wlg_batch_ms(wlg_batch *batch);   /* device time of the last run */
void wlg_batch_destroy(wlg_batch *batch);

/* device bytes one measurement needs (admission by HBM footprint), with the
 * default layout and with the lean layout (WLG_TUNE_LEAN = 1) */
size_t wlg_footprint(const int dims[4], const wlg_parms *p, int precision);
size_t wlg_footprint_lean(const int dims[4], const wlg_parms *p, int precision);

/* launch-shape parameters of a context (0 = library default).  For each
 * precision: the contraction kernel (1 streaming form, 2 chain form with msl
 * lanes per (site,k), 3 chain form with one lane per (site,k), msl <= 4,
 * 4 fused time-slice form, msl <= 4, 5 chain form with the levels of a
 * (site,k) split over two lanes, msl = 2 or 4, else the one-lane form, 6 the
 * fused form for the FP32 T blocks (T <= 8) of the mixed path and the one-lane
 * chain form for the T beyond, msl <= 4, 7 as 6 with the one-lane chain form
 * at T = 9 only and the fused form for T >= 10 forming the open loops afresh at
 * every T from the stored FP32 copies of the FP64 time transporters, 8 the
 * fused form for the FP32 T blocks and the fused form over the stored
 * transporters for every T >= 9; on the FP64 path forms 6 and 7 are form 5
 * and form 8 is form 4; in forms 4 and 8 the first fused launch of each R
 * also forms the next R's spatial transporters; defaults 4 on the FP64 path
 * and 7 on the mixed path, both with 8 start slices) and
 * its workgroup parameter (chunks per
 * (x0,k) for the streaming form, persistent workgroups for the chain forms,
 * workgroups per group of start slices for the fused form; form 6 uses it for
 * both of its launch kinds); for the fused forms also the time-slice depth (T
 * values per launch, 1..16) and the start slices per workgroup (4 or 8; one
 * wave each).  Results do not depend on
 * them beyond floating-point rounding and summation order. */
enum {
   WLG_TUNE_FP64_CONTRACT_KERNEL = 0, WLG_TUNE_FP64_CONTRACT_WG,
   WLG_TUNE_MIXED_CONTRACT_KERNEL, WLG_TUNE_MIXED_CONTRACT_WG,
   WLG_TUNE_FP64_CONTRACT_DEPTH, WLG_TUNE_FP64_CONTRACT_STARTS,
   WLG_TUNE_MIXED_CONTRACT_DEPTH, WLG_TUNE_MIXED_CONTRACT_STARTS,
   WLG_TUNE_APE_FORM,   /* APE kernels: 1 one thread per link, 2 per-link staple sums then
                           lincomb and projection, 3 one thread per site, 4 per-site staple
                           sums then lincomb and projection, 5 one thread per link on FP32
                           links packed two elements per 16 bytes (mixed path, default
                           layout; otherwise form 1).  Default: 5 on the mixed path, 1 on
                           the FP64 path */
   WLG_TUNE_LEAN,       /* 1: least device memory (wlg_footprint_lean): fused time-slice
                           contraction for every T, time transporters kept only at the
                           starts of its T blocks, APE steps through the level buffers */
   WLG_TUNE_MIXED_RESTART, /* mixed path, fused time-slice form for every T: 1 runs the T
                           blocks from T = 9 in FP32, each restarted from an FP32 copy of
                           the FP64 time transporter at its first T, so the time links past
                           that T are accumulated in FP32 inside the open loops (0, the
                           default: those blocks carry the open loops in FP64) */
   WLG_TUNE_FUSED_TRACES, /* fused time-slice launches: 1 takes the msl^2 traces of a step
                           on the vector units, 2 (the default) on the matrix cores (4x4
                           MFMA outer products per pair, pair sums in FP64) */
   WLG_TUNE_COUNT
};
int wlg_tune(wlg_ctx *ctx, int param, int value);
int wlg_tune_get(wlg_ctx *ctx, int param);

int wlg_synchronize(wlg_ctx *ctx);
/* release the scratch memory of the context (the resident field stays) */
int wlg_trim(wlg_ctx *ctx);
void wlg_timers_reset(wlg_ctx *ctx);
double wlg_timer_ms(wlg_ctx *ctx, int stage);
/* number of contraction launches and their summed duration (ms), measured with
 * HIP events on the context's stream */
int wlg_contract_stats(wlg_ctx *ctx, long *launches, double *ms);
/* the same by launch kind: the fused time-slice launches, the per-(R,T) chain
 * (or streaming) launches, the chain launch of each R >= 2 that also forms its
 * spatial transporters, the finish launch of each R, and the fused launches
 * over the stored FP32 transporters of contraction form 7 (its launch count
 * and summed duration, measured with HIP events around each group of launches) */
enum { WLG_L_FUSED = 0, WLG_L_CHAIN, WLG_L_CHAIN_BUILD, WLG_L_FINISH, WLG_L_STORED, WLG_L_COUNT };
int wlg_launch_stats(wlg_ctx *ctx, int kind, long *launches, double *ms);

#ifdef __cplusplus
}
#endif
#endif
