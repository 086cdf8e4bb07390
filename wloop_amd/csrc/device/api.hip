// api.hip -- the device-level C ABI (include/wloop_gpu.h): contexts, device
// memory, streams and the measurement pipeline of ms10 set_data
// (main/ms10.c:633-646): HYP -> operator levels -> transporters -> contraction.
//
// Device memory of a context: the gauge field (4 link slots) plus one scratch
// arena carved per phase.  HYP uses 24 slots (its level-1 and level-2
// fields); the measurement uses the operator levels, APE ping-pong slots, the
// time transporters of every T, two generations of spatial transporters and
// the per-workgroup partial sums.  A "slot" is one link per site.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <vector>

#include "kernels.h"
#include "wloop_gpu.h"

using namespace wl;

static thread_local char g_err[512] = "";
static thread_local int g_capturing = 0;

static int set_err(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}

#define CK(call)                                                                          \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return set_err("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define CKL()                                                                             \
  do {                                                                                    \
    hipError_t e_ = hipGetLastError();                                                    \
    if (e_ != hipSuccess)                                                                 \
      return set_err("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

struct wlg_ctx {
  int device;
  hipStream_t st;
  Lat L;
  int bc;
  double2* U;         // gauge field, 4 slots FP64
  double2* Uh;        // HYP-smeared copy made by wlg_measure (input stays intact)
  char* arena;        // scratch
  size_t arena_bytes;
  double* wl_dev;
  size_t wl_bytes;
  double* partial;
  size_t partial_bytes;
  // operator levels of the last measurement (FP64, 3-slot fields)
  const double2* lev[5];
  const float2* lev32[5];
  int msl, lev_fp32;   // msl = 0: no valid levels
  // spatial transporters S_R of the last R >= 2 of the last measurement (the
  // reference's pus[msl+sl] after wloop_sum), valid while slong_ok
  const void* slong[5];
  int slong_ok;
  int batch_refs;      // batches whose captured graph uses this context's buffers
  // timing
  std::vector<hipEvent_t> ev;
  size_t ev_used;
  double t_ms[WLG_T_COUNT];
  int tune[WLG_TUNE_COUNT];  // 0 = default
  long contract_launches;
  double contract_ms;
  long kind_launches[WLG_L_COUNT];  // contraction launches by kind (wlg_launch_stats)
  double kind_ms[WLG_L_COUNT];
  // graph capture: no host synchronisation, timers or allocation; W goes to
  // the pinned buffer wl_pinned
  int graph;
  int cur_msl;  // msl of the measurement being prepared or run
  double* wl_pinned;
  size_t wl_pinned_bytes;
};

extern "C" const char* wlg_error(void) { return g_err; }

int wl_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return -1;
}

int wl_resident_field(wlg_ctx* c, int* dev, const Lat** L, const double2** U, hipStream_t* st,
                      int* open_bc) {
  if (!c) return set_err("no context");
  *dev = c->device;
  *L = &c->L;
  *U = c->U;
  *st = c->st;
  *open_bc = c->bc == 0;
  return 0;
}

static size_t slot64(const Lat& L) { return (size_t)9 * L.V * sizeof(double2); }
static size_t slot32(const Lat& L) { return (size_t)9 * L.V * sizeof(float2); }

// Grow a device buffer of context c to at least `need` bytes.  While a batch
// graph holds the context (its kernels were captured with the current buffer
// addresses) a buffer is never reallocated: the call fails instead.
static int ensure(wlg_ctx* c, char** p, size_t* have, size_t need) {
  if (*have >= need) return 0;
  if (g_capturing) return set_err("device buffer of %zu bytes not prepared before graph capture", need);
  if (c->batch_refs > 0)
    return set_err("the context belongs to %d batch(es) whose graph uses its buffers: destroy the "
                   "batch before measuring it with parameters that need %zu bytes more",
                   c->batch_refs, need - *have);
  if (*p) CK(hipFree(*p));
  *p = nullptr;
  *have = 0;
  CK(hipMalloc((void**)p, need));
  *have = need;
  return 0;
}

// The scratch arena of c, grown to `need` bytes, for an operation that
// overwrites it: the operator levels and spatial transporters of the last
// measurement live there, so they are no longer available afterwards.
static int scratch(wlg_ctx* c, size_t need) {
  c->msl = 0;
  c->slong_ok = 0;
  return ensure(c, &c->arena, &c->arena_bytes, need);
}

static hipEvent_t next_event(wlg_ctx* c) {
  if (c->ev_used == c->ev.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    c->ev.push_back(e);
  }
  return c->ev[c->ev_used++];
}

static int record(wlg_ctx* c, hipEvent_t* e) {
  if (c->graph) {
    *e = nullptr;
    return 0;
  }
  *e = next_event(c);
  if (!*e) return set_err("hipEventCreate failed");
  CK(hipEventRecord(*e, c->st));
  return 0;
}

extern "C" int wlg_create(wlg_ctx** out, const int dims[4], int bc, int device) {
  *out = nullptr;
  for (int m = 0; m < 4; m++)
    if (dims[m] < 2) return set_err("lattice extent N%d = %d must be >= 2", m, dims[m]);
  if (bc != 0 && bc != 3)
    return set_err("only open (0) and periodic (3) boundary conditions are supported");
  int ndev = 0;
  CK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return set_err("device %d not present (%d devices)", device, ndev);
  CK(hipSetDevice(device));
  wlg_ctx* c = new wlg_ctx();
  c->device = device;
  c->bc = bc;
  lat_set(c->L, dims);
  if (c->L.V >= (1L << 28)) {  // 32-bit byte offsets of a site within a field plane
    const long V = c->L.V;
    delete c;
    return set_err("lattice volume %ld exceeds 2^28-1 sites", V);
  }
  hipError_t e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&c->U, 4 * slot64(c->L));
  if (e != hipSuccess) {
    set_err("context allocation failed: %s", hipGetErrorString(e));
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
    return -1;
  }
  *out = c;
  return 0;
}

extern "C" void wlg_destroy(wlg_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->st);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->U) (void)hipFree(c->U);
  if (c->Uh) (void)hipFree(c->Uh);
  if (c->arena) (void)hipFree(c->arena);
  if (c->wl_dev) (void)hipFree(c->wl_dev);
  if (c->wl_pinned) (void)hipHostFree(c->wl_pinned);
  if (c->partial) (void)hipFree(c->partial);
  (void)hipStreamDestroy(c->st);
  delete c;
}

extern "C" int wlg_synchronize(wlg_ctx* c) {
  CK(hipSetDevice(c->device));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

// host AoS [V][nl][18] <-> device planes, staged through the arena
static int upload_field(wlg_ctx* c, const double* h, int nl, double2* dst) {
  size_t bytes = (size_t)c->L.V * nl * 18 * sizeof(double);
  if (scratch(c, bytes)) return -1;
  CK(hipMemcpyAsync(c->arena, h, bytes, hipMemcpyHostToDevice, c->st));
  launch_aos_to_soa(c->L.V, nl, (const double*)c->arena, dst, c->st);
  CKL();
  return 0;
}

static int download_field(wlg_ctx* c, const double2* src, int nl, double* h) {
  size_t bytes = (size_t)c->L.V * nl * 18 * sizeof(double);
  if (scratch(c, bytes)) return -1;
  launch_soa_to_aos(c->L.V, nl, src, (double*)c->arena, c->st);
  CKL();
  CK(hipMemcpyAsync(h, c->arena, bytes, hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_upload(wlg_ctx* c, const double* u) {
  CK(hipSetDevice(c->device));
  if (upload_field(c, u, 4, c->U)) return -1;
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_download(wlg_ctx* c, double* u) {
  CK(hipSetDevice(c->device));
  return download_field(c, c->U, 4, u);
}

extern "C" int wlg_generate(wlg_ctx* c, uint64_t seed, double eps) {
  CK(hipSetDevice(c->device));
  if (scratch(c, slot64(c->L))) return -1;
  launch_generate(c->L, c->U, (double2*)c->arena, seed, eps, c->bc == 0, c->st);
  CKL();
  CK(hipStreamSynchronize(c->st));
  return 0;
}

// HYP of `in` into `out` (out may equal in: hyp_links smears udfld in place)
static int run_hyp(wlg_ctx* c, const double2* in, double2* out, const double a[3], int proj,
                   int piter, int time, int space, int prec) {
  size_t s = slot64(c->L);
  if (out != in) {
    launch_copy((long)36 * c->L.V, in, out, c->st);
    CKL();
  }
  if (a[0] == 0.0) return 0;  // hyp_links.c:405-406
  if (scratch(c, 24 * s)) return -1;
  // decorated links of levels 1 and 2: 12 slots each (FP32 on the mixed path)
  if (prec == WLG_MIXED)
    launch_hyp<float, double, float>(c->L, in, out, (float2*)c->arena, (float2*)(c->arena + 12 * s),
                                     a, proj, piter, time, space, c->bc == 0, c->st);
  else
    launch_hyp<double, double, double>(c->L, in, out, (double2*)c->arena,
                                       (double2*)(c->arena + 12 * s), a, proj, piter, time, space,
                                       c->bc == 0, c->st);
  CKL();
  if (c->bc == 0 && !time) {
    launch_zero_bc(c->L, out, c->st);  // set_bc (hyp_links.c:447)
    CKL();
  }
  return 0;
}

extern "C" int wlg_hyp(wlg_ctx* c, const double alpha[3], int proj, int piter, int time,
                       int space, int prec) {
  CK(hipSetDevice(c->device));
  if (run_hyp(c, c->U, c->U, alpha, proj, piter, time, space, prec)) return -1;
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_ape_step(wlg_ctx* c, double alpha, int proj, int proj_iter, int four_d,
                            int prec) {
  CK(hipSetDevice(c->device));
  size_t s = slot64(c->L);
  if (scratch(c, 4 * s)) return -1;
  double2* tmp = (double2*)c->arena;
  if (four_d) {
    launch_ape4(c->L, c->U, tmp, alpha, proj, proj_iter, c->st);
    CKL();
    CK(hipMemcpyAsync(c->U, tmp, 4 * s, hipMemcpyDeviceToDevice, c->st));
  } else {
    const double2* in = c->U + 9 * c->L.V;
    if (prec == WLG_MIXED)
      launch_ape<float, double>(c->L, in, tmp, alpha, proj, proj_iter, c->st);
    else
      launch_ape<double, double>(c->L, in, tmp, alpha, proj, proj_iter, c->st);
    CKL();
    CK(hipMemcpyAsync(c->U + 9 * c->L.V, tmp, 3 * s, hipMemcpyDeviceToDevice, c->st));
  }
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_hyp_spatial_step(wlg_ctx* c, double a1, double a2, int proj, int proj_iter,
                                    int prec) {
  CK(hipSetDevice(c->device));
  size_t s = slot64(c->L);
  if (scratch(c, 9 * s)) return -1;
  double2* D = (double2*)c->arena;
  double2* out = (double2*)(c->arena + 6 * s);
  const double2* F = c->U + 9 * c->L.V;
  if (prec == WLG_MIXED)
    launch_hyp3<float, double>(c->L, F, D, out, a1, a2, proj, proj_iter, c->st);
  else
    launch_hyp3<double, double>(c->L, F, D, out, a1, a2, proj, proj_iter, c->st);
  CKL();
  CK(hipMemcpyAsync(c->U + 9 * c->L.V, out, 3 * s, hipMemcpyDeviceToDevice, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_staple_sum(wlg_ctx* c, const double* a, const double* b, int mu, int nu,
                              double* v) {
  CK(hipSetDevice(c->device));
  if (mu < 0 || mu > 3 || nu < 0 || nu > 3 || mu == nu)
    return set_err("staple_sum: need 0 <= mu,nu <= 3 and mu != nu (got %d,%d)", mu, nu);
  size_t s = slot64(c->L);
  size_t aos = (size_t)c->L.V * 4 * 18 * sizeof(double);
  if (scratch(c, 12 * s + aos)) return -1;
  double2* A = (double2*)c->arena;
  double2* B = (double2*)(c->arena + 4 * s);
  double2* Vv = (double2*)(c->arena + 8 * s);
  char* stage = c->arena + 12 * s;
  const double* src[3] = {a, b, v};
  double2* dst[3] = {A, B, Vv};
  for (int q = 0; q < 3; q++) {
    CK(hipMemcpyAsync(stage, src[q], aos, hipMemcpyHostToDevice, c->st));
    launch_aos_to_soa(c->L.V, 4, (const double*)stage, dst[q], c->st);
    CKL();
  }
  launch_staple_sum(c->L, A, B, Vv, mu, nu, c->st);
  CKL();
  launch_soa_to_aos(c->L.V, 4, Vv, (double*)stage, c->st);
  CKL();
  CK(hipMemcpyAsync(v, stage, aos, hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));
  return 0;
}

static int check_parms(const wlg_ctx* c, const wlg_parms* p) {
  const Lat& L = c->L;
  if (p->msl < 1 || p->msl > 5) return set_err("max smearing level out of range (msl=%d)", p->msl);
  if (p->mwlt < 1 || p->mwlt > L.n[0] / 2 + 1)
    return set_err("max wloop time extension out of range (mwlt=%d)", p->mwlt);
  for (int m = 1; m < 4; m++)
    if (p->mwlr < 1 || p->mwlr > L.n[m] / 2 + 1)
      return set_err("max wloop spatial extension out of range (mwlr=%d)", p->mwlr);
  if (p->op_smear != 0 && p->op_smear != 1) return set_err("unknown operator smearing %d", p->op_smear);
  for (int i = 0; i < p->msl; i++)
    if (p->nss[i] < 0) return set_err("negative smearing step count nss[%d]=%d", i, p->nss[i]);
  return 0;
}

// first T whose time transporter is accumulated in FP64 on the mixed path
static const int kTimeFp64From = 9;

struct Layout {
  size_t lev, tmp, src32, lt, lt32, lt64, s, total;
  int lev_alias;  // level 0 is the input field itself (nss[0] = 0, lean layout): no storage
  int lean;
  bool lean_lt;   // time transporters kept only at the starts of the fused T blocks
};

// contraction variant: kernel form, its workgroup parameter and, for the
// fused time-slice form, the T depth per launch and start slices per workgroup
struct Contract {
  int kernel, wg, depth, starts, restart;
  int mx;  // fused launches: traces on the matrix cores
};

// T blocks of the fused time-slice form: [Ta, Ta+d-1] covering 1..mwlt; on the
// mixed path the blocks of T <= 8 run in FP32, those from T = 9 in FP64, or,
// restarted (restart = 1), in FP32 from an FP32 copy of the FP64 L_Ta, so that
// no open loop passes through more FP32 products than those of T <= 8
// kind: BLK_FP32 / BLK_FP64 carry the open loops through the block in that
// precision; BLK_STORED forms them afresh at every T from the stored FP32
// copies of the FP64-accumulated time transporters (mixed path)
enum { BLK_FP32 = 0, BLK_FP64 = 1, BLK_STORED = 2 };
struct DepthBlock {
  int Ta, d, kind;
};

static std::vector<DepthBlock> depth_blocks(int prec, int mwlt, int depth, int restart) {
  std::vector<DepthBlock> b;
  const int split = prec == WLG_MIXED ? (mwlt < kTimeFp64From - 1 ? mwlt : kTimeFp64From - 1) : 0;
  for (int T = 1; T <= split; T += depth)
    b.push_back({T, split - T + 1 < depth ? split - T + 1 : depth, BLK_FP32});
  for (int T = split + 1; T <= mwlt; T += depth)
    b.push_back({T, mwlt - T + 1 < depth ? mwlt - T + 1 : depth,
                 (prec == WLG_MIXED && restart) ? BLK_FP32 : BLK_FP64});
  return b;
}

// FP64 time transporters kept for the starts of FP64 blocks on the mixed path
static int depth_fp64_starts(int prec, int mwlt, int depth) {
  if (prec != WLG_MIXED) return 0;
  int n = 0;
  for (const DepthBlock& b : depth_blocks(prec, mwlt, depth, 0)) n += b.kind == BLK_FP64;
  return n;
}

// Arena layout of a measurement.  nb64: FP64 time transporters kept for the
// starts of the FP64 blocks of the fused form on the mixed path.  lean (fused
// contraction for every T): the time transporters are kept only at the starts
// of the fused T blocks (plus two in flight while the chain is built), the
// APE steps of a level alternate between one 3-slot buffer and the level's own
// buffer, and a level 0 that is the input field itself takes no storage.
static Layout wloop_layout(const Lat& L, const wlg_parms* p, int prec, int nb64, int lean,
                           const std::vector<DepthBlock>& blocks, bool fused_all) {
  size_t s64 = slot64(L), s32 = slot32(L);
  Layout o;
  size_t off = 0;
  const size_t sl = (prec == WLG_MIXED) ? s32 : s64;  // storage of levels and transporters
  o.lean = lean;
  o.lev_alias = lean && p->nss[0] == 0;
  o.lev = off; off += (size_t)3 * (p->msl - o.lev_alias) * sl;
  // APE: 2x3 ping-pong + 3 staple sums (lean: 3, through the level buffer);
  // 3-d HYP: its 6-slot intermediate + ping-pong
  o.tmp = off; off += (p->op_smear == 1 ? 12 : (lean ? 3 : 9)) * sl;
  int n32 = 0, n64 = 0;  // block starts past T = 1 that need storage of their own
  for (const DepthBlock& b : blocks)
    if (b.Ta > 1 && b.kind != BLK_STORED) (b.kind == BLK_FP64 ? n64 : n32)++;
  const bool lean_lt = lean && fused_all;  // every T in a fused block: starts only
  if (prec == WLG_MIXED) {
    o.src32 = off; off += 3 * s32;
    o.lt32 = off; off += (size_t)(lean_lt ? 3 + n32 : p->mwlt) * s32;
    o.lt64 = off; off += (size_t)(2 + nb64) * s64;
    o.lt = 0;
    o.s = off; off += (size_t)2 * 3 * p->msl * s32;
  } else {
    o.src32 = o.lt32 = o.lt64 = 0;
    o.lt = off; off += (size_t)(lean_lt ? 2 + n64 : (p->mwlt > 1 ? p->mwlt - 1 : 0)) * s64;
    o.s = off; off += (size_t)2 * 3 * p->msl * s64;
  }
  o.lean_lt = lean_lt;
  o.total = off;
  return o;
}

static int p_msl_gt4(const wlg_ctx* c) { return c->cur_msl > 4; }

// T blocks a contraction form runs with the fused time-slice kernel (internal
// form index = tuning value - 1): every T (index 3); for indices 5 and 6 on the
// mixed path the FP32 blocks T <= 8, and for index 6 also the blocks from
// T = 10 over the stored transporters; for index 7 the FP32 blocks and every T
// from 9 over the stored transporters (on the FP64 path: every T); the
// remaining T run one launch per (R,T) with the chain form (the first of them
// forming S_R)
static std::vector<DepthBlock> fused_blocks(const Contract& k, int prec, int mwlt) {
  std::vector<DepthBlock> b;
  if (k.kernel == 3) return depth_blocks(prec, mwlt, k.depth, k.restart);
  if ((k.kernel == 5 || k.kernel == 6) && prec == WLG_MIXED)
    for (const DepthBlock& x : depth_blocks(prec, mwlt, k.depth, 0))
      if (x.kind == BLK_FP32) b.push_back(x);
  if (k.kernel == 6 && prec == WLG_MIXED)
    for (int T = kTimeFp64From + 1; T <= mwlt; T += 16)
      b.push_back({T, mwlt - T + 1 < 16 ? mwlt - T + 1 : 16, BLK_STORED});
  if (k.kernel == 7) {  // form 8: the FP32 blocks, then every T from 9 over the stored transporters
    if (prec != WLG_MIXED) return depth_blocks(prec, mwlt, k.depth, 0);
    for (const DepthBlock& x : depth_blocks(prec, mwlt, k.depth, 0))
      if (x.kind == BLK_FP32) b.push_back(x);
    for (int T = kTimeFp64From; T <= mwlt; T += 16)
      b.push_back({T, mwlt - T + 1 < 16 ? mwlt - T + 1 : 16, BLK_STORED});
  }
  return b;
}

static Contract contract_choice(const wlg_ctx* c, int prec) {
  const bool mixed = prec == WLG_MIXED;
  const int* t = c->tune;
  const int tk = t[mixed ? WLG_TUNE_MIXED_CONTRACT_KERNEL : WLG_TUNE_FP64_CONTRACT_KERNEL];
  const int tw = t[mixed ? WLG_TUNE_MIXED_CONTRACT_WG : WLG_TUNE_FP64_CONTRACT_WG];
  const int td = t[mixed ? WLG_TUNE_MIXED_CONTRACT_DEPTH : WLG_TUNE_FP64_CONTRACT_DEPTH];
  const int ts = t[mixed ? WLG_TUNE_MIXED_CONTRACT_STARTS : WLG_TUNE_FP64_CONTRACT_STARTS];
  Contract k;
  // defaults (the autotuner's trial winners at 48^3x96): the fused time-slice
  // form for FP64 (one launch per R and block of 16 T, 8 start slices and 64
  // workgroups per start group: 1,050 against 1,655 ms per configuration for
  // the two-lane chain); on the mixed path form 7 (fused FP32 blocks for
  // T <= 8, the one-lane chain at T = 9 forming S_R, the fused form over the
  // stored transporters from T = 10) with 8 start slices and 512 workgroups per
  // start group (form 8, which forms S_{R+1} inside the fused FP32 launch, is
  // within 0.5%; 4 start slices are 2% slower)
  k.kernel = tk > 0 ? tk - 1 : (mixed ? 6 : 3);
  if (c->tune[WLG_TUNE_LEAN] == 1) k.kernel = 3;     // lean layout: fused form for every T
  if (k.kernel >= 2 && p_msl_gt4(c)) k.kernel = 0;  // one-lane and fused forms need msl <= 4
  static const int wg_default[8] = {8, 512, 512, 64, 512, 512, 512, 512};  // per form, measured at 48^3x96
  k.wg = tw > 0 ? tw : wg_default[k.kernel];
  k.depth = td > 0 ? td : 16;
  k.starts = ts > 0 ? ts : 8;
  k.restart = mixed && t[WLG_TUNE_MIXED_RESTART] == 1;
  // traces of the fused launches on the matrix cores unless asked otherwise
  // (48^3x96: FP64 fused launch 31.4 against 36.8 ms; FP32 the same time and
  // the same bits)
  k.mx = t[WLG_TUNE_FUSED_TRACES] != 1;
  // LDS of the fused form: at most 160 KiB per workgroup on gfx950 (FP64
  // time links on the FP64 blocks of the mixed path)
  const int tsb = mixed ? 4 : 8;
  while (k.starts > 4 && contract_depth_lds(c->cur_msl, k.starts, tsb, 8, 0, 16) > 163840) k.starts /= 2;
  return k;
}

// partial sums of the (R,T) launches of one R (W of an R is finished before
// the next R is contracted)
static size_t partial_count(const Lat& L, const wlg_parms* p, const Contract& k) {
  return (size_t)p->mwlt * L.n[0] * contract_slots(L, k.kernel, k.wg) * p->msl * p->msl;
}

static size_t footprint(const int dims[4], const wlg_parms* p, int prec, int lean) {
  Lat L;
  lat_set(L, dims);
  const bool fused = lean && p->msl <= 4;
  // default launch shapes: depth-16 blocks of the fused form; partial sums per
  // x0: the most among the library defaults and the tuning candidates
  // (streaming form: 3 x 32 chunks; chain forms: up to 2048 persistent
  // workgroups; fused form: up to 128 workgroups per start group)
  const Contract k = fused ? Contract{3, 128, 16, 8, 0, 0} : Contract{1, 2048, 16, 8, 0, 0};
  const std::vector<DepthBlock> blocks =
      fused ? depth_blocks(prec, p->mwlt, 16, 0) : std::vector<DepthBlock>();
  Layout o = wloop_layout(L, p, prec, depth_fp64_starts(prec, p->mwlt, 16), lean, blocks, fused);
  size_t hyp = 24 * slot64(L);
  size_t arena = o.total > hyp ? o.total : hyp;
  size_t wls = (size_t)p->msl * p->msl * p->mwlt * p->mwlr * dims[0];
  return 8 * slot64(L) + arena + wls * sizeof(double) + partial_count(L, p, k) * sizeof(double);
}

extern "C" size_t wlg_footprint(const int dims[4], const wlg_parms* p, int prec) {
  return footprint(dims, p, prec, 0);
}

extern "C" size_t wlg_footprint_lean(const int dims[4], const wlg_parms* p, int prec) {
  return footprint(dims, p, prec, 1);
}

// APE kernel form of the operator levels: 1 one thread per link, 2 per-link
// staple sums then lincomb/projection, 3 one thread per site for its three
// links, 4 per-site staple sums then lincomb/projection
// (default: 5 on the mixed path, measured 240 against 251 ms for the 96 steps
// at 48^3x96, 1 on the FP64 path; the lean layout has no scratch slots for
// forms 2, 4 and 5)
static int ape_form(const wlg_ctx* c, int prec) {
  int t = c->tune[WLG_TUNE_APE_FORM];
  if (t == 0) t = (prec == WLG_MIXED) ? 5 : 1;
  if (t == 5 && prec != WLG_MIXED) t = 1;  // the packed layout is for FP32 links
  if (c->tune[WLG_TUNE_LEAN] == 1) {
    if (t == 2 || t == 4) t -= 1;
    if (t == 5) t = 1;
  }
  return t;
}

// operator levels (smear.c:99-121): F_0 = spatial links of U, then nss[sl]
// further APE (or 3-d HYP) steps per level, snapshot after each level.
// TS = double: plain path (levels in FP64, staples in FP64).
// TS = float: mixed path (levels stored in FP32, staples in FP32, the
// projection after every step in FP64).  src is the 3-slot spatial field.
template <class TS>
static int run_levels_t(wlg_ctx* c, const v2<TS>* src, const wlg_parms* p, const Layout& o,
                        char* levbase, char* tmpbase, const v2<TS>** lev) {
  const Lat& L = c->L;
  const int msl = p->msl;
  const size_t lvl = (size_t)27 * L.V;
  int proj = p->proj_iter > 0;
  v2<TS>* levbuf = (v2<TS>*)levbase;
  v2<TS>* tmp0 = (v2<TS>*)tmpbase;
  v2<TS>* tmp1 = tmp0 + lvl;
  v2<TS>* tmpD = tmp0 + 2 * lvl;  // 3-d HYP intermediate (6 slots) / APE staple sums (3)
  if constexpr (sizeof(TS) == 4) {
    // APE form 5: the steps run on the packed FP32 layout in two 3-slot
    // buffers in the scratch slots (2 x 15 V float4 of the 9 x 9 V float2,
    // 256-byte aligned); each level is unpacked into its own storage.  The
    // links and the arithmetic are those of form 1, so the levels have the
    // same bits.
    if (p->op_smear == 0 && !o.lean && ape_form(c, WLG_MIXED) == 5) {
      const int proj = p->proj_iter > 0;
      float4* pk = (float4*)(((uintptr_t)tmpbase + 255) & ~(uintptr_t)255);
      float4* P[2] = {pk, pk + 15 * L.V};
      const float4* cur = nullptr;
      int which = 0;
      for (int sl = 0; sl < msl; sl++) {
        const int n = p->nss[sl];
        if (n == 0) {
          lev[sl] = (sl == 0) ? src : lev[sl - 1];
          continue;
        }
        if (!cur) {
          launch_pack_links(L.V, 3, (const float2*)((sl == 0) ? src : lev[sl - 1]), P[0], c->st);
          CKL();
          cur = P[0];
          which = 1;
        }
        for (int step = 0; step < n; step++) {
          launch_ape_packed(L, cur, P[which], p->alpha_ape, proj, p->proj_iter, c->st);
          CKL();
          cur = P[which];
          which ^= 1;
        }
        v2<TS>* mine = levbuf + (sl - o.lev_alias) * lvl;
        launch_unpack_links(L.V, 3, cur, (float2*)mine, c->st);
        CKL();
        lev[sl] = mine;
      }
      return 0;
    }
  }
  for (int sl = 0; sl < msl; sl++) {
    const v2<TS>* cur = (sl == 0) ? src : lev[sl - 1];
    int n = p->nss[sl];
    if (n == 0) {
      lev[sl] = cur;
      continue;
    }
    v2<TS>* mine = levbuf + (sl - o.lev_alias) * lvl;  // this level's storage
    const bool through_level = o.lean && p->op_smear == 0;
    for (int step = 0; step < n; step++) {
      v2<TS>* dst;
      if (through_level)  // alternate with tmp0 so that the last step lands in mine
        dst = ((n - 1 - step) % 2 == 0) ? mine : tmp0;
      else
        dst = (step == n - 1) ? mine : (step % 2 ? tmp1 : tmp0);
      if (p->op_smear == 1)
        launch_hyp3<TS, TS>(L, cur, tmpD, dst, p->alpha_hyp[0], p->alpha_hyp[1], proj,
                            p->proj_iter, c->st);
      else
        switch (ape_form(c, sizeof(TS) == 4 ? WLG_MIXED : WLG_FP64)) {
          case 2: launch_ape_split<TS, TS>(L, cur, tmpD, dst, p->alpha_ape, proj, p->proj_iter, 0, c->st); break;
          case 3: launch_ape_site<TS, TS>(L, cur, dst, p->alpha_ape, proj, p->proj_iter, c->st); break;
          case 4: launch_ape_split<TS, TS>(L, cur, tmpD, dst, p->alpha_ape, proj, p->proj_iter, 1, c->st); break;
          default: launch_ape<TS, TS>(L, cur, dst, p->alpha_ape, proj, p->proj_iter, c->st); break;
        }
      CKL();
      cur = dst;
    }
    lev[sl] = mine;
  }
  return 0;
}

static int run_levels(wlg_ctx* c, const double2* U, const wlg_parms* p, int prec, const Layout& o,
                      const double2** lev64, const float2** lev32) {
  const Lat& L = c->L;
  char* A = c->arena;
  c->msl = p->msl;
  c->lev_fp32 = (prec == WLG_MIXED);
  if (prec == WLG_MIXED) {
    float2* src = (float2*)(A + o.src32);
    launch_d2f((long)27 * L.V, U + 9 * L.V, src, c->st);  // F_0 in FP32
    CKL();
    if (run_levels_t<float>(c, src, p, o, A + o.lev, A + o.tmp, lev32)) return -1;
    for (int sl = 0; sl < p->msl; sl++) c->lev32[sl] = lev32[sl];
  } else {
    if (run_levels_t<double>(c, U + 9 * L.V, p, o, A + o.lev, A + o.tmp, lev64)) return -1;
    for (int sl = 0; sl < p->msl; sl++) c->lev[sl] = lev64[sl];
  }
  return 0;
}

// pointers of one fused-form block: L_Ta and the time links, in the
// compute precision of the block
struct DepthPtrs {
  const void* lta;
  const void* u0;
};

template <class TS, class TC>
static int run_transport_contract(wlg_ctx* c, const wlg_parms* p, const v2<TS>* const* lev,
                                  const v2<TS>* const* LT, v2<TS>* sbuf0, v2<TS>* sbuf1,
                                  const Contract& k, const std::vector<DepthBlock>& blocks,
                                  const std::vector<DepthPtrs>& bptr, double norm) {
  const Lat& L = c->L;
  const int msl = p->msl;
  const size_t lvl = (size_t)27 * L.V;  // one 3-slot field in 2-vectors
  const size_t pstride = (size_t)L.n[0] * contract_slots(L, k.kernel, k.wg) * msl * msl;
  const v2<TS>* S[5];
  const v2<TS>* prev[5];
  v2<TS>* bufs[2] = {sbuf0, sbuf1};
  // T values the fused blocks do not cover run one chain launch per (R,T)
  std::vector<char> covered(p->mwlt + 1, 0);
  for (const DepthBlock& b : blocks)
    for (int T = b.Ta; T < b.Ta + b.d && T <= p->mwlt; T++) covered[T] = 1;
  std::vector<int> tchain;
  for (int T = 1; T <= p->mwlt; T++)
    if (!covered[T]) tchain.push_back(T);
  // forms 5 and 6 continue with the one-lane chain (mixed) / two-lane chain (FP64)
  const int kc = (k.kernel == 5 || k.kernel == 6) ? (sizeof(TS) == 4 ? 2 : 4) : k.kernel;
  // forms 4 and 8 (every T fused): the first fused launch of each R forms the
  // next R's spatial transporters for its start slices (S_2 in the launches of
  // R = 1, from the levels), so no launch of its own forms them
  const bool ahead = (k.kernel == 3 || k.kernel == 7) && tchain.empty();
  // with the one-lane (FP32 level links) or two-lane (FP64) chain form for the
  // per-(R,T) launches, the first of them forms the spatial transporters of its
  // R (no separate pass over F, S_{R-1} and S_R), ahead of every other launch
  // of the R
  const bool build = ((sizeof(TS) == 4 && kc == 2) || (sizeof(TS) == 8 && kc == 4)) && msl <= 4 &&
                     !tchain.empty();
  for (int r = 0; r < p->mwlr; r++) {
    const int R = r + 1;
    // events per R: start, spatial transporters, chain launch forming S_R,
    // fused launches, fused launches over stored transporters, chain
    // launches, finish (the timers read them in this order)
    hipEvent_t e0, e1, eb, ef, es, ec, e2;
    if (record(c, &e0)) return -1;
    const bool build_r = build && R > 1;
    if (R == 1) {
      for (int i = 0; i < msl; i++) S[i] = lev[i];
    } else if (ahead) {
      for (int i = 0; i < msl; i++) S[i] = bufs[R % 2] + i * lvl;
    } else {
      v2<TS>* out = bufs[R % 2];
      for (int i = 0; i < msl; i++) prev[i] = (R == 2) ? lev[i] : bufs[(R - 1) % 2] + i * lvl;
      if (!build_r) {
        launch_space_transporter<TS>(L, msl, lev, prev, out, c->st);
        CKL();
      }
      for (int i = 0; i < msl; i++) S[i] = out + i * lvl;
    }
    if (record(c, &e1)) return -1;
    size_t cnext = 0;  // next entry of tchain
    if (build_r) {
      const int T = tchain[0];
      launch_contract_build<TS, TC>(L, msl, R, T, LT[T - 1], lev, prev, bufs[R % 2],
                                    c->partial + (size_t)(T - 1) * pstride, k.wg, kc == 4, c->st);
      CKL();
      c->contract_launches += 1;
      c->kind_launches[WLG_L_CHAIN_BUILD] += 1;
      cnext = 1;
    }
    if (record(c, &eb)) return -1;
    // fused blocks in two groups, the blocks over stored transporters last
    v2<TS>* snext = (ahead && R < p->mwlr) ? bufs[(R + 1) % 2] : nullptr;
    for (int group = 0; group < 2; group++) {
      long nl = 0;
      for (size_t q = 0; q < blocks.size(); q++) {
        const DepthBlock& b = blocks[q];
        if ((b.kind == BLK_STORED) != (group == 1)) continue;
        double* part = c->partial + (size_t)(b.Ta - 1) * pstride;
        nl++;
        v2<TS>* sn = snext;  // formed by the first launch of the R
        snext = nullptr;
        if constexpr (sizeof(TS) == 4) {
          if (b.kind == BLK_STORED)
            launch_contract_stored(L, msl, R, b.Ta, b.d, k.starts, k.wg, (const float2*)bptr[q].lta, S,
                                   lev, sn, part, (long)pstride, k.mx, c->st);
          else if (b.kind == BLK_FP64)
            launch_contract_depth<float, double>(L, msl, R, b.Ta, b.d, k.starts, k.wg,
                                                 (const double2*)bptr[q].lta,
                                                 (const double2*)bptr[q].u0, S, lev, sn, part,
                                                 (long)pstride, k.mx, c->st);
          else
            launch_contract_depth<float, float>(L, msl, R, b.Ta, b.d, k.starts, k.wg,
                                                (const float2*)bptr[q].lta,
                                                (const float2*)bptr[q].u0, S, lev, sn, part,
                                                (long)pstride, k.mx, c->st);
        } else {
          launch_contract_depth<double, double>(L, msl, R, b.Ta, b.d, k.starts, k.wg,
                                                (const double2*)bptr[q].lta,
                                                (const double2*)bptr[q].u0, S, lev, sn, part,
                                                (long)pstride, k.mx, c->st);
        }
        CKL();
      }
      c->contract_launches += nl;
      c->kind_launches[group == 0 ? WLG_L_FUSED : WLG_L_STORED] += nl;
      if (record(c, group == 0 ? &ef : &es)) return -1;
    }
    for (size_t q = cnext; q < tchain.size(); q++) {
      const int T = tchain[q];
      launch_contract<TS, TC>(L, msl, R, T, LT[T - 1], S, c->partial + (size_t)(T - 1) * pstride, kc,
                              k.wg, c->st);
      CKL();
    }
    c->contract_launches += (long)(tchain.size() - cnext);
    c->kind_launches[WLG_L_CHAIN] += (long)(tchain.size() - cnext);
    if (record(c, &ec)) return -1;
    // W of this R (normalisation and wl_idx store, wloop.c:535, :584-597)
    launch_finish_wloop(L.n[0], msl, p->mwlt, p->mwlr, r, contract_slots(L, k.kernel, k.wg),
                        c->partial, norm, c->wl_dev, c->st);
    CKL();
    c->kind_launches[WLG_L_FINISH] += 1;
    if (record(c, &e2)) return -1;
  }
  // the transporters of the last R stay in the arena (pus[msl+sl] of the
  // reference after wloop_sum, wloop.c:207-258)
  if (p->mwlr >= 2) {
    for (int i = 0; i < msl; i++) c->slong[i] = S[i];
    c->slong_ok = 1;
  }
  return 0;
}

static int run_wloop(wlg_ctx* c, const double2* U, const wlg_parms* p, int prec, double* wl_host,
                     hipEvent_t e_start) {
  const Lat& L = c->L;
  if (check_parms(c, p)) return -1;
  const int msl = p->msl;
  c->cur_msl = p->msl;
  const Contract k = contract_choice(c, prec);
  const std::vector<DepthBlock> blocks = fused_blocks(k, prec, p->mwlt);
  int nb64 = 0;
  for (const DepthBlock& b : blocks) nb64 += (prec == WLG_MIXED) && b.kind == BLK_FP64;
  Layout o = wloop_layout(L, p, prec, nb64, c->tune[WLG_TUNE_LEAN] == 1, blocks, k.kernel == 3);
  if (scratch(c, o.total)) return -1;
  size_t wls = (size_t)msl * msl * p->mwlt * p->mwlr * L.n[0];
  if (ensure(c, (char**)&c->wl_dev, &c->wl_bytes, wls * sizeof(double))) return -1;
  if (ensure(c, (char**)&c->partial, &c->partial_bytes, partial_count(L, p, k) * sizeof(double)))
    return -1;
  const size_t s64 = slot64(L), s32 = slot32(L);
  char* A = c->arena;

  const double2* lev[5];
  const float2* lev32p[5];
  if (run_levels(c, U, p, prec, o, lev, lev32p)) return -1;
  hipEvent_t e_lev;
  if (record(c, &e_lev)) return -1;

  const double norm = 9.0 * L.n[1] * L.n[2] * L.n[3];  // wloop.c:535

  // ---- time transporters L_T (wloop.c:153-204) ----
  // per-(R,T) forms: every T; fused form: the starts of its T blocks only
  int tmax = p->mwlt;
  if (k.kernel == 3) {
    tmax = 1;
    for (const DepthBlock& b : blocks) tmax = b.Ta > tmax ? b.Ta : tmax;
  }
  std::vector<DepthPtrs> bptr(blocks.size());
  size_t ev_before_rt = 0;
  hipEvent_t e_tt;
  if (prec == WLG_MIXED) {
    float2* lt32 = (float2*)(A + o.lt32);
    double2* ch[2] = {(double2*)(A + o.lt64), (double2*)(A + o.lt64 + s64)};
    double2* starts64 = (double2*)(A + o.lt64 + 2 * s64);  // FP64 L_Ta of the FP64 blocks
    std::vector<const float2*> lts(p->mwlt, nullptr);
    std::vector<const double2*> l64(p->mwlt + 1, nullptr);
    launch_d2f((long)9 * L.V, U, lt32, c->st);  // L_1 = U_0 in FP32
    CKL();
    lts[0] = lt32;
    int m64 = 0, m32 = 0, n32 = 0;
    for (const DepthBlock& b : blocks) n32 += b.kind == BLK_FP32 && b.Ta > 1;
    for (int T = 2; T <= tmax; T++) {
      float2* out32 = lt32 + (size_t)(T - 1) * 9 * L.V;
      bool start32 = false;  // an FP32 block starts at T
      for (const DepthBlock& b : blocks) start32 = start32 || (b.kind == BLK_FP32 && b.Ta == T);
      if (o.lean_lt)  // [L_1][block starts][two in flight]
        out32 = lt32 + (size_t)(start32 ? 1 + m32++ : 1 + n32 + T % 2) * 9 * L.V;
      if (T < kTimeFp64From) {
        launch_time_transporter<float, float, float, float>(L, lt32, 0, lts[T - 2], out32, nullptr,
                                                            c->st);
        lts[T - 1] = out32;
      } else {
        bool start = false;
        for (const DepthBlock& b : blocks) start = start || (b.kind == BLK_FP64 && b.Ta == T);
        double2* dst = start ? starts64 + (size_t)(m64++) * 9 * L.V : ch[T % 2];
        float2* o32 = (k.kernel == 3 && !start32) ? nullptr : out32;
        if (T == kTimeFp64From)
          launch_time_transporter<double, double, float, double>(L, U, 0, lts[T - 2], dst, o32,
                                                                 c->st);
        else
          launch_time_transporter<double, double, double, double>(L, U, 0, l64[T - 1], dst, o32,
                                                                  c->st);
        l64[T] = dst;
        if (o32) lts[T - 1] = out32;
      }
      CKL();
    }
    for (size_t q = 0; q < blocks.size(); q++) {
      if (blocks[q].kind == BLK_FP64) bptr[q] = {l64[blocks[q].Ta], U};
      else bptr[q] = {lts[blocks[q].Ta - 1], blocks[q].kind == BLK_FP32 ? lt32 : nullptr};
    }
    if (record(c, &e_tt)) return -1;
    ev_before_rt = c->ev_used;
    float2* sb = (float2*)(A + o.s);
    if (run_transport_contract<float, float>(c, p, lev32p, lts.data(), sb, sb + 3 * msl * 9 * L.V,
                                             k, blocks, bptr, norm))
      return -1;
    (void)s32;
  } else {
    std::vector<const double2*> lts(p->mwlt, nullptr);
    lts[0] = U;  // slot 0 of the gauge field
    double2* ltb = (double2*)(A + o.lt);
    int m64 = 0, n64 = 0;
    for (const DepthBlock& b : blocks) n64 += b.Ta > 1;
    for (int T = 2; T <= tmax; T++) {
      double2* out = ltb + (size_t)(T - 2) * 9 * L.V;
      if (o.lean_lt) {  // [block starts][two in flight]
        bool start = false;
        for (const DepthBlock& b : blocks) start = start || b.Ta == T;
        out = ltb + (size_t)(start ? m64++ : n64 + T % 2) * 9 * L.V;
      }
      launch_time_transporter<double, double, double, double>(L, U, 0, lts[T - 2], out, nullptr,
                                                              c->st);
      CKL();
      lts[T - 1] = out;
    }
    for (size_t q = 0; q < blocks.size(); q++) bptr[q] = {lts[blocks[q].Ta - 1], U};
    if (record(c, &e_tt)) return -1;
    ev_before_rt = c->ev_used;
    double2* sb = (double2*)(A + o.s);
    if (run_transport_contract<double, double>(c, p, lev, lts.data(), sb, sb + 3 * msl * 9 * L.V,
                                               k, blocks, bptr, norm))
      return -1;
  }
  size_t ev_after_rt = c->ev_used;

  if (c->graph) {
    CK(hipMemcpyAsync(c->wl_pinned, c->wl_dev, wls * sizeof(double), hipMemcpyDeviceToHost, c->st));
    return 0;
  }
  hipEvent_t e_end;
  if (record(c, &e_end)) return -1;
  if (wl_host) CK(hipMemcpyAsync(wl_host, c->wl_dev, wls * sizeof(double), hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));

  // ---- timers ----
  float ms = 0.f;
  if (e_start) {
    CK(hipEventElapsedTime(&ms, e_start, e_lev));
    c->t_ms[WLG_T_LEVELS] += ms;
  }
  CK(hipEventElapsedTime(&ms, e_lev, e_tt));
  c->t_ms[WLG_T_TIME_TR] += ms;
  static const int kinds[5] = {WLG_L_CHAIN_BUILD, WLG_L_FUSED, WLG_L_STORED, WLG_L_CHAIN, WLG_L_FINISH};
  for (size_t q = ev_before_rt; q + 6 < ev_after_rt; q += 7) {  // one R (see run_transport_contract)
    float a = 0.f, b = 0.f;
    CK(hipEventElapsedTime(&a, c->ev[q], c->ev[q + 1]));
    CK(hipEventElapsedTime(&b, c->ev[q + 1], c->ev[q + 6]));
    c->t_ms[WLG_T_SPACE_TR] += a;
    c->t_ms[WLG_T_CONTRACT] += b;
    c->contract_ms += b;
    for (int m = 0; m < 5; m++) {
      float d = 0.f;
      CK(hipEventElapsedTime(&d, c->ev[q + 1 + m], c->ev[q + 2 + m]));
      c->kind_ms[kinds[m]] += d;
    }
  }
  CK(hipEventElapsedTime(&ms, c->ev[ev_after_rt - 1], e_end));
  c->t_ms[WLG_T_FINISH] += ms;
  return 0;
}

extern "C" int wlg_wloop(wlg_ctx* c, const wlg_parms* p, int prec, double* wl) {
  CK(hipSetDevice(c->device));
  c->ev_used = 0;
  hipEvent_t e0;
  if (record(c, &e0)) return -1;
  return run_wloop(c, c->U, p, prec, wl, e0);
}

extern "C" int wlg_measure(wlg_ctx* c, const wlg_parms* p, int prec, double* wl) {
  CK(hipSetDevice(c->device));
  if (check_parms(c, p)) return -1;
  c->ev_used = 0;
  hipEvent_t e0, e1;
  if (record(c, &e0)) return -1;
  if (!c->Uh) CK(hipMalloc((void**)&c->Uh, 4 * slot64(c->L)));
  int piter = p->proj_iter;
  if (run_hyp(c, c->U, c->Uh, p->alpha_action, piter > 0, piter, 1, p->hyp_time_only ? 0 : 1,
              prec))
    return -1;
  if (record(c, &e1)) return -1;
  if (run_wloop(c, c->Uh, p, prec, wl, e1)) return -1;
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, e0, e1));
  c->t_ms[WLG_T_HYP] += ms;
  return 0;
}

extern "C" int wlg_level_download(wlg_ctx* c, int sl, double* lev) {
  CK(hipSetDevice(c->device));
  if (c->msl == 0)
    return set_err("no operator levels on the device: none were built, or a later operation on "
                   "the context reused their memory");
  if (sl < 0 || sl >= 2 * c->msl) return set_err("level %d not available (msl=%d)", sl, c->msl);
  if (sl >= c->msl && !c->slong_ok)
    return set_err("no spatial transporters of R >= 2 from the last measurement (slot %d)", sl);
  size_t bytes = (size_t)c->L.V * 3 * 18 * sizeof(double);
  double* stage = nullptr;
  double2* wide = nullptr;
  CK(hipMalloc((void**)&stage, bytes));
  hipError_t e = hipSuccess;
  const bool up = sl >= c->msl;
  const double2* src = up ? (const double2*)c->slong[sl - c->msl] : c->lev[sl];
  if (c->lev_fp32) {
    e = hipMalloc((void**)&wide, bytes);
    if (e == hipSuccess) {
      launch_f2d((long)27 * c->L.V, up ? (const float2*)c->slong[sl - c->msl] : c->lev32[sl], wide,
                 c->st);
      e = hipGetLastError();
    }
    src = wide;
  }
  if (e == hipSuccess) {
    launch_soa_to_aos(c->L.V, 3, src, stage, c->st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(lev, stage, bytes, hipMemcpyDeviceToHost, c->st);
  if (e == hipSuccess) e = hipStreamSynchronize(c->st);
  (void)hipFree(stage);
  if (wide) (void)hipFree(wide);
  if (e != hipSuccess) return set_err("level download failed: %s", hipGetErrorString(e));
  return 0;
}

extern "C" void wlg_timers_reset(wlg_ctx* c) {
  for (int i = 0; i < WLG_T_COUNT; i++) c->t_ms[i] = 0.0;
  c->contract_ms = 0.0;
  c->contract_launches = 0;
  for (int m = 0; m < WLG_L_COUNT; m++) {
    c->kind_launches[m] = 0;
    c->kind_ms[m] = 0.0;
  }
}

extern "C" double wlg_timer_ms(wlg_ctx* c, int stage) {
  return (stage >= 0 && stage < WLG_T_COUNT) ? c->t_ms[stage] : -1.0;
}

extern "C" int wlg_contract_stats(wlg_ctx* c, long* launches, double* ms) {
  *launches = c->contract_launches;
  *ms = c->contract_ms;
  return 0;
}

extern "C" int wlg_launch_stats(wlg_ctx* c, int kind, long* launches, double* ms) {
  if (kind < 0 || kind >= WLG_L_COUNT) return set_err("unknown launch kind %d", kind);
  *launches = c->kind_launches[kind];
  *ms = c->kind_ms[kind];
  return 0;
}

extern "C" int wlg_levels(wlg_ctx* c, const wlg_parms* p, int prec) {
  CK(hipSetDevice(c->device));
  if (check_parms(c, p)) return -1;
  Layout o = wloop_layout(c->L, p, prec, 0, c->tune[WLG_TUNE_LEAN] == 1, {}, false);
  if (scratch(c, o.total)) return -1;
  const double2* lev[5];
  const float2* lev32[5];
  if (run_levels(c, c->U, p, prec, o, lev, lev32)) return -1;
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_set_bc(wlg_ctx* c) {
  CK(hipSetDevice(c->device));
  if (c->bc == 0) {
    launch_zero_bc(c->L, c->U, c->st);
    CKL();
  }
  CK(hipStreamSynchronize(c->st));
  return 0;
}

extern "C" int wlg_trim(wlg_ctx* c) {
  CK(hipSetDevice(c->device));
  if (c->batch_refs > 0)
    return set_err("the context belongs to %d batch(es): destroy them before releasing its memory",
                   c->batch_refs);
  CK(hipStreamSynchronize(c->st));
  if (c->arena) CK(hipFree(c->arena));
  if (c->partial) CK(hipFree(c->partial));
  if (c->Uh) CK(hipFree(c->Uh));
  c->arena = nullptr;
  c->partial = nullptr;
  c->Uh = nullptr;
  c->arena_bytes = c->partial_bytes = 0;
  c->msl = 0;
  c->slong_ok = 0;
  return 0;
}

extern "C" int wlg_plaquette(wlg_ctx* c, double* sum) {
  CK(hipSetDevice(c->device));
  int nblk = plaquette_blocks(c->L);
  if (scratch(c, (size_t)nblk * sizeof(double))) return -1;
  launch_plaquette(c->L, c->U, c->bc == 0, (double*)c->arena, nblk, c->st);
  CKL();
  std::vector<double> part(nblk);
  CK(hipMemcpyAsync(part.data(), c->arena, nblk * sizeof(double), hipMemcpyDeviceToHost, c->st));
  CK(hipStreamSynchronize(c->st));
  double s = 0.0;
  for (int i = 0; i < nblk; i++) s += part[i];
  *sum = s;
  return 0;
}

extern "C" int wlg_tune(wlg_ctx* c, int param, int value) {
  if (param < 0 || param >= WLG_TUNE_COUNT) return set_err("unknown tuning parameter %d", param);
  if (c->batch_refs > 0 && c->tune[param] != value)
    return set_err("the context belongs to %d batch(es) captured with its launch shapes: destroy "
                   "them before changing a tuning parameter", c->batch_refs);
  if (value < 0) return set_err("tuning value must be >= 0 (0 = default)");
  if ((param == WLG_TUNE_FP64_CONTRACT_KERNEL || param == WLG_TUNE_MIXED_CONTRACT_KERNEL) &&
      value > 8)
    return set_err("contraction kernel must be 0 (default), 1 (streaming), 2 (chain, msl lanes "
                   "per pair), 3 (chain, one lane per pair), 4 (fused time slices), 5 (chain, "
                   "two lanes per pair), 6 (fused time slices for the FP32 T blocks of the "
                   "mixed path, chain form beyond), 7 (as 6, with the T blocks from 10 fused "
                   "over the stored transporters) or 8 (fused time slices for the FP32 T blocks, "
                   "fused over the stored transporters from T = 9)");
  if ((param == WLG_TUNE_FP64_CONTRACT_DEPTH || param == WLG_TUNE_MIXED_CONTRACT_DEPTH) &&
      value > 16)
    return set_err("time-slice depth must be 0 (default) or 1..16");
  if ((param == WLG_TUNE_FP64_CONTRACT_STARTS || param == WLG_TUNE_MIXED_CONTRACT_STARTS) &&
      value != 0 && value != 4 && value != 8)
    return set_err("start slices per workgroup must be 0 (default), 4 or 8");
  if (param == WLG_TUNE_APE_FORM && value > 5)
    return set_err("APE form must be 0 (default), 1 (one thread per link), 2 (staples per link, "
                   "then projection), 3 (one thread per site), 4 (staples per site, then "
                   "projection) or 5 (one thread per link on packed FP32 links, mixed path)");
  if (param == WLG_TUNE_LEAN && value > 1)
    return set_err("lean layout must be 0 (default layout) or 1 (least device memory)");
  if (param == WLG_TUNE_FUSED_TRACES && value > 2)
    return set_err("fused traces must be 0 (default), 1 (vector units) or 2 (matrix cores)");
  if (param == WLG_TUNE_MIXED_RESTART && value > 1)
    return set_err("mixed restart must be 0 (FP64 blocks from T = 9) or 1 (FP32 blocks from T = 9 "
                   "restarted from the FP64 time transporter)");
  if ((param == WLG_TUNE_FP64_CONTRACT_WG || param == WLG_TUNE_MIXED_CONTRACT_WG) && value > 65536)
    return set_err("workgroup parameter must be 0 (default) or 1..65536");
  c->tune[param] = value;
  return 0;
}

extern "C" int wlg_tune_get(wlg_ctx* c, int param) {
  return (param >= 0 && param < WLG_TUNE_COUNT) ? c->tune[param] : -1;
}

// ---- batches: several contexts measured by one captured graph ---------------

// every buffer a measurement touches, allocated ahead of graph capture
static int prepare_measure(wlg_ctx* c, const wlg_parms* p, int prec) {
  if (check_parms(c, p)) return -1;
  const Lat& L = c->L;
  if (!c->Uh) CK(hipMalloc((void**)&c->Uh, 4 * slot64(L)));
  c->cur_msl = p->msl;
  const Contract k = contract_choice(c, prec);
  const std::vector<DepthBlock> blocks = fused_blocks(k, prec, p->mwlt);
  int nb64 = 0;
  for (const DepthBlock& b : blocks) nb64 += (prec == WLG_MIXED) && b.kind == BLK_FP64;
  Layout o = wloop_layout(L, p, prec, nb64, c->tune[WLG_TUNE_LEAN] == 1, blocks, k.kernel == 3);
  size_t hyp = 24 * slot64(L);
  if (scratch(c, o.total > hyp ? o.total : hyp)) return -1;
  size_t wls = (size_t)p->msl * p->msl * p->mwlt * p->mwlr * L.n[0];
  if (ensure(c, (char**)&c->wl_dev, &c->wl_bytes, wls * sizeof(double))) return -1;
  if (ensure(c, (char**)&c->partial, &c->partial_bytes, partial_count(L, p, k) * sizeof(double)))
    return -1;
  if (c->wl_pinned_bytes < wls * sizeof(double)) {
    if (c->wl_pinned) CK(hipHostFree(c->wl_pinned));
    c->wl_pinned = nullptr;
    c->wl_pinned_bytes = 0;
    CK(hipHostMalloc((void**)&c->wl_pinned, wls * sizeof(double), hipHostMallocDefault));
    c->wl_pinned_bytes = wls * sizeof(double);
  }
  return 0;
}

struct wlg_batch {
  int n, prec, device;
  std::vector<wlg_ctx*> ctx;
  std::vector<wlg_parms> parms;
  hipStream_t st;
  hipGraph_t graph;
  hipGraphExec_t exec;
  hipEvent_t fork, t0, t1;
  std::vector<hipEvent_t> join;
  double last_ms;
};

static void batch_free(wlg_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->t0) (void)hipEventDestroy(b->t0);
  if (b->t1) (void)hipEventDestroy(b->t1);
  for (hipEvent_t e : b->join)
    if (e) (void)hipEventDestroy(e);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

extern "C" int wlg_batch_create(wlg_batch** out, int n, wlg_ctx* const* ctxs,
                                const wlg_parms* parms, int prec) {
  *out = nullptr;
  if (n < 1) return set_err("a batch needs at least one context");
  int dev = ctxs[0]->device;
  for (int i = 0; i < n; i++) {
    if (ctxs[i]->device != dev) return set_err("all contexts of a batch must be on one device");
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return set_err("context %d appears twice in the batch", i);
  }
  CK(hipSetDevice(dev));
  for (int i = 0; i < n; i++)
    if (prepare_measure(ctxs[i], &parms[i], prec)) return -1;
  wlg_batch* b = new wlg_batch();
  b->n = n;
  b->prec = prec;
  b->device = dev;
  b->ctx.assign(ctxs, ctxs + n);
  b->parms.assign(parms, parms + n);
  b->join.assign(n, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreate(&b->t0);
  if (e == hipSuccess) e = hipEventCreate(&b->t1);
  for (int i = 0; i < n && e == hipSuccess; i++) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    set_err("batch setup failed: %s", hipGetErrorString(e));
    batch_free(b);
    return -1;
  }
  // capture: fork from the batch stream into every context's stream, run its
  // measurement there, join back.  Lattices that fill the chip by themselves
  // (kFillVolume sites or more) follow one another (at 48^3x96 two concurrent
  // branches ran 891 ms per configuration against 868 one after the other);
  // smaller ones run as concurrent branches beside them.
  const long kFillVolume = 1L << 22;
  e = hipStreamBeginCapture(b->st, hipStreamCaptureModeRelaxed);
  if (e == hipSuccess) e = hipEventRecord(b->fork, b->st);
  int rc = 0;
  int last_big = -1;
  g_capturing = 1;
  for (int i = 0; i < n && e == hipSuccess && rc == 0; i++) {
    wlg_ctx* c = ctxs[i];
    const bool big = c->L.V >= kFillVolume;
    e = hipStreamWaitEvent(c->st, big && last_big >= 0 ? b->join[last_big] : b->fork, 0);
    if (big) last_big = i;
    if (e != hipSuccess) break;
    c->graph = 1;
    const wlg_parms* p = &parms[i];
    rc = run_hyp(c, c->U, c->Uh, p->alpha_action, p->proj_iter > 0, p->proj_iter, 1,
                 p->hyp_time_only ? 0 : 1, prec);
    if (rc == 0) rc = run_wloop(c, c->Uh, p, prec, nullptr, nullptr);
    c->graph = 0;
    if (rc == 0) e = hipEventRecord(b->join[i], c->st);
    if (rc == 0 && e == hipSuccess) e = hipStreamWaitEvent(b->st, b->join[i], 0);
  }
  g_capturing = 0;
  hipError_t e2 = hipStreamEndCapture(b->st, &b->graph);
  if (rc != 0 || e != hipSuccess || e2 != hipSuccess) {
    if (rc == 0) set_err("graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    batch_free(b);
    return -1;
  }
  e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    set_err("graph instantiation failed: %s", hipGetErrorString(e));
    batch_free(b);
    return -1;
  }
  for (int i = 0; i < n; i++) ctxs[i]->batch_refs++;
  *out = b;
  return 0;
}

extern "C" int wlg_batch_run(wlg_batch* b, double* const* wl) {
  CK(hipSetDevice(b->device));
  CK(hipEventRecord(b->t0, b->st));
  CK(hipGraphLaunch(b->exec, b->st));
  CK(hipEventRecord(b->t1, b->st));
  CK(hipStreamSynchronize(b->st));
  float ms = 0.f;
  CK(hipEventElapsedTime(&ms, b->t0, b->t1));
  b->last_ms = ms;
  for (int i = 0; i < b->n; i++) {
    if (!wl || !wl[i]) continue;
    const wlg_parms* p = &b->parms[i];
    size_t wls = (size_t)p->msl * p->msl * p->mwlt * p->mwlr * b->ctx[i]->L.n[0];
    memcpy(wl[i], b->ctx[i]->wl_pinned, wls * sizeof(double));
  }
  return 0;
}

extern "C" double wlg_batch_ms(wlg_batch* b) { return b->last_ms; }

extern "C" void wlg_batch_destroy(wlg_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->st);
  for (wlg_ctx* c : b->ctx) c->batch_refs--;
  batch_free(b);
}
