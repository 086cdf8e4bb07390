// smear.hip -- HYP (4-d, 3 levels), 3-d HYP and APE smearing kernels.
//
// One thread computes one output link.  Staples are summed in the compute
// type TC (double on the plain path, float on the mixed path); the
// lincomb (1-a)U + b*S and the approximate SU(3) projection always run in
// FP64.  Link storage type TS is double or float.
//
// Reference algorithm: modules/smear/hyp_links.c:154-373 (HYP levels),
// modules/smear/hyp_spatial_links.c:140-283 (3-d HYP),
// modules/smear/ape_links.c:147-262 (APE step), staple formula
// modules/lnk/lnkstaple.c:43-56, :318-382.
#include "su3.hpp"
#include "kernels.h"

namespace wl {

// Site and output kind (direction, or direction pair) of this thread in a
// stencil launch.  The launch is one-dimensional with ny kinds per block of 256
// sites, the kind fastest, and every XCD takes a contiguous range of positions
// (xcd_order): the outputs of a site block and of its neighbour blocks, which
// read the same links, are computed close together in time on one L2.  Ordered
// the other way (every site of one kind, then the next kind) the APE step
// fetched 8x its algorithmic bytes from HBM.
__device__ __forceinline__ site_t stencil_site(int ny, int& y) {
  const int b = xcd_order(blockIdx.x, gridDim.x);
  y = b % ny;
  return (site_t)(b / ny) * 256 + threadIdx.x;  // every stencil launch has 256 threads
}

// Wave priority of the stencil kernels.  A thread first loads links and sums
// staples (short products between loads), then combines and projects in FP64
// (thousands of cycles with no memory access).  The staple phase runs at a
// raised priority, so that a wave with loads to issue is not held behind waves
// in their projections; the projections at the base priority then fill the
// SIMD while the loads of the other waves are in flight (at 48^3x96: 7% off
// the APE step, 3-4% off the HYP levels; the results are unchanged).
__device__ __forceinline__ void staple_priority() { __builtin_amdgcn_s_setprio(2); }
__device__ __forceinline__ void projection_priority() { __builtin_amdgcn_s_setprio(0); }

// v += A_nu(x) B_mu(x+nu) A_nu(x+mu)^dag + A_nu(x-nu)^dag B_mu(x-nu) A_nu(x-nu+mu)
// ln: slot of the nu-links in A, lm: slot of the mu-links in B.  PK: A and B
// in the packed FP32 layout.
template <class TC, class TS, bool PK = false>
__device__ __forceinline__ void staple_add(M3<TC>& v, const Lat& L, site_t s, const int x[4],
                                           const v2<TS>* __restrict__ A, int ln,
                                           const v2<TS>* __restrict__ B, int lm, int mu, int nu) {
  M3<TS> a, b, c;
  const site_t xpn = fwd(L, s, x, nu, 1);
  const site_t xpm = fwd(L, s, x, mu, 1);
  ld_link<PK>(b, B, L.V, lm, xpn);
  ld_link<PK>(c, A, L.V, ln, xpm);
  M3<TC> t;
  if constexpr (PK) {
    // packed FP32 links: the third link of each half-staple is loaded after
    // the product of the first two, which fits the APE step in 91 VGPRs (5
    // waves per SIMD instead of 4)
    t = mul_dag(convert<TC>(b), convert<TC>(c));
    __builtin_amdgcn_sched_barrier(0);
    ld_link<PK>(a, A, L.V, ln, s);
  } else {
    ld_link<PK>(a, A, L.V, ln, s);
    t = mul_dag(convert<TC>(b), convert<TC>(c));
  }
  mul_acc(v, convert<TC>(a), t);
  const site_t xmn = bwd(L, s, x, nu);
  const site_t xmnpm = fwd(L, xmn, x, mu, 1);  // (x - nu has the mu coordinate of x: nu != mu)
  ld_link<PK>(b, B, L.V, lm, xmn);
  ld_link<PK>(c, A, L.V, ln, xmnpm);
  if constexpr (PK) {
    t = mul(convert<TC>(b), convert<TC>(c));
    __builtin_amdgcn_sched_barrier(0);
    ld_link<PK>(a, A, L.V, ln, xmn);
  } else {
    ld_link<PK>(a, A, L.V, ln, xmn);
    t = mul(convert<TC>(b), convert<TC>(c));
  }
  dag_mul_acc(v, convert<TC>(a), t);
}

// out = P((1-a)U + b S), S in TC, U in TS, out in TO (TS unless given),
// combination in FP64; PK: U and out in the packed FP32 layout
template <class TC, class TS, class TO = TS, bool PK = false>
__device__ __forceinline__ void decorate_store(const M3<TC>& S, const v2<TS>* __restrict__ U, int lu,
                                               v2<TO>* __restrict__ out, int lo, long V, site_t s,
                                               double alpha, double beta, int proj, int piter) {
  projection_priority();
  M3<TS> u;
  ld_link<PK>(u, U, V, lu, s);
  M3<double> w;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    w.re[e] = alpha * (double)u.re[e] + beta * (double)S.re[e];
    w.im[e] = alpha * (double)u.im[e] + beta * (double)S.im[e];
  }
  if (proj) project(w, piter);
  st_link<PK>(convert<TO>(w), out, V, lo, s);
}

__device__ __forceinline__ int part4(int a, int b) {
  // pair partition {0,1},{2,3} -> 0 ; {0,2},{1,3} -> 1 ; {0,3},{1,2} -> 2
  int d = a + b;
  if (d > 3) d = 6 - d;
  return d - 1;
}

// The decorated links of levels 1 and 2 are stored in TD: the link type TS of
// the field, or FP32 on the mixed path (the staple sums that read them run in
// FP32 there).
//
// Level 1: V1[part{mu,eta}]_mu = P((1-a3) U_mu + a3/2 S_{mu eta}[U,U]);
// kind y = 3*mu + (i-1), eta = (mu+i)&3.  Output slot part*4+mu.
template <class TC, class TS, class TD>
__global__ __launch_bounds__(256) void hyp_level1(Lat L, const v2<TS>* __restrict__ U,
                                                  v2<TD>* __restrict__ D1, double alpha,
                                                  double beta, int proj, int piter) {
  int y;
  const site_t s = stencil_site(12, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = y / 3, eta = (mu + 1 + y % 3) & 3;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  staple_add<TC, TS>(S, L, s, x, U, eta, U, mu, mu, eta);
  decorate_store<TC, TS, TD>(S, U, mu, D1, 4 * part4(mu, eta) + mu, L.V, s, alpha, beta, proj, piter);
}

// Level 2: V2[part{mu,nu}]_mu = P((1-a2)U_mu + a2/4 sum_rho S_{mu rho}[A,B]),
// rho-links A = V1[part{mu,nu}], mu-links B = V1[part{nu,rho}], rho ascending.
template <class TC, class TS, class TD>
__global__ __launch_bounds__(256) void hyp_level2(Lat L, const v2<TS>* __restrict__ U,
                                                  const v2<TD>* __restrict__ D1,
                                                  v2<TD>* __restrict__ D2, double alpha,
                                                  double beta, int proj, int piter) {
  int y;
  const site_t s = stencil_site(12, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = y / 3, nu = (mu + 1 + y % 3) & 3;
  int n = part4(mu, nu);
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  for (int rho = 0; rho < 4; rho++) {
    if (rho == mu || rho == nu) continue;
    staple_add<TC, TD>(S, L, s, x, D1, 4 * n + rho, D1, 4 * part4(nu, rho) + mu, mu, rho);
  }
  decorate_store<TC, TS, TD>(S, U, mu, D2, 4 * n + mu, L.V, s, alpha, beta, proj, piter);
}

// Level 3: V_mu = P((1-a1)U_mu + a1/6 sum_{i=1..3} S_{mu nu}[V2[part{nu,mu}]]),
// nu = (mu+i)&3, written to out (which may be U itself: each thread reads
// only its own U_mu(x) besides D2).  Kind y = mu - mu0 of ny.
template <class TC, class TS, class TD>
__global__ __launch_bounds__(256) void hyp_level3(Lat L, const v2<TS>* U, v2<TS>* out,
                                                  const v2<TD>* __restrict__ D2, int mu0, int ny,
                                                  double alpha, double beta, int proj,
                                                  int piter, int open_bc) {
  int y;
  const site_t s = stencil_site(ny, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = mu0 + y;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  for (int i = 1; i < 4; i++) {
    int nu = (mu + i) & 3;
    int n = part4(nu, mu);
    staple_add<TC, TD>(S, L, s, x, D2, 4 * n + nu, D2, 4 * n + mu, mu, nu);
  }
  decorate_store<TC, TS>(S, U, mu, out, mu, L.V, s, alpha, beta, proj, piter);
  if (open_bc && mu == 0 && x[0] == L.n[0] - 1) {
    M3<TS> z;
    set_zero(z);
    store_u(z, out, L.V, 0, s);  // set_bc: U_0(x0 = N0-1) = 0
  }
}

// 3-d HYP level 1 on a spatial field F (slots k = mu-1):
// V1[slot(mu,eta)]_mu = P((1-a2)U + a2/2 S_{mu eta}[F,F]);  slot(mu,eta) as
// hyp_spatial_links.c:81-86 ({1,2},{2,1},{3,1} -> 0, else 1).
__device__ __forceinline__ int part3(int a, int b) {
  return ((a == 1 && b == 2) || (a == 2 && b == 1) || (a == 3 && b == 1)) ? 0 : 1;
}

template <class TC, class TS>
__global__ __launch_bounds__(256) void hyp3_level1(Lat L, const v2<TS>* __restrict__ F,
                                                   v2<TS>* __restrict__ D, double alpha,
                                                   double beta, int proj, int piter) {
  int y;
  const site_t s = stencil_site(6, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = 1 + y / 2, eta = ((mu + y % 2) % 3) + 1;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  staple_add<TC, TS>(S, L, s, x, F, eta - 1, F, mu - 1, mu, eta);
  decorate_store<TC, TS>(S, F, mu - 1, D, 3 * part3(mu, eta) + mu - 1, L.V, s, alpha, beta, proj,
                         piter);
}

// 3-d HYP level 2: nu-links from slot(nu,mu), mu-links from slot(mu,nu),
// nu = ((mu+i)%3)+1 for i = 0,1 (hyp_spatial_links.c:187-234).
template <class TC, class TS>
__global__ __launch_bounds__(256) void hyp3_level2(Lat L, const v2<TS>* __restrict__ F,
                                                   const v2<TS>* __restrict__ D,
                                                   v2<TS>* __restrict__ out, double alpha,
                                                   double beta, int proj, int piter) {
  int y;
  const site_t s = stencil_site(3, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = 1 + y;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  for (int i = 0; i < 2; i++) {
    int nu = ((mu + i) % 3) + 1;
    staple_add<TC, TS>(S, L, s, x, D, 3 * part3(nu, mu) + nu - 1, D, 3 * part3(mu, nu) + mu - 1, mu,
                       nu);
  }
  decorate_store<TC, TS>(S, F, mu - 1, out, mu - 1, L.V, s, alpha, beta, proj, piter);
}

// One APE step on spatial links (ape_links.c:147-262 with time = 0):
// out_mu = P(in_mu + alpha * sum_{eta != mu, spatial, ascending} S_{mu eta}[in,in]).
// in and out are 3-slot spatial fields (slot k = mu-1); Jacobi update.
// register targets (waves per SIMD without spills): 5 with packed FP32 links
// (staple_add's load order; 2.5% faster than 4), 4 with FP32 links, 3 with
// FP64 links and FP32 staples, 2 in FP64 (which fits 3)
// PK: in and out in the packed FP32 layout (5 float4 planes per slot)
template <class TC, class TS, bool PK = false>
__global__ __launch_bounds__(256, PK ? 5 : sizeof(TS) == 4 ? 4 : (sizeof(TC) == 4 ? 3 : 2)) void ape_step(Lat L, const v2<TS>* __restrict__ in,
                                                v2<TS>* __restrict__ out, double alpha, int proj,
                                                int piter) {
  int y;
  const site_t s = stencil_site(3, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = 1 + y;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  for (int eta = 1; eta < 4; eta++) {
    if (eta == mu) continue;
    staple_add<TC, TS, PK>(S, L, s, x, in, eta - 1, in, mu - 1, mu, eta);
  }
  decorate_store<TC, TS, TS, PK>(S, in, mu - 1, out, mu - 1, L.V, s, 1.0, alpha, proj, piter);
}

// standard <-> packed FP32 layout of nslot link slots, one thread per (slot,
// site); the values are copied unchanged
__global__ __launch_bounds__(256) void pack_links(long V, int nslot, const float2* __restrict__ src,
                                                  float4* __restrict__ dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nslot * V) return;
  const int l = (int)(i / V);
  const long s = i - (long)l * V;
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const float2 a = src[(long)(9 * l + 2 * q) * V + s];
    const float2 b = q < 4 ? src[(long)(9 * l + 2 * q + 1) * V + s] : make_float2(0.f, 0.f);
    dst[(long)(5 * l + q) * V + s] = make_float4(a.x, a.y, b.x, b.y);
  }
}

__global__ __launch_bounds__(256) void unpack_links(long V, int nslot, const float4* __restrict__ src,
                                                    float2* __restrict__ dst) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nslot * V) return;
  const int l = (int)(i / V);
  const long s = i - (long)l * V;
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const float4 v = src[(long)(5 * l + q) * V + s];
    dst[(long)(9 * l + 2 * q) * V + s] = make_float2(v.x, v.y);
    if (q < 4) dst[(long)(9 * l + 2 * q + 1) * V + s] = make_float2(v.z, v.w);
  }
}

// ---- APE step, one thread per site -------------------------------------------
// The three spatial links of a site are smeared by one thread.  The up
// staples of links mu and eta in the plane (mu,eta), mu < eta, share one
// product:
//     up staple of mu  = U_eta(x) Z,      Z = U_mu(x+eta) U_eta(x+mu)^dag
//     up staple of eta = U_mu(x)  Z^dag
// so each plane costs 1 + 2 + 2x2 = 7 products instead of 8 (21 per site),
// and every link the staples of a site read is loaded by one thread.  The
// planes are taken in the order (1,2), (1,3), (2,3); per link the staples are
// then summed with eta ascending, up before down, as in ape_step.

// v += U_eta(x-eta)^dag U_mu(x-eta) U_eta(x-eta+mu)
template <class TC, class TS>
__device__ __forceinline__ void down_staple(M3<TC>& v, const Lat& L, site_t s, const int x[4],
                                            const v2<TS>* __restrict__ in, int mu, int eta) {
  M3<TS> a, b, c;
  const site_t xme = bwd(L, s, x, eta);
  int y[4] = {x[0], x[1], x[2], x[3]};
  y[eta] = (x[eta] == 0) ? L.n[eta] - 1 : x[eta] - 1;
  load_u(b, in, L.V, mu - 1, xme);
  load_u(c, in, L.V, eta - 1, fwd(L, xme, y, mu, 1));
  load_u(a, in, L.V, eta - 1, xme);
  const M3<TC> t = mul(convert<TC>(b), convert<TC>(c));
  dag_mul_acc(v, convert<TC>(a), t);
}

// the up and down staples of links mu and eta (mu < eta) in their plane
template <class TC, class TS>
__device__ __forceinline__ void plane_staples(M3<TC>& Smu, M3<TC>& Seta, const Lat& L, site_t s,
                                              const int x[4], const v2<TS>* __restrict__ in, int mu,
                                              int eta) {
  M3<TC> z;
  {
    M3<TS> b, c;
    const site_t s1 = after(s, Seta);
    load_u(b, in, L.V, mu - 1, fwd(L, s1, x, eta, 1));
    load_u(c, in, L.V, eta - 1, fwd(L, s1, x, mu, 1));
    z = mul_dag(convert<TC>(b), convert<TC>(c));
  }
  M3<TS> a;
  load_u(a, in, L.V, eta - 1, s);
  mul_acc(Smu, convert<TC>(a), z);
  load_u(a, in, L.V, mu - 1, s);
  mul_dag_acc(Seta, convert<TC>(a), z);
  down_staple<TC, TS>(Smu, L, after(s, Seta), x, in, mu, eta);
  down_staple<TC, TS>(Seta, L, after(s, Smu), x, in, eta, mu);
}

// the staple sums S[mu-1] of the three spatial links of a site
template <class TC, class TS>
__device__ __forceinline__ void site_staples(M3<TC> S[3], const Lat& L, site_t s, const int x[4],
                                             const v2<TS>* __restrict__ in) {
  set_zero(S[0]);
  set_zero(S[1]);
  set_zero(S[2]);
  plane_staples<TC, TS>(S[0], S[1], L, s, x, in, 1, 2);
  plane_staples<TC, TS>(S[0], S[2], L, s, x, in, 1, 3);
  plane_staples<TC, TS>(S[1], S[2], L, s, x, in, 2, 3);
}

__device__ __forceinline__ site_t block_site() {
  return (site_t)xcd_order(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
}

template <class TC, class TS>
__global__ __launch_bounds__(256, sizeof(TS) == 4 ? 3 : 2) void ape_site(Lat L, const v2<TS>* __restrict__ in,
                                                                      v2<TS>* __restrict__ out,
                                                                      double alpha, int proj,
                                                                      int piter) {
  const site_t s = block_site();
  if (s >= L.V) return;
  staple_priority();
  int x[4];
  coords(L, s, x);
  M3<TC> S[3];
  site_staples<TC, TS>(S, L, s, x, in);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    __builtin_amdgcn_sched_barrier(0);  // one link's projection at a time
    decorate_store<TC, TS>(S[k], in, k, out, k, L.V, s, 1.0, alpha, proj, piter);
  }
}

// the staple sums of ape_site, stored (3 slots of TC) for ape_project
template <class TC, class TS>
__global__ __launch_bounds__(256, sizeof(TC) == 4 ? 3 : 2) void ape_site_staples(Lat L, const v2<TS>* __restrict__ in,
                                                                              v2<TC>* __restrict__ S) {
  const site_t s = block_site();
  if (s >= L.V) return;
  staple_priority();
  int x[4];
  coords(L, s, x);
  M3<TC> v[3];
  site_staples<TC, TS>(v, L, s, x, in);
#pragma unroll
  for (int k = 0; k < 3; k++) store_u(v[k], S, L.V, k, s);
}

// The APE step in two kernels: the staple sums of every spatial link (in TC,
// stored in TC), then the FP64 lincomb and projection.  The first is light in
// registers and bound by link-load latency, the second is dense FP64 work on
// two streamed links; each runs at the occupancy its own register use allows.
// Same arithmetic as ape_step.
template <class TC, class TS>
__global__ __launch_bounds__(256, sizeof(TC) == 4 ? 4 : 2) void ape_staples(Lat L, const v2<TS>* __restrict__ in,
                                                      v2<TC>* __restrict__ S) {
  int y;
  const site_t s = stencil_site(3, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = 1 + y;
  int x[4];
  coords(L, s, x);
  M3<TC> v;
  set_zero(v);
  for (int eta = 1; eta < 4; eta++) {
    if (eta == mu) continue;
    staple_add<TC, TS>(v, L, s, x, in, eta - 1, in, mu - 1, mu, eta);
  }
  store_u(v, S, L.V, mu - 1, s);
}

template <class TC, class TS>
__global__ __launch_bounds__(256, 3) void ape_project(Lat L, const v2<TS>* __restrict__ in,
                                                      const v2<TC>* __restrict__ S,
                                                      v2<TS>* __restrict__ out, double alpha,
                                                      int proj, int piter) {
  int y;
  const site_t s = stencil_site(3, y);
  if (s >= L.V) return;
  int mu = 1 + y;
  M3<TC> v;
  load_u(v, S, L.V, mu - 1, s);
  decorate_store<TC, TS>(v, in, mu - 1, out, mu - 1, L.V, s, 1.0, alpha, proj, piter);
}

// 4-d APE step (ape_links.c with time = 1) on a 4-slot field.
template <class TC, class TS>
__global__ __launch_bounds__(256) void ape_step4(Lat L, const v2<TS>* __restrict__ in,
                                                 v2<TS>* __restrict__ out, double alpha, int proj,
                                                 int piter) {
  int y;
  const site_t s = stencil_site(4, y);
  if (s >= L.V) return;
  staple_priority();
  int mu = y;
  int x[4];
  coords(L, s, x);
  M3<TC> S;
  set_zero(S);
  for (int eta = 0; eta < 4; eta++) {
    if (eta == mu) continue;
    staple_add<TC, TS>(S, L, s, x, in, eta, in, mu, mu, eta);
  }
  decorate_store<TC, TS>(S, in, mu, out, mu, L.V, s, 1.0, alpha, proj, piter);
}

// Generic staple sum on a 4-slot field: v_mu += S_{mu nu}[A,B] (staple_sum.c:53-87).
__global__ __launch_bounds__(256) void staple_sum_kernel(Lat L, const double2* __restrict__ A,
                                                         const double2* __restrict__ B,
                                                         double2* __restrict__ v, int mu, int nu) {
  const site_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int x[4];
  coords(L, s, x);
  M3<double> S;
  load_u(S, v, L.V, mu, s);
  staple_add<double, double>(S, L, s, x, A, nu, B, mu, mu, nu);
  store_u(S, v, L.V, mu, s);
}

// ---- launchers --------------------------------------------------------------

// stencil launches: ny output kinds per block of 256 sites, one-dimensional
// (stencil_site)
static inline dim3 grid_for(long V, int ny) { return dim3((unsigned)(ny * ((V + 255) / 256)), 1, 1); }

template <class TC, class TS, class TD>
void launch_hyp(const Lat& L, const v2<TS>* U, v2<TS>* out, v2<TD>* D1, v2<TD>* D2,
                const double a[3], int proj, int piter, int time, int space, int open_bc,
                hipStream_t st) {
  hipLaunchKernelGGL((hyp_level1<TC, TS, TD>), grid_for(L.V, 12), dim3(256), 0, st, L, U, D1,
                     1.0 - a[2], a[2] / 2.0, proj, piter);
  hipLaunchKernelGGL((hyp_level2<TC, TS, TD>), grid_for(L.V, 12), dim3(256), 0, st, L, U, D1, D2,
                     1.0 - a[1], a[1] / 4.0, proj, piter);
  int mu0 = 1 - (time > 0), mu1 = 1 + 3 * (space > 0);
  if (mu1 > mu0)
    hipLaunchKernelGGL((hyp_level3<TC, TS, TD>), grid_for(L.V, mu1 - mu0), dim3(256), 0, st, L, U,
                       out, D2, mu0, mu1 - mu0, 1.0 - a[0], a[0] / 6.0, proj, piter, open_bc);
}

template <class TC, class TS>
void launch_hyp3(const Lat& L, const v2<TS>* F, v2<TS>* D, v2<TS>* out, double a1, double a2,
                 int proj, int piter, hipStream_t st) {
  hipLaunchKernelGGL((hyp3_level1<TC, TS>), grid_for(L.V, 6), dim3(256), 0, st, L, F, D,
                     1.0 - a2, a2 / 2.0, proj, piter);
  hipLaunchKernelGGL((hyp3_level2<TC, TS>), grid_for(L.V, 3), dim3(256), 0, st, L, F, D, out,
                     1.0 - a1, a1 / 4.0, proj, piter);
}

template <class TC, class TS>
void launch_ape(const Lat& L, const v2<TS>* in, v2<TS>* out, double alpha, int proj, int piter,
                hipStream_t st) {
  hipLaunchKernelGGL((ape_step<TC, TS>), grid_for(L.V, 3), dim3(256), 0, st, L, in, out, alpha,
                     proj, piter);
}

void launch_ape_packed(const Lat& L, const float4* in, float4* out, double alpha, int proj, int piter,
                       hipStream_t st) {
  hipLaunchKernelGGL((ape_step<float, float, true>), grid_for(L.V, 3), dim3(256), 0, st, L,
                     (const float2*)in, (float2*)out, alpha, proj, piter);
}

void launch_pack_links(long V, int nslot, const float2* src, float4* dst, hipStream_t st) {
  hipLaunchKernelGGL(pack_links, dim3((unsigned)(((long)nslot * V + 255) / 256)), dim3(256), 0, st, V,
                     nslot, src, dst);
}

void launch_unpack_links(long V, int nslot, const float4* src, float2* dst, hipStream_t st) {
  hipLaunchKernelGGL(unpack_links, dim3((unsigned)(((long)nslot * V + 255) / 256)), dim3(256), 0, st,
                     V, nslot, src, dst);
}

template <class TC, class TS>
void launch_ape_site(const Lat& L, const v2<TS>* in, v2<TS>* out, double alpha, int proj, int piter,
                     hipStream_t st) {
  hipLaunchKernelGGL((ape_site<TC, TS>), grid_for(L.V, 1), dim3(256), 0, st, L, in, out, alpha, proj,
                     piter);
}

template <class TC, class TS>
void launch_ape_split(const Lat& L, const v2<TS>* in, v2<TC>* S, v2<TS>* out, double alpha, int proj,
                      int piter, int site, hipStream_t st) {
  if (site)
    hipLaunchKernelGGL((ape_site_staples<TC, TS>), grid_for(L.V, 1), dim3(256), 0, st, L, in, S);
  else
    hipLaunchKernelGGL((ape_staples<TC, TS>), grid_for(L.V, 3), dim3(256), 0, st, L, in, S);
  hipLaunchKernelGGL((ape_project<TC, TS>), grid_for(L.V, 3), dim3(256), 0, st, L, in, S, out, alpha,
                     proj, piter);
}

void launch_ape4(const Lat& L, const double2* in, double2* out, double alpha, int proj, int piter,
                 hipStream_t st) {
  hipLaunchKernelGGL((ape_step4<double, double>), grid_for(L.V, 4), dim3(256), 0, st, L, in, out,
                     alpha, proj, piter);
}

void launch_staple_sum(const Lat& L, const double2* A, const double2* B, double2* v, int mu, int nu,
                       hipStream_t st) {
  hipLaunchKernelGGL(staple_sum_kernel, grid_for(L.V, 1), dim3(256), 0, st, L, A, B, v, mu, nu);
}

template void launch_hyp<double, double, double>(const Lat&, const double2*, double2*, double2*,
                                                  double2*, const double[3], int, int, int, int,
                                                  int, hipStream_t);
template void launch_hyp<float, double, float>(const Lat&, const double2*, double2*, float2*,
                                               float2*, const double[3], int, int, int, int, int,
                                               hipStream_t);
template void launch_hyp3<double, double>(const Lat&, const double2*, double2*, double2*, double,
                                          double, int, int, hipStream_t);
template void launch_hyp3<float, double>(const Lat&, const double2*, double2*, double2*, double,
                                         double, int, int, hipStream_t);
template void launch_ape<double, double>(const Lat&, const double2*, double2*, double, int, int,
                                         hipStream_t);
template void launch_ape<float, double>(const Lat&, const double2*, double2*, double, int, int,
                                        hipStream_t);
template void launch_ape<float, float>(const Lat&, const float2*, float2*, double, int, int,
                                       hipStream_t);
template void launch_ape_split<double, double>(const Lat&, const double2*, double2*, double2*, double,
                                               int, int, int, hipStream_t);
template void launch_ape_split<float, float>(const Lat&, const float2*, float2*, float2*, double, int,
                                             int, int, hipStream_t);
template void launch_ape_site<double, double>(const Lat&, const double2*, double2*, double, int, int,
                                              hipStream_t);
template void launch_ape_site<float, float>(const Lat&, const float2*, float2*, double, int, int,
                                            hipStream_t);
template void launch_hyp3<float, float>(const Lat&, const float2*, float2*, float2*, double, double,
                                        int, int, hipStream_t);

}  // namespace wl
