// wloop.hip -- straight-line transporters and the Wilson-loop contraction.
//
// Reference algorithm: modules/wloop/wloop.c
//   compute_u0_long  :153-204   L_T(x) = U_0(x) L_{T-1}(x+0)
//   compute_uk_long  :207-258   S_R(x,k) = F_k(x) S_{R-1}(x+k,k)
//   wloop_fixed_t_r_sl :436-508 Re tr[(L_T(x) S_j(x+T0))^dag (S_i(x) L_T(x+Rk))]
//   wloops           :524-608   normalisation 1/(9 N1 N2 N3), wl_idx store
//
// The contraction shares products between level pairs: per (x,k) it forms
// A_j = L_T(x) S_j(x+T0) and B_i = S_i(x) L_T(x+Rk) once (2 msl products)
// and takes the msl^2 traces; the reference recomputes both products for
// every level pair.  Each product has the same operands in the same order
// as in the reference.
//
// Reduction is deterministic: wavefront butterfly, then the 4 waves of a
// workgroup in fixed order into one partial per (R,T,x0,k,chunk,i,j); a
// final kernel sums the partials in fixed order.
#include "su3.hpp"
#include "kernels.h"

namespace wl {

// ---- time transporter ---------------------------------------------------
// out(x) = U0(x) * prev(x + 0);  U0 and prev are 1-slot fields (slot l0 of U)
template <class TC, class TU, class TP, class TO>
__global__ __launch_bounds__(256) void time_transporter(Lat L, const v2<TU>* __restrict__ U,
                                                        int lu, const v2<TP>* __restrict__ prev,
                                                        v2<TO>* __restrict__ out,
                                                        v2<float>* __restrict__ out32) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int x[4];
  coords(L, s, x);
  M3<TU> u;
  M3<TP> p;
  load(u, U, L.V, lu, s);
  load(p, prev, L.V, 0, fwd(L, s, x, 0, 1));
  M3<TC> w = mul(convert<TC>(u), convert<TC>(p));
  store(convert<TO>(w), out, L.V, 0, s);
  if (out32) store(convert<float>(w), out32, L.V, 0, s);
}

// ---- spatial transporters of all levels --------------------------------
// S_R^i(x,k) = F^i_k(x) * S_{R-1}^i(x+k,k);  blockIdx.y = 3*i + k
template <class TS>
struct LevelPtrs {
  const v2<TS>* p[5];
};

template <class TS>
__global__ __launch_bounds__(256) void space_transporter(Lat L, LevelPtrs<TS> F,
                                                         LevelPtrs<TS> prev,
                                                         v2<TS>* __restrict__ out) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int i = blockIdx.y / 3, k = blockIdx.y % 3;
  int x[4];
  coords(L, s, x);
  M3<TS> f, p;
  load(f, F.p[i], L.V, k, s);
  load(p, prev.p[i], L.V, k, fwd(L, s, x, k + 1, 1));
  store(mul(f, p), out + (long)i * 27 * L.V, L.V, k, s);
}

// ---- contraction ----------------------------------------------------------
// grid: x = chunk, y = k, z = x0.  Each workgroup strides over the spatial
// slice and leaves MSL*MSL partial sums.
template <int MSL, class TS, class TC, class TA>
__global__ __launch_bounds__(256) void contract(Lat L, int R, int T, const v2<TS>* __restrict__ LT,
                                                LevelPtrs<TS> S, TA* __restrict__ partial) {
  const int k = blockIdx.y, x0 = blockIdx.z, nchunk = gridDim.x;
  const long slice = L.st[0];
  const long base = (long)x0 * slice;
  const long sbase = (long)((x0 + T) % L.n[0]) * slice;
  TA acc[MSL * MSL];
#pragma unroll
  for (int q = 0; q < MSL * MSL; q++) acc[q] = TA(0);

  for (long xs = (long)blockIdx.x * 256 + threadIdx.x; xs < slice; xs += (long)nchunk * 256) {
    const long s = base + xs;
    int x[4];
    coords(L, s, x);
    const long sr = fwd(L, s, x, k + 1, R);
    M3<TS> lt, tmp;
    load(lt, LT, L.V, 0, s);
    M3<TC> ltc = convert<TC>(lt);
    M3<TC> A[MSL];
#pragma unroll
    for (int j = 0; j < MSL; j++) {
      load(tmp, S.p[j], L.V, k, sbase + xs);
      A[j] = mul(ltc, convert<TC>(tmp));
    }
    load(lt, LT, L.V, 0, sr);
    ltc = convert<TC>(lt);
#pragma unroll
    for (int i = 0; i < MSL; i++) {
      load(tmp, S.p[i], L.V, k, s);
      M3<TC> B = mul(convert<TC>(tmp), ltc);
#pragma unroll
      for (int j = 0; j < MSL; j++) acc[i * MSL + j] += (TA)retr_dag_mul(A[j], B);
    }
  }

  // wavefront butterfly, then the 4 waves in order
  __shared__ TA red[4][MSL * MSL];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MSL * MSL; q++) {
    TA v = acc[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < MSL * MSL) {
    TA v = red[0][threadIdx.x];
    v += red[1][threadIdx.x];
    v += red[2][threadIdx.x];
    v += red[3][threadIdx.x];
    long idx = (((long)x0 * 3 + k) * nchunk + blockIdx.x) * (MSL * MSL) + threadIdx.x;
    partial[idx] = v;
  }
}

// Sum partials of one (R,T) block: out W_ij(R,T;x0) into wl (wl_idx order).
// partial layout per (R,T): [x0][k][chunk][i*msl+j]
__global__ __launch_bounds__(256) void finish_wloop(int N0, int msl, int mwlt, int mwlr, int nchunk,
                                                    const double* __restrict__ partial,
                                                    double norm, double* __restrict__ wl) {
  long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long nout = (long)msl * msl * mwlt * mwlr * N0;
  if (id >= nout) return;
  // id enumerates wl_idx order: i + msl*(j + msl*(t + mwlt*(r + mwlr*x0)))
  long r_ = id;
  int i = (int)(r_ % msl); r_ /= msl;
  int j = (int)(r_ % msl); r_ /= msl;
  int t = (int)(r_ % mwlt); r_ /= mwlt;
  int r = (int)(r_ % mwlr); r_ /= mwlr;
  int x0 = (int)r_;
  const int mm = msl * msl;
  const double* p = partial + ((long)(r * mwlt + t) * N0 + x0) * 3 * nchunk * mm;
  double sum = 0.0;
  for (int k = 0; k < 3; k++)
    for (int c = 0; c < nchunk; c++) sum += p[((long)k * nchunk + c) * mm + i * msl + j];
  wl[id] = sum / norm;
}

// ---- launchers ------------------------------------------------------------

static inline dim3 grid_for(long V, int ny) { return dim3((unsigned)((V + 255) / 256), ny, 1); }

template <class TC, class TU, class TP, class TO>
void launch_time_transporter(const Lat& L, const v2<TU>* U, int lu, const v2<TP>* prev, v2<TO>* out,
                             float2* out32, hipStream_t st) {
  hipLaunchKernelGGL((time_transporter<TC, TU, TP, TO>), grid_for(L.V, 1), dim3(256), 0, st, L, U,
                     lu, prev, out, out32);
}

template <class TS>
void launch_space_transporter(const Lat& L, int msl, const v2<TS>* const* F,
                              const v2<TS>* const* prev, v2<TS>* out, hipStream_t st) {
  LevelPtrs<TS> f, p;
  for (int i = 0; i < 5; i++) {
    f.p[i] = i < msl ? F[i] : nullptr;
    p.p[i] = i < msl ? prev[i] : nullptr;
  }
  hipLaunchKernelGGL((space_transporter<TS>), grid_for(L.V, 3 * msl), dim3(256), 0, st, L, f, p,
                     out);
}

int contract_chunks(const Lat& L) {
  long slice = L.st[0];
  long maxc = (slice + 255) / 256;
  long want = (2048 + 3L * L.n[0] - 1) / (3L * L.n[0]);
  long c = want < maxc ? want : maxc;
  return (int)(c < 1 ? 1 : c);
}

template <class TS, class TC>
void launch_contract(const Lat& L, int msl, int R, int T, const v2<TS>* LT,
                     const v2<TS>* const* S, double* partial, int nchunk, hipStream_t st) {
  LevelPtrs<TS> sp;
  for (int i = 0; i < 5; i++) sp.p[i] = i < msl ? S[i] : nullptr;
  dim3 grid(nchunk, 3, L.n[0]);
  switch (msl) {
#define CASE(M)                                                                              \
  case M:                                                                                    \
    hipLaunchKernelGGL((contract<M, TS, TC, double>), grid, dim3(256), 0, st, L, R, T, LT, sp, \
                       partial);                                                             \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5)
#undef CASE
  }
}

void launch_finish_wloop(int N0, int msl, int mwlt, int mwlr, int nchunk, const double* partial,
                         double norm, double* wl, hipStream_t st) {
  long nout = (long)msl * msl * mwlt * mwlr * N0;
  hipLaunchKernelGGL(finish_wloop, grid_for(nout, 1), dim3(256), 0, st, N0, msl, mwlt, mwlr,
                     nchunk, partial, norm, wl);
}

template void launch_time_transporter<double, double, double, double>(
    const Lat&, const double2*, int, const double2*, double2*, float2*, hipStream_t);
template void launch_time_transporter<float, float, float, float>(
    const Lat&, const float2*, int, const float2*, float2*, float2*, hipStream_t);
template void launch_time_transporter<double, double, float, double>(
    const Lat&, const double2*, int, const float2*, double2*, float2*, hipStream_t);
template void launch_space_transporter<double>(const Lat&, int, const double2* const*,
                                               const double2* const*, double2*, hipStream_t);
template void launch_space_transporter<float>(const Lat&, int, const float2* const*,
                                              const float2* const*, float2*, hipStream_t);
template void launch_contract<double, double>(const Lat&, int, int, int, const double2*,
                                              const double2* const*, double*, int, hipStream_t);
template void launch_contract<float, float>(const Lat&, int, int, int, const float2*,
                                            const float2* const*, double*, int, hipStream_t);

}  // namespace wl
