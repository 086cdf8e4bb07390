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
// as in the reference (with fused multiply-adds).
//
// Reduction is deterministic: wavefront butterfly, per-wave accumulators in
// LDS, the waves of a workgroup in fixed order into one partial per
// (R,T,x0,workgroup,i,j); a final kernel sums the partials in fixed order.
#include <type_traits>

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
  const site_t s = blockIdx.x * blockDim.x + threadIdx.x;
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

// Links to build the spatial transporters of one R inside a contraction
// launch: S_R(x,k) = F(x,k) S_{R-1}(x+k,k) for every level, written to out.
template <class TS>
struct BuildPtrs {
  const v2<TS>* f[5];
  const v2<TS>* prev[5];
  v2<TS>* out[5];
};

// level pointer j for a per-lane j, selected from the kernel-argument
// registers (indexing the argument block with a per-lane index makes the
// compiler load the pointer from memory before every use)
template <class TS>
__device__ __forceinline__ const v2<TS>* level_ptr(const LevelPtrs<TS>& S, int j) {
  const v2<TS>* r = S.p[0];
  r = j == 1 ? S.p[1] : r;
  r = j == 2 ? S.p[2] : r;
  r = j == 3 ? S.p[3] : r;
  r = j == 4 ? S.p[4] : r;
  return r;
}

template <class TS>
__global__ __launch_bounds__(256) void space_transporter(Lat L, LevelPtrs<TS> F,
                                                         LevelPtrs<TS> prev,
                                                         v2<TS>* __restrict__ out) {
  const site_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int i = blockIdx.y / 3, k = blockIdx.y % 3;
  int x[4];
  coords(L, s, x);
  M3<TS> f, p;
  load(f, F.p[i], L.V, k, s);
  load(p, prev.p[i], L.V, k, fwd(L, s, x, k + 1, 1));
  store(mul(f, p), out + (long)i * 27 * L.V, L.V, k, s);
}

// ---- contraction, streaming form ---------------------------------------
// One thread per (site, k) of one source slice; it forms A_j = L_T(x) S_j(x+T0)
// and B_i = S_i(x) L_T(x+Rk) and the msl^2 traces Re tr(A_j^dag B_i).
// grid: x = chunk of the spatial slice, y = k, z = position of the source
// slice along the chains x0 -> x0+T (mod N0), so that a slice read as sink
// S_j(x+T0) is read as source S_i(x) by a workgroup dispatched next to it.
// Partial sums: one per (x0, k, chunk, i, j).
template <int MSL, class TS, class TC>
__global__ __launch_bounds__(256) void contract_stream(Lat L, int R, int T, int g,
                                                       const v2<TS>* __restrict__ LT,
                                                       LevelPtrs<TS> S,
                                                       double* __restrict__ partial) {
  const int len = L.n[0] / g;
  const int k = blockIdx.y, nchunk = gridDim.x;
  const int x0 = (int)(((long)(blockIdx.z / len) + (long)(blockIdx.z % len) * T) % L.n[0]);
  const long slice = L.st[0];
  const long base = (long)x0 * slice;
  const long sbase = (long)((x0 + T) % L.n[0]) * slice;
  double acc[MSL * MSL];
#pragma unroll
  for (int q = 0; q < MSL * MSL; q++) acc[q] = 0.0;

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
      for (int j = 0; j < MSL; j++) acc[i * MSL + j] += (double)retr_dag_mul(A[j], B);
    }
  }

  // wavefront butterfly, then the 4 waves in order
  __shared__ double red[4][MSL * MSL];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < MSL * MSL; q++) {
    double v = acc[q];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < MSL * MSL) {
    double v = red[0][threadIdx.x];
    v += red[1][threadIdx.x];
    v += red[2][threadIdx.x];
    v += red[3][threadIdx.x];
    long idx = (((long)x0 * 3 + k) * nchunk + blockIdx.x) * (MSL * MSL) + threadIdx.x;
    partial[idx] = v;
  }
}

// ---- contraction, chain form --------------------------------------------
// W_ij(R,T;x0) = sum_{x,k} Re tr[S_j(x+T0)^dag D_i(x)],
//   D_i(x) = L_T(x)^dag S_i(x) L_T(x+Rk)
// which equals the reference's Re tr[(L_T(x) S_j(x+T0))^dag (S_i(x) L_T(x+Rk))].
//
// Chains: along x0 -> x0+T (mod N0) the sink links S_j(x0+T) of one step are
// the source links of the next.  A thread owns one (site, k, level i) and
// walks a chain of source slices c, c+T, c+2T, ...: it loads the msl sink
// links once, takes msl traces against its carried D_i, then builds the
// next D_i from its own sink link.  Every level link and time transporter is
// read once per (R,T) launch (plus one extra slice per chain), instead of
// twice.  Lanes of one (site,k) sit next to each other, so their identical
// sink loads coalesce.
//
// Work units are (block of pairs, chain); a persistent workgroup takes units
// w, w+nwg, ... in order.  Per step, the msl sums of each lane are reduced
// over the pairs of the wave (butterfly) and added into a per-wave LDS
// accumulator indexed by x0; at the end the waves are combined in fixed
// order into one partial per (x0, workgroup, i, j).  Deterministic.
template <int MSL, int MSLP, int WAVES, class TS, class TC>
__global__ __launch_bounds__(256) void contract_chain(Lat L, int R, int T, int g,
                                                      const v2<TS>* __restrict__ LT,
                                                      LevelPtrs<TS> S,
                                                      double* __restrict__ partial) {
  constexpr int PPW = 64 / MSLP;      // (site,k) pairs per wave
  constexpr int PPB = PPW * WAVES;    // pairs per block unit
  constexpr int MM = MSL * MSL;
  extern __shared__ double lds_acc[];  // [WAVES][N0][MM]
  const int N0 = L.n[0];
  const int len = N0 / g;
  const long slice = L.st[0];
  const long npairs = 3 * slice;
  const long nblocks = (npairs + PPB - 1) / PPB;
  const long units = nblocks * g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = lane % MSLP;
  const int pw = lane / MSLP;
  double* acc_w = lds_acc + (long)wave * N0 * MM;

  for (int q = lane; q < N0 * MM; q += 64) acc_w[q] = 0.0;

  for (long u = blockIdx.x; u < units; u += gridDim.x) {
    const int c = (int)(u / nblocks);
    const long b = u % nblocks;
    const long pr = b * PPB + (long)wave * PPW + pw;
    const bool live = (pw < PPW) && (i < MSL) && (pr < npairs);
    const int k = live ? (int)(pr / slice) : 0;
    const long xs = live ? pr % slice : 0;
    int x[4];
    coords(L, xs, x);  // x0 = 0 coordinates of the spatial site
    M3<TC> D;
    set_zero(D);
    {
      const long s0 = (long)c * slice + xs;
      x[0] = c;
      M3<TS> si, l0, lr;
      if (live) {
        load(si, level_ptr(S, i), L.V, k, s0);
        load(l0, LT, L.V, 0, s0);
        load(lr, LT, L.V, 0, fwd(L, s0, x, k + 1, R));
        D = mul(dag_mul(convert<TC>(l0), convert<TC>(si)), convert<TC>(lr));
      }
    }
    int x0 = c;
    for (int n = 0; n < len; n++) {
      int x1 = x0 + T;
      if (x1 >= N0) x1 -= N0;
      const long s1 = (long)x1 * slice + xs;
      double tr[MSL];
      M3<TS> own;
      set_zero(own);
#pragma unroll
      for (int j = 0; j < MSL; j++) {
        M3<TS> sj;
        if (live) load(sj, S.p[j], L.V, k, s1);
        else set_zero(sj);
        tr[j] = (double)retr_dag_mul(convert<TC>(sj), D);
        if (j == i) own = sj;
      }
      // reduce over the pairs of the wave: lanes with equal i
#pragma unroll
      for (int j = 0; j < MSL; j++) {
        double v = tr[j];
#pragma unroll
        for (int off = MSLP; off < 64; off <<= 1) v += __shfl_xor(v, off, 64);
        tr[j] = v;
      }
      if (lane < MSL) {
#pragma unroll
        for (int j = 0; j < MSL; j++) acc_w[x0 * MM + lane * MSL + j] += tr[j];
      }
      if (n + 1 < len && live) {
        x[0] = x1;
        M3<TS> l1, lr;
        load(l1, LT, L.V, 0, s1);
        load(lr, LT, L.V, 0, fwd(L, s1, x, k + 1, R));
        D = mul(dag_mul(convert<TC>(l1), convert<TC>(own)), convert<TC>(lr));
      }
      x0 = x1;
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < N0 * MM; q += blockDim.x) {
    double v = lds_acc[q];
#pragma unroll
    for (int w = 1; w < WAVES; w++) v += lds_acc[(long)w * N0 * MM + q];
    const int xo = q / MM, ij = q % MM;
    partial[((long)xo * gridDim.x + blockIdx.x) * MM + ij] = v;
  }
}

// ---- contraction, one-lane chain form ----------------------------------
// As the chain form, but one lane owns a (site, k) pair with the open loops
// D_i of all msl levels, so every sink link is loaded by one lane only.  Per
// step the msl^2 traces (computed in TC) of the 64 lanes of a wave are
// transposed through LDS and summed in FP64 (16 lanes per value, then two
// butterfly steps); the 4 wave sums are added in fixed order into a per-
// workgroup FP64 accumulator indexed by x0 (one barrier per step: all waves
// of a workgroup walk the same chain).  Work units are blocks of 256 sites of
// one direction k, the three directions of a block next to each other, and
// each XCD takes a contiguous range of units: the time transporters L_T(x) and
// L_T(x + R k) of a block are fetched into that XCD's L2 once for the three
// directions.  With BUILD the launch also forms the spatial transporters of its
// R: every level link of a slice is computed as it is loaded, S_R(x,k) =
// F(x,k) S_{R-1}(x+k,k) (the product and operand order of space_transporter),
// and stored for the other launches of the R; the chains visit every slice, so
// every S_R is formed once (the start slice of a chain twice, by the same
// lane).  Requires msl <= 4.
template <int MSL, class TS, class TC, bool BUILD>
__global__ __launch_bounds__(256, 2) void contract_chain1(Lat L, int R, int T, int g, int gb,
                                                       const v2<TS>* __restrict__ LT,
                                                       LevelPtrs<TS> S, BuildPtrs<TS> B,
                                                       double* __restrict__ partial) {
  constexpr int MM = MSL * MSL;
  static_assert(MM <= 16, "one-lane chain form needs msl <= 4");
  extern __shared__ double lds_c1[];
  const int N0 = L.n[0];
  double* acc = lds_c1;                                   // [N0][MM]
  double* red2 = lds_c1 + N0 * MM;                        // [2][4][16]
  TC* redf = (TC*)(red2 + 2 * 4 * 16);                    // [4][64][16], traces in TC
  const int len = N0 / g;
  const long slice = L.st[0];
  const long nbk = (slice + 255) / 256;  // blocks of 256 sites per direction
  const long nblocks = 3 * nbk;
  const long units = nblocks * g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  TC* rw = redf + wave * 64 * 16;

  for (int q = threadIdx.x; q < N0 * MM; q += 256) acc[q] = 0.0;
  __syncthreads();
  int par = 0;
  for (long u = xcd_order(blockIdx.x, gridDim.x); u < units; u += gridDim.x) {
    const int c = (int)(u / nblocks);
    // chain c: segment c / gb of the base chain c % gb (slices = c % gb mod gb)
    const int cs = (int)(((long)(c % gb) + (long)(c / gb) * len * T) % N0);
    const long bl = u % nblocks;
    const int k = (int)(bl % 3);
    const long xs0 = (bl / 3) * 256 + threadIdx.x;
    const bool live = xs0 < slice;
    const long xs = live ? xs0 : 0;
    int x[4];
    coords(L, xs, x);
    // level link i of the pair at site s (BUILD: formed and stored as S_R)
    auto level = [&](M3<TS>& t, int i, long s) {
      if constexpr (BUILD) {
        M3<TS> f, q;
        load(f, B.f[i], L.V, k, s);
        load(q, B.prev[i], L.V, k, fwd(L, s, x, k + 1, 1));
        t = mul(f, q);
        if (live) store(t, B.out[i], L.V, k, s);
      } else {
        load(t, S.p[i], L.V, k, s);
      }
    };
    M3<TC> D[MSL], Sx[MSL];
    {
      const long s0 = (long)cs * slice + xs;
      M3<TS> a, b, t;
      load(a, LT, L.V, 0, s0);
      load(b, LT, L.V, 0, fwd(L, s0, x, k + 1, R));
      M3<TC> ac = convert<TC>(a), bc = convert<TC>(b);
#pragma unroll
      for (int i = 0; i < MSL; i++) {
        level(t, i, s0);
        D[i] = mul(dag_mul(ac, convert<TC>(t)), bc);
      }
    }
    int x0 = cs;
    int x1 = x0 + T;
    if (x1 >= N0) x1 -= N0;
#pragma unroll
    for (int j = 0; j < MSL; j++) {
      M3<TS> t;
      level(t, j, (long)x1 * slice + xs);
      Sx[j] = convert<TC>(t);
    }
    for (int n = 0; n < len; n++) {
      const long s1 = (long)x1 * slice + xs;
      M3<TS> l1, l1r;
      if (n + 1 < len) {
        load(l1, LT, L.V, 0, s1);
        load(l1r, LT, L.V, 0, fwd(L, s1, x, k + 1, R));
      }
      // traces of this step, transposed through LDS, FP64 sums
#pragma unroll
      for (int i = 0; i < MSL; i++)
#pragma unroll
        for (int j = 0; j < MSL; j++)
          rw[lane * 16 + i * MSL + j] = live ? retr_dag_mul(Sx[j], D[i]) : TC(0);
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      const int q = lane & 15, grp = lane >> 4;
      double sred = 0.0;
      if (q < MM) {
#pragma unroll
        for (int m = 0; m < 16; m++) sred += (double)rw[(grp * 16 + m) * 16 + q];
      }
      sred += __shfl_xor(sred, 16, 64);
      sred += __shfl_xor(sred, 32, 64);
      if (lane < 16) red2[(par * 4 + wave) * 16 + lane] = sred;
      __syncthreads();
      if (threadIdx.x < MM) {
        const double* r2 = red2 + par * 64;
        acc[x0 * MM + threadIdx.x] += ((r2[threadIdx.x] + r2[16 + threadIdx.x]) + r2[32 + threadIdx.x]) +
                                      r2[48 + threadIdx.x];
      }
      par ^= 1;
      if (n + 1 < len) {
        // next source slice x1: D_j = L(x1)^dag S_j(x1) L(x1+Rk)
        M3<TC> a = convert<TC>(l1), b = convert<TC>(l1r);
#pragma unroll
        for (int j = 0; j < MSL; j++) D[j] = mul(dag_mul(a, Sx[j]), b);
        x0 = x1;
        x1 = x0 + T;
        if (x1 >= N0) x1 -= N0;
#pragma unroll
        for (int j = 0; j < MSL; j++) {
          M3<TS> t;
          level(t, j, (long)x1 * slice + xs);
          Sx[j] = convert<TC>(t);
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < N0 * MM; q += 256) {
    const int xo = q / MM, ij = q % MM;
    partial[((long)xo * gridDim.x + blockIdx.x) * MM + ij] = acc[q];
  }
}

// ---- lane exchange within a row of 16 lanes (DPP) ------------------------
// partner of lane p: 15 - p (mirror), 7 - p within each half of 8 (half
// mirror), p ^ 2 and p ^ 1 (quad permutations)
constexpr int kRowMirror = 0x140, kRowHalfMirror = 0x141, kQuadXor2 = 0x4E, kQuadXor1 = 0xB1;
// rotation within a row: lane p reads lane (p - n) mod 16 (row_ror:n)
constexpr int kRowRor4 = 0x124, kRowRor8 = 0x128, kRowRor12 = 0x12C;
typedef float mfma_f32x4 __attribute__((ext_vector_type(4)));

// sink 2-vectors per level in a ring slot of the fused form (16 pairs x 9
// elements, padded by 16 dwords)
constexpr int sink_stride(size_t v2_bytes) { return 144 + (v2_bytes == 8 ? 8 : 4); }

// (every lane of these patterns has a source lane, so the moves need no
// fallback value)
template <int CTRL> __device__ __forceinline__ double row_dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ __forceinline__ float row_dpp(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ double lane_xor1(double v) { return row_dpp<kQuadXor1>(v); }
__device__ __forceinline__ float lane_xor1(float v) { return row_dpp<kQuadXor1>(v); }

// ---- contraction, two-lane chain form ------------------------------------
// As the one-lane chain form, but the msl levels of a (site, k) pair are split
// over two neighbouring lanes: lane 2p+h carries D_i and loads the sink links
// S_j of the levels h MH .. h MH + MH-1 (MH = msl/2), which halves the
// registers per lane (two waves per SIMD in FP64 instead of one).  The traces
// with the partner's sink links take them one row at a time through a DPP
// exchange of lanes 2p and 2p+1, in the operation order of retr_dag_mul, so
// every trace has the bits the one-lane form gives it.  Every link is still
// read once per (R,T).  Work units are blocks of 128 sites of one direction k
// (32 pairs per wave), the three directions of a block next to each other,
// each XCD on a contiguous range of units.  The traces of a step are summed
// over the pairs of a wave as in the one-lane chain form; with WACC each wave
// adds them into its own LDS accumulator (the waves are combined in fixed
// order at the end), so the waves of a workgroup do not wait for each other
// at every step (4.0 against 4.2 ms per (R,T) in FP64 at 48^3x96).  Requires
// msl = 2 or 4.
template <int MSL, class TS, class TC, bool WACC, bool BUILD>
__global__ __launch_bounds__(256, 2) void contract_chain2(Lat L, int R, int T, int g, int gb,
                                                          const v2<TS>* __restrict__ LT,
                                                          LevelPtrs<TS> S, BuildPtrs<TS> B,
                                                          double* __restrict__ partial) {
  constexpr int MM = MSL * MSL;
  constexpr int MH = MSL / 2;
  static_assert(MSL == 2 || MSL == 4, "two-lane chain form needs msl = 2 or 4");
  extern __shared__ double lds_c2[];
  const int N0 = L.n[0];
  // WACC: one accumulator per wave, [4][N0][MM], and no workgroup barrier per
  // step; else one per workgroup, [N0][MM], with the wave sums staged in red2
  constexpr int NACC = WACC ? 4 : 1;
  double* acc = lds_c2;                                   // [NACC][N0][MM]
  double* red2 = lds_c2 + NACC * N0 * MM;                 // [2][4][16]
  TC* redf = (TC*)(red2 + 2 * 4 * 16);                    // [4][32][16], traces in TC
  const int len = N0 / g;
  const long slice = L.st[0];
  const long nbk = (slice + 127) / 128;  // blocks of 128 sites per direction
  const long nblocks = 3 * nbk;
  const long units = nblocks * g;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane & 1, pr = lane >> 1;
  TC* rw = redf + wave * 32 * 16;
  // this lane's levels i = h MH + m (BUILD: the level links F, the transporters
  // S_{R-1} and the output S_R of those levels)
  const v2<TS>* sp[MH];
  const v2<TS>* bf[MH];
  const v2<TS>* bp[MH];
  v2<TS>* bo[MH];
#pragma unroll
  for (int m = 0; m < MH; m++) {
    sp[m] = level_ptr(S, h * MH + m);
    if constexpr (BUILD) {
      LevelPtrs<TS> f, q, o;
#pragma unroll
      for (int i = 0; i < 5; i++) {
        f.p[i] = B.f[i];
        q.p[i] = B.prev[i];
        o.p[i] = B.out[i];
      }
      bf[m] = level_ptr(f, h * MH + m);
      bp[m] = level_ptr(q, h * MH + m);
      bo[m] = const_cast<v2<TS>*>(level_ptr(o, h * MH + m));
    }
  }

  for (int q = threadIdx.x; q < NACC * N0 * MM; q += 256) acc[q] = 0.0;
  __syncthreads();
  double* accw = acc + (WACC ? wave * N0 * MM : 0);
  int par = 0;
  for (long u = xcd_order(blockIdx.x, gridDim.x); u < units; u += gridDim.x) {
    const int c = (int)(u / nblocks);
    // chain c: segment c / gb of the base chain c % gb (slices = c % gb mod gb)
    const int cs = (int)(((long)(c % gb) + (long)(c / gb) * len * T) % N0);
    const long bl = u % nblocks;
    const int k = (int)(bl % 3);
    const long xs0 = (bl / 3) * 128 + (threadIdx.x >> 1);
    const bool live = xs0 < slice;
    const long xs = live ? xs0 : 0;
    int x[4];
    coords(L, xs, x);
    // level link m of this lane at site s (BUILD: formed as S_R(x,k) =
    // F(x,k) S_{R-1}(x+k,k), the operands of space_transporter, and stored)
    auto level = [&](M3<TS>& t, int m, long s) {
      if constexpr (BUILD) {
        M3<TS> f, q;
        load(f, bf[m], L.V, k, s);
        load(q, bp[m], L.V, k, fwd(L, s, x, k + 1, 1));
        t = mul(f, q);
        if (live) store(t, bo[m], L.V, k, s);
      } else {
        load(t, sp[m], L.V, k, s);
      }
    };
    M3<TC> D[MH], Sx[MH];
    {
      const long s0 = (long)cs * slice + xs;
      M3<TS> a, b, t;
      load(a, LT, L.V, 0, s0);
      load(b, LT, L.V, 0, fwd(L, s0, x, k + 1, R));
      M3<TC> ac = convert<TC>(a), bc = convert<TC>(b);
#pragma unroll
      for (int m = 0; m < MH; m++) {
        level(t, m, s0);
        D[m] = mul(dag_mul(ac, convert<TC>(t)), bc);
      }
    }
    int x0 = cs;
    int x1 = x0 + T;
    if (x1 >= N0) x1 -= N0;
#pragma unroll
    for (int m = 0; m < MH; m++) {
      M3<TS> t;
      level(t, m, (long)x1 * slice + xs);
      Sx[m] = convert<TC>(t);
    }
    for (int n = 0; n < len; n++) {
      // traces with this lane's sink links, then with the partner's
#pragma unroll
      for (int mi = 0; mi < MH; mi++) {
        const int i = h * MH + mi;
#pragma unroll
        for (int mj = 0; mj < MH; mj++) {
          const int j = h * MH + mj;
          rw[pr * 16 + i * MSL + j] = live ? retr_dag_mul(Sx[mj], D[mi]) : TC(0);
        }
      }
#pragma unroll
      for (int mj = 0; mj < MH; mj++) {
        TC rp[MH][3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          TC pre[3], pim[3];
#pragma unroll
          for (int e = 0; e < 3; e++) {
            pre[e] = lane_xor1(Sx[mj].re[3 * r + e]);
            pim[e] = lane_xor1(Sx[mj].im[3 * r + e]);
          }
#pragma unroll
          for (int mi = 0; mi < MH; mi++) {
            TC t = pre[0] * D[mi].re[3 * r];
            t = fmadd(pim[0], D[mi].im[3 * r], t);
#pragma unroll
            for (int e = 1; e < 3; e++) {
              t = fmadd(pre[e], D[mi].re[3 * r + e], t);
              t = fmadd(pim[e], D[mi].im[3 * r + e], t);
            }
            rp[mi][r] = t;
          }
        }
        const int j = (1 - h) * MH + mj;
#pragma unroll
        for (int mi = 0; mi < MH; mi++) {
          const int i = h * MH + mi;
          rw[pr * 16 + i * MSL + j] = live ? rp[mi][0] + rp[mi][1] + rp[mi][2] : TC(0);
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      {
        const int q = lane & 15, grp = lane >> 4;
        double sred = 0.0;
        if (q < MM) {
#pragma unroll
          for (int m = 0; m < 8; m++) sred += (double)rw[(grp * 8 + m) * 16 + q];
        }
        sred += __shfl_xor(sred, 16, 64);
        sred += __shfl_xor(sred, 32, 64);
        if constexpr (WACC) {
          if (lane < MM) accw[x0 * MM + lane] += sred;
        } else {
          if (lane < 16) red2[(par * 4 + wave) * 16 + lane] = sred;
        }
      }
      if constexpr (!WACC) {
        __syncthreads();
        if (threadIdx.x < MM) {
          const double* r2 = red2 + par * 64;
          acc[x0 * MM + threadIdx.x] += ((r2[threadIdx.x] + r2[16 + threadIdx.x]) + r2[32 + threadIdx.x]) +
                                        r2[48 + threadIdx.x];
        }
        par ^= 1;
      }
      if (n + 1 < len) {
        // next source slice x1: D_i = L(x1)^dag S_i(x1) L(x1+Rk)
        const long s1 = (long)x1 * slice + xs;
        M3<TS> l1, l1r;
        load(l1, LT, L.V, 0, s1);
        load(l1r, LT, L.V, 0, fwd(L, s1, x, k + 1, R));
        M3<TC> a = convert<TC>(l1), b = convert<TC>(l1r);
#pragma unroll
        for (int m = 0; m < MH; m++) D[m] = mul(dag_mul(a, Sx[m]), b);
        x0 = x1;
        x1 = x0 + T;
        if (x1 >= N0) x1 -= N0;
#pragma unroll
        for (int m = 0; m < MH; m++) {
          M3<TS> t;
          level(t, m, (long)x1 * slice + xs);
          Sx[m] = convert<TC>(t);
        }
      }
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < N0 * MM; q += 256) {
    const int xo = q / MM, ij = q % MM;
    double v = acc[q];
    if constexpr (WACC) v = ((v + acc[N0 * MM + q]) + acc[2 * N0 * MM + q]) + acc[3 * N0 * MM + q];
    partial[((long)xo * gridDim.x + blockIdx.x) * MM + ij] = v;
  }
}

// ---- contraction, fused time-slice form ----------------------------------
// Whether a fused launch sweeps the start groups (one workgroup per slot
// taking every start group of its pair blocks in turn, the ring sliding on):
// for T from 1 with as many T as start slices per workgroup, the ring at the
// end of a start group holds all but one of the slices the next start group
// starts with, so the fill of W + 1 slices at each pair block becomes one
// slice, and each slice of a pair block is fetched from HBM once.
// (Only the FP32 form is built with the sweep; other launches keep one start
// group per workgroup.)
__host__ __device__ constexpr bool depth_sweep(bool stored, bool fp32, int Ta, int d, int W) {
  return !stored && fp32 && Ta == 1 && d == W;
}

// One launch covers T = Ta .. Ta+d-1 of one R.  A lane owns (pair p of a block
// of 16 (site,k) pairs, level i) for one start slice x0 and carries
//     D_i(T) = L_T(x)^dag S_i(x) L_T(x+Rk)
// from T = Ta (built from the stored L_Ta) to Ta+d-1 through
//     D_i(T+1) = U_0(x+T0)^dag D_i(T) U_0(x+Rk+T0),
// i.e. L_{T+1}(x) = L_T(x) U_0(x+T0), and takes Re tr[S_j(x+T0)^dag D_i(T)]
// for every level j at every T (the reference's trace, wloop.c:436-508).
// The transporters of the intermediate T are never stored.
//
// The W waves of a workgroup own the start slices x0g .. x0g+W-1 of the same
// pair block.  At step n wave w needs, from time slice x0g+Ta+n+w, the sink
// links S_j of the block and the time links at x and x+Rk; a ring of W+1
// LDS slots holds the W slices of the step; the slice the next step adds is
// stored into the free slot at the top of the step, from registers loaded
// during the step before.  Level and time links are fetched from HBM about
// (W+d-1)/(W d) times per (x0,T) instead of once, and a global load has a
// whole step to arrive before it is waited for.
//
// Lane layout: lane = 16 i + p (a row of 16 lanes per level).  The traces of a
// step are summed over the 16 pairs of a row in FP64 by a butterfly of DPP
// exchanges (partners 15-p, the half mirror, p^2, p^1) that hands half of the
// carried values to the partner in each of its first two stages, then added
// into a per-wave LDS accumulator acc[T][i][j]; one
// partial per (T, x0, workgroup slot, i, j) is stored at the end.  Workgroup
// b serves start group b / nslot and the blocks of 16 sites b % nslot +
// m nslot, each for k = 0, 1, 2 in turn (the time links or transporters at x
// of a site block are fetched from HBM for its first direction and from L2
// for the other two), so the workgroups that read the same pair blocks sit on
// the same XCD when nslot is a multiple of 8.  Deterministic.  Requires
// msl <= 4, d <= 16.
//
// SWEEP (FP32, T from 1, d = W; depth_sweep): the grid has nslot workgroups,
// and workgroup b takes every start group of its pair blocks in turn.  Ring
// offset o of start group g sits in slot (o + rot) mod (W + 1), rot = -g mod
// (W + 1): a slot keeps its slice from one start group to the next, and the
// ring at the end of a start group holds ring offsets -1 .. W - 2 of the next
// but W - 1, which the last refill fetch of the start group brings in (stored
// after its last step).  The sums of a start group are added into the
// workgroup's partial slots (written by its first pair block) after its last
// step, pair block after pair block as without the sweep.
//
// STORED (mixed path, T >= 10, or T >= 9 in form 8): the time transporters of
// the block's T are the
// stored FP32 copies of the FP64-accumulated L_T (LTa + (T-Ta) 9V), and
//     D_i(T) = L_T(x)^dag S_i(x) L_T(x+Rk)
// is formed afresh at every T from them (the arithmetic of the chain forms, so
// no time link is accumulated in FP32).  The ring then holds the sinks only;
// each wave keeps the source links S_i(x) of its start slice in registers and
// its own L_T at x and x+Rk in a two-slot LDS buffer, loaded one step ahead.
//
// MX (matrix-core traces): lane = 4 p + i instead.  The msl^2 traces of a pair
// are one 4x4 outer-product accumulation over the 18 real components,
//     C_p[i][j] = sum_e Re D_i,e Re S_j,e + Im D_i,e Im S_j,e,
// with lane 4p+i supplying D_i (A operand) and S_i (B operand, level j = i):
// v_mfma_f32_4x4x1f32 (16 blocks of 4x4, one per pair) for FP32 open loops,
// the output C_p[i][j] in accumulator register i of lane 4p+j, each trace a
// sum of 18 products in FP32 (two chains of 9, real and imaginary parts, added
// at the end).  The sum over the 16 pairs
// is in FP64: lane bits 3 and 2 (the pairs of a row of 16 lanes) by two DPP
// stages that hand half of the carried values to the partner, then lane bits 4
// and 5 (the four rows) by v_mfma_f64_4x4x4f64 with a B operand of ones, whose
// output lane 16j + 4i + c holds the sum of input lanes 16r + 4i + j over r.
// FP64 open loops use v_mfma_f64_4x4x4f64 directly: block b of lanes 16r+4b+i
// sums the products of the four pairs 4r+b (r = 0..3) in FP64, and two DPP
// stages sum the blocks.  Each lane reads one sink link per step instead of
// msl, and the trace arithmetic runs on the matrix cores beside the VALU work
// of the other waves.
//
// waves per SIMD the LDS of the form allows at W = 4 (FP64 time links) or
// W = 4 / 8 (FP32 time links), as register target
template <int MSL, int W, class TS, class TC, bool STORED, bool MX, bool SWEEP>
__global__ __launch_bounds__(64 * W, sizeof(TC) == 8 ? 2 : (STORED && W < 8 ? 3 : 4)) void contract_depth(Lat L, int R, int Ta, int d, int nslot,
                                                         const v2<TC>* __restrict__ LTa,
                                                         const v2<TC>* __restrict__ U0,
                                                         LevelPtrs<TS> S,
                                                         LevelPtrs<TS> F,
                                                         v2<TS>* __restrict__ snext,
                                                         double* __restrict__ partial,
                                                         long pstride) {
  static_assert(MSL <= 4, "fused time-slice form needs msl <= 4");
  constexpr int MM = MSL * MSL;
  constexpr int NT = 64 * W;
  constexpr int NS = MSL * 9 * 16;             // sink 2-vectors of a slice: [j][e][p]
  constexpr int NU = 2 * 9 * 16;               // time-link 2-vectors: [x / x+Rk][e][p]
  // sink 2-vectors per level in a ring slot: 144, padded so that the four
  // levels of a pair (lanes 4p+i of the matrix-core form) sit in different
  // LDS banks (level stride 48 dwords mod 64 for FP32, 16 for FP64)
  constexpr int SJ = sink_stride(sizeof(v2<TS>));
  constexpr int NSP = MSL * SJ;
  constexpr int PS = (NS + NT - 1) / NT, PU = STORED ? 0 : (NU + NT - 1) / NT;
  constexpr int PL = STORED ? (NU + 63) / 64 : 0;  // L_T 2-vectors per lane (own wave)
  constexpr int SLOT = NSP * (int)sizeof(v2<TS>) + (STORED ? 0 : NU * (int)sizeof(v2<TC>));  // bytes
  extern __shared__ double lds_d[];
  double* acc = lds_d;                                   // [W][d][16]: (i, j) at 4 i + j
  long* xr_tab = (long*)(lds_d + W * d * 16);            // [16] x+Rk of the block's pairs
  char* ring = (char*)(xr_tab + 16);                     // [W+1] slots
  v2<TC>* ltb = (v2<TC>*)(ring + (W + 1) * SLOT);        // STORED: [W][NU]
  const int N0 = L.n[0];
  const long slice = L.st[0];
  const long V = L.V;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform, in an SGPR
  const int i = MX ? (lane & 3) : (lane >> 4), p = MX ? (lane >> 2) : (lane & 15);
  // SWEEP (depth_sweep): the workgroup takes every start group of its pair
  // blocks in turn, with nslot workgroups in the grid; otherwise it takes one
  // start group, with start groups x nslot workgroups
  static_assert(!(SWEEP && STORED), "no sweep over the stored transporters");
  constexpr bool sweep = SWEEP;
  const int G = sweep ? (N0 + W - 1) / W : 1;
  const int jw = sweep ? (int)blockIdx.x : (int)(blockIdx.x % nslot);
  const int x0g0 = sweep ? 0 : (int)(blockIdx.x / nslot) * W;
  int x0g = x0g0, x0 = x0g + w;  // the start group and this wave's start slice
  bool wave_live = x0 < N0;
  const long nbk = (slice + 15) / 16;
  const long nb = 3 * nbk;

  // this thread's elements of a ring slot: sink element q = (j*9 + e)*16 + pp
  // and time-link element q = (h*9 + e)*16 + pp (h = 1: the site x + R k)
  constexpr int PU1 = PU > 0 ? PU : 1;
  int sj_[PS], sp_[PS], uh_[PU1], up_[PU1];
  long uo_[PU1];
  const v2<TS>* sb_[PS];  // level plane of the element, at the element's pair
#pragma unroll
  for (int m = 0; m < PS; m++) {
    const int q = threadIdx.x + m * NT;
    sj_[m] = q < NS ? q / 144 : -1;
    sp_[m] = q & 15;  // (ring index: q + sj_ (SJ - 144))
    sb_[m] = level_ptr(S, sj_[m] >= 0 ? sj_[m] : 0) + (long)((q >> 4) % 9) * V + sp_[m];
  }
#pragma unroll
  for (int m = 0; m < PU; m++) {
    const int q = threadIdx.x + m * NT;
    uh_[m] = q < NU ? q / 144 : -1;
    up_[m] = q & 15;
    uo_[m] = (long)((q >> 4) % 9) * V;
  }

  for (int q = threadIdx.x; q < W * d * 16; q += NT) acc[q] = 0.0;

  // pair block bb = 3 (site block) + k: 3 jw, 3 jw + 1, 3 jw + 2, then the
  // site block jw + nslot, ...
  for (long bb = 3L * jw; bb < nb; bb += (bb % 3 == 2) ? 3L * nslot - 2 : 1) {
    const int k = (int)(bb % 3);
    const long xs0 = (bb / 3) * 16;
    const long xs = xs0 + p;
    const bool pair_live = xs < slice;
    bool live = false;  // (per start group)
    int x[4];
    coords(L, pair_live ? xs : 0, x);
    const long xr = fwd(L, pair_live ? xs : 0, x, k + 1, R);  // spatial index of x + R k
    const long kV = (long)9 * k * V;
    // the start group g (x0g0 + g W) and this wave's start slice
    int rot = 0, g = 0;
    auto set_group = [&](int gg) {
      x0g = x0g0 + gg * W;
      x0 = x0g + w;
      wave_live = x0 < N0;
      live = wave_live && i < MSL && pair_live;
      rot = (W + 1 - gg % (W + 1)) % (W + 1);
    };
    // the next R's spatial transporters at this wave's start slice (snext:
    // msl 3-slot fields), S_{R+1}(x,k) = F(x,k) S_R(x+k,k) in the product and
    // operand order of space_transporter: the start slices and pair blocks of a
    // launch cover every (site, k) once, so the launch forms every S_{R+1} once
    // and no separate pass over F, S_R and S_{R+1} is needed.  (Loading the
    // operands early, to overlap the ring fill or the block's last step, made
    // the compiler spill 30-60 registers.)
    auto form_next = [&]() {
      if (snext != nullptr && live) {
        const long xsl = (long)x0 * slice;
        M3<TS> f, sr;
        load(f, level_ptr(F, i), V, k, xsl + xs);
        load(sr, level_ptr(S, i), V, k, xsl + fwd(L, xs, x, k + 1, 1));
        store(mul(f, sr), snext + (long)i * 27 * V, V, k, xsl + xs);
      }
    };
    set_group(0);
    form_next();  // (before the fill of the first start group, whose loads it then overlaps)

    // one time slice of the block (sinks S_j and the links of usrc at x and
    // x + R k) into registers / from registers into a ring slot.  Offsets and
    // validity of this thread's elements are fixed per pair block.  Every
    // thread loads every one of its element positions unconditionally (a
    // position with no element, or past the end of the slice, reads a valid
    // link: element 0 of its plane) and the store into the ring skips
    // positions with no element, so the loads sit in no branch and are waited
    // for only at the store, one step later.  Ring positions of pairs past the
    // end of the slice hold such links: those pairs carry D = 0, so they
    // contribute nothing whatever (finite) link is there.
    struct Slice {  // this thread's elements of one slice, in registers
      v2<TS> s[PS];
      v2<TC> u[PU1];
    };
    Slice nxt;  // the slice of the next refill, fetched one step ahead
    const v2<TS>* sa_[PS];  // the element's link at the pair block's first slice
    long uoff_[PU1];
    bool sok_[PS], uok_[PU1];
    auto offsets = [&]() {
#pragma unroll
      for (int m = 0; m < PS; m++) {
        sok_[m] = sj_[m] >= 0 && xs0 + sp_[m] < slice;
        sa_[m] = sb_[m] + kV + (sok_[m] ? xs0 : -(long)sp_[m]);
      }
#pragma unroll
      for (int m = 0; m < PU; m++) {
        uok_[m] = uh_[m] >= 0 && xs0 + up_[m] < slice;
        uoff_[m] = uo_[m] + (uok_[m] ? (uh_[m] > 0 ? xr_tab[up_[m]] : xs0 + up_[m]) : 0);
      }
    };
    auto fetch_at = [&](Slice& f, long ys, const v2<TC>* __restrict__ usrc) {
#pragma unroll
      for (int m = 0; m < PS; m++) f.s[m] = sa_[m][ys];
#pragma unroll
      for (int m = 0; m < PU; m++) f.u[m] = usrc[uoff_[m] + ys];
    };
    // ring offset o: time slice x0g+Ta+o
    auto fetch = [&](Slice& f, int o) { fetch_at(f, (long)((x0g + Ta + o) % N0) * slice, U0); };
    auto put = [&](int slot, const Slice& f) {
      char* dst = ring + slot * SLOT;
#pragma unroll
      for (int m = 0; m < PS; m++) {
        v2<TS> v;  // (component-wise: a copy of the whole element kept the slice in scratch)
        v.x = f.s[m].x;
        v.y = f.s[m].y;
        if (sj_[m] >= 0) ((v2<TS>*)dst)[threadIdx.x + m * NT + sj_[m] * (SJ - 144)] = v;
      }
#pragma unroll
      for (int m = 0; m < PU; m++) {
        v2<TC> v;
        v.x = f.u[m].x;
        v.y = f.u[m].y;
        if (uh_[m] >= 0) ((v2<TC>*)(dst + NSP * (int)sizeof(v2<TS>)))[threadIdx.x + m * NT] = v;
      }
    };
    // the fills of a pair block's first slices: slice j < cnt (time slice
    // ys_of(j), time links from usrc) into ring slot slot_of(j).  The loads of
    // FILL slices are all issued before the first of their stores: a slice's
    // stores wait for its loads (and, behind the conditional stores, for every
    // load before them), so filled one slice after another the slices' HBM
    // round trips would follow one another.  FILL: as many slices as the
    // register budget of the occupancy target leaves room for (FP32, W = 8:
    // 5 slices of 6 registers)
    constexpr int SLICE_REGS = (PS * (int)sizeof(v2<TS>) + PU * (int)sizeof(v2<TC>)) / 4;
    constexpr int FILL_REGS = sizeof(TC) == 8 ? 48 : (W >= 8 ? 30 : 20);
    constexpr int FILL = FILL_REGS / SLICE_REGS > 1 ? FILL_REGS / SLICE_REGS : 1;
    auto fill = [&](int cnt, auto ys_of, auto slot_of, const v2<TC>* __restrict__ usrc) {
      for (int j0 = 0; j0 < cnt; j0 += FILL) {
        Slice f[FILL];
#pragma unroll
        for (int j = 0; j < FILL; j++)
          if (j0 + j < cnt) fetch_at(f[j], ys_of(j0 + j), usrc);
#pragma unroll
        for (int j = 0; j < FILL; j++)
          if (j0 + j < cnt) put(slot_of(j0 + j), f[j]);
      }
    };
    auto slice_at = [&](int o) { return (long)((x0g + Ta + o) % N0) * slice; };  // ring offset o

    if (w == 0 && i == 0) xr_tab[p] = xr;
    __syncthreads();  // xr_tab; the previous block's last reads of the ring
    offsets();
    // STORED: this wave's L_T elements q = lane + 64 m = (h*9 + e)*16 + pp at
    // its start slice, loaded unconditionally (element 0 of the plane where
    // there is none: only pairs past the end of the slice, whose D is 0, read
    // those)
    v2<TC> pl[PL > 0 ? PL : 1];
    unsigned lo_[PL > 0 ? PL : 1];  // element offsets within one L_T (< 9V < 2^32)
    bool lok_[PL > 0 ? PL : 1];
    M3<TC> Msrc;
    auto lt_fetch = [&](int n) {  // L_{Ta+n} of this wave's pairs into registers
#pragma unroll
      for (int m = 0; m < PL; m++) pl[m] = (LTa + (long)n * 9 * V)[lo_[m]];
    };
    // from registers into this wave's buffer (read only by this wave, at the
    // top of the next step: LDS accesses of a wave take effect in program order)
    auto lt_put = [&]() {
#pragma unroll
      for (int m = 0; m < PL; m++) {
        v2<TC> v;
        v.x = pl[m].x;
        v.y = pl[m].y;
        if (lane + 64 * m < NU) ltb[w * NU + lane + 64 * m] = v;
      }
    };
    M3<TC> D;
    // D_i(T+1) = U_0(x+T0)^dag D_i(T) U_0(x+Rk+T0) from the ring slot's time
    // links (every lane: D = 0 stays 0)
    auto update = [&](const v2<TC>* ru) {
      M3<TC> u;
#pragma unroll
      for (int e = 0; e < 9; e++) {
        v2<TC> c = ru[e * 16];
        u.re[e] = c.x;
        u.im[e] = c.y;
      }
      M3<TC> h = dag_mul(u, D);
#pragma unroll
      for (int e = 0; e < 9; e++) {
        v2<TC> c = ru[(9 + e) * 16];
        u.re[e] = c.x;
        u.im[e] = c.y;
      }
      D = mul(h, u);
    };
    // ring slot of ring offset o (relative to the start group): (o + rot) mod
    // (W + 1); a sweep shifts it by one slot per start group, so that a slot
    // keeps its slice from one start group to the next
    auto step = [&](int n, auto more_t) {
      constexpr bool more = decltype(more_t)::value;
      if constexpr (more) {
        put((n + W + rot) % (W + 1), nxt);
        // (sweep: at n = d - 2 the slice the next start group lacks; where no
        // slice is needed the slice before is fetched again, from L2: in the
        // sweep a fetch under a branch made the compiler wait for it within
        // the step)
        if constexpr (SWEEP)
          fetch(nxt, (n + 2 < d || g + 1 < G) ? n + W + 1 : n + W);
        else if (n + 2 < d)
          fetch(nxt, n + W + 1);
      }
      const char* slot = ring + ((n + w + rot) % (W + 1)) * SLOT;
      const v2<TS>* rs = (const v2<TS>*)slot + p;
      const v2<TC>* ru = (const v2<TC>*)(slot + NSP * (int)sizeof(v2<TS>)) + p;
      if constexpr (STORED) {
        // D_i(T) = L_T(x)^dag S_i(x) L_T(x + R k) from this wave's buffer
        const v2<TC>* lb = ltb + w * NU + p;
        M3<TC> a;
#pragma unroll
        for (int e = 0; e < 9; e++) {
          v2<TC> c = lb[e * 16];
          a.re[e] = c.x;
          a.im[e] = c.y;
        }
        const M3<TC> h = dag_mul(a, Msrc);
#pragma unroll
        for (int e = 0; e < 9; e++) {
          v2<TC> c = lb[(9 + e) * 16];
          a.re[e] = c.x;
          a.im[e] = c.y;
        }
        D = mul(h, a);
        // (the traces' sink reads stay below: hoisted above the formation of
        // D they hold 18 more registers there)
        __builtin_amdgcn_sched_barrier(0);
      }
      double* accn = acc + (w * d + n) * 16;
      if constexpr (MX) {
        // traces on the matrix cores; lane 4p+i holds the sink of level i
        // (levels i >= msl read level 0: their outputs are not stored)
        const v2<TS>* rj = rs + (i < MSL ? i : 0) * SJ;
        double v;
        int vij;
        if constexpr (sizeof(TC) == 4) {
          // two independent accumulations (real and imaginary parts) halve
          // the chain of dependent outer products (four chains measured the
          // same as two)
          mfma_f32x4 c = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int e = 0; e < 9; e++) {
            const v2<TS> sv = rj[e * 16];
            c = __builtin_amdgcn_mfma_f32_4x4x1f32(D.re[e], (float)sv.x, c, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_4x4x1f32(D.im[e], (float)sv.y, ci, 0, 0, 0);
          }
          c += ci;
          if constexpr (more && !STORED) update(ru);
          // lane 4p+j: c[i] = trace (i, j) of pair p.  Lane bit 3: keep the
          // traces i = 2 b3 + {0,1}, hand the other two to lane ^ 8
          const bool b3 = (lane & 8) != 0, b2 = (lane & 4) != 0;
          const double t0 = c[0], t1 = c[1], t2 = c[2], t3 = c[3];
          double k0 = b3 ? t2 : t0, k1 = b3 ? t3 : t1;
          const double s0 = b3 ? t0 : t2, s1 = b3 ? t1 : t3;
          k0 += row_dpp<kRowRor8>(s0);
          k1 += row_dpp<kRowRor8>(s1);
          // lane bit 2: keep i = 2 b3 + b2, the partner is lane + 4 (bit 2
          // clear) or lane - 4 (bit 2 set)
          double kk = b2 ? k1 : k0;
          const double ss = b2 ? k0 : k1;
          const double up = row_dpp<kRowRor12>(ss), dn = row_dpp<kRowRor4>(ss);
          kk += b2 ? dn : up;
          // rows: output lane 16j + 4i + c = sum over r of lane 16r + 4i + j
          v = __builtin_amdgcn_mfma_f64_4x4x4f64(kk, 1.0, 0.0, 0, 0, 0);
          vij = ((lane >> 2) & 3) * 4 + (lane >> 4);
        } else {
          double c = 0.0, ci = 0.0;
#pragma unroll
          for (int e = 0; e < 9; e++) {
            const v2<TS> sv = rj[e * 16];
            c = __builtin_amdgcn_mfma_f64_4x4x4f64((double)D.re[e], (double)sv.x, c, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f64_4x4x4f64((double)D.im[e], (double)sv.y, ci, 0, 0, 0);
          }
          c += ci;
          if constexpr (more && !STORED) update(ru);
          // lane 16i + 4b + j: trace (i, j) summed over the pairs 4r + b;
          // sum over b (lane bits 3 and 2)
          c += row_dpp<kRowRor8>(c);
          c += row_dpp<kRowRor4>(c);
          v = c;
          vij = (lane >> 4) * 4 + (lane & 3);
        }
        // every lane of an entry holds the same sum: the read-modify-write of
        // the lanes of one entry stores the same value
        accn[vij] += v;
      } else {
      // traces against the sinks S_j(x + (Ta+n) 0): products in the compute
      // type, each trace converted to FP64 before it is summed
      double tr[4];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        tr[j] = 0.0;
        if (j < MSL) {
          M3<TS> sj;
#pragma unroll
          for (int e = 0; e < 9; e++) {
            v2<TS> v = rs[j * SJ + e * 16];
            sj.re[e] = v.x;
            sj.im[e] = v.y;
          }
          tr[j] = (double)retr_dag_mul_chain(convert<TC>(sj), D);
        }
      }
      // sum over the 16 pairs of the row in FP64; lane p ends with the sum
      // for j = p >> 2
      const bool b3 = (p & 8) != 0, b2 = (p & 4) != 0;
      double s0 = b3 ? tr[0] : tr[2], s1 = b3 ? tr[1] : tr[3];
      double k0 = b3 ? tr[2] : tr[0], k1 = b3 ? tr[3] : tr[1];
      k0 += row_dpp<kRowMirror>(s0);
      k1 += row_dpp<kRowMirror>(s1);
      const double sh = b2 ? k0 : k1;
      double v = b2 ? k1 : k0;
      v += row_dpp<kRowHalfMirror>(sh);
      v += row_dpp<kQuadXor2>(v);
      v += row_dpp<kQuadXor1>(v);
      if (wave_live && i < MSL && (p & 3) == 0 && (p >> 2) < MSL) accn[i * 4 + (p >> 2)] += v;
      if constexpr (more && !STORED) update(ru);
      }
      if constexpr (STORED && more) {
        lt_put();
        if (n + 2 < d) lt_fetch(n + 2);
      }
      if constexpr (more) __syncthreads();
    };
    if (!STORED && Ta == 1) {
      fill(W + 1, [&](int j) { return slice_at(j - 1); }, [](int j) { return (j + W) % (W + 1); }, U0);
    }
    for (g = 0; g < (SWEEP ? G : 1); g++) {
      if (g > 0) {
        set_group(g);
        form_next();
      }

      if constexpr (STORED) {
        const long xslice = (long)(wave_live ? x0 : 0) * slice;
  #pragma unroll
        for (int m = 0; m < PL; m++) {
          const int q = lane + 64 * m;
          const int hh = q / 144, e = (q >> 4) % 9, pp = q & 15;
          lok_[m] = q < NU && xs0 + pp < slice;
          lo_[m] = lok_[m] ? (unsigned)((long)e * V + xslice + (hh ? xr_tab[pp] : xs0 + pp)) : 0u;
        }
        set_zero(Msrc);
        if (live) {
          M3<TS> t;
          load(t, level_ptr(S, i), V, k, xslice + xs);
          Msrc = convert<TC>(t);
        }
      }
      set_zero(D);
      if constexpr (STORED) {
        // sinks of the slices x0g+Ta+o in ring slots o = 0 .. W-1; L_Ta into
        // the buffer, L_{Ta+1} into registers
        fill(W, slice_at, [](int j) { return j; }, U0);
        lt_fetch(0);
        lt_put();
        if (d > 1) lt_fetch(1);
        __syncthreads();
      } else {
      // the start slices x0g+w with S_i(x) and L_Ta at x and x + R k: for Ta = 1
      // (L_1 = U_0) ring offsets -1 .. W-2, filled with offsets -1 .. W-1;
      // otherwise slots 0 .. W-1, refilled with offsets 0 .. W-1 afterwards
      // (Ta = 1: filled before the first start group; in a sweep the ring
      // then holds the slices of every later one, the last of them stored at
      // the end of the start group before)
      if (Ta != 1) {
        fill(W, [&](int j) { return (long)((x0g + j) % N0) * slice; }, [](int j) { return j; }, LTa);
      }
      __syncthreads();
      // D_i(Ta) = L_Ta(x)^dag S_i(x) L_Ta(x + R k)
      if (live) {
        const char* slot = ring + (Ta == 1 ? (w + W + rot) % (W + 1) : w) * SLOT;
        const v2<TS>* rs = (const v2<TS>*)slot + p;
        const v2<TC>* ru = (const v2<TC>*)(slot + NSP * (int)sizeof(v2<TS>)) + p;
        M3<TC> a;
        M3<TS> t;
  #pragma unroll
        for (int e = 0; e < 9; e++) {
          v2<TS> v = rs[i * SJ + e * 16];
          t.re[e] = v.x;
          t.im[e] = v.y;
          v2<TC> c = ru[e * 16];
          a.re[e] = c.x;
          a.im[e] = c.y;
        }
        const M3<TC> h = dag_mul(a, convert<TC>(t));
  #pragma unroll
        for (int e = 0; e < 9; e++) {
          v2<TC> c = ru[(9 + e) * 16];
          a.re[e] = c.x;
          a.im[e] = c.y;
        }
        D = mul(h, a);
      }
      __syncthreads();  // ring slots of the start slices are refilled next
      if (Ta != 1) {
        fill(W, slice_at, [](int j) { return j; }, U0);
        __syncthreads();
      }
      }  // not STORED

      // the slice of ring offset n+W (read by wave W-1 at step n+1) is fetched
      // into registers during step n-1 and stored at the top of step n into the
      // slot wave 0 read at step n-1, so a global load has a whole step to land.
      // The last step is peeled (no refill, no update of D), so that the body of
      // every step is one branch-free block: the traces and the update of D can
      // be interleaved by the compiler (the matrix-core traces then run beside
      // the products of the update).  Lanes that are not live carry D = 0 and add
      // zeros; D stays finite because every ring position holds a link or zero.
      if (d > 1) fetch(nxt, W);
      for (int n = 0; n + 1 < d; n++) step(n, std::true_type());
      step(d - 1, std::false_type());
      __syncthreads();  // the ring of this start group is read for the last time; acc is complete
      // (sweep) the slice the next start group lacks (ring offset 2W - 1 of
      // this one, fetched at step d - 2) into the slot of offset W - 2, read
      // for the last time at step W - 2
      if (sweep && g + 1 < G) put((2 * W - 1 + rot) % (W + 1), nxt);
      if (sweep) {
        // this start group's sums into the workgroup's partial slots (the
        // first pair block of the workgroup writes them), in pair-block order
        // as without the sweep; each thread clears the entries it took
        const bool first = bb == 3L * jw;
        for (int q = threadIdx.x; q < W * d * MM; q += NT) {
          const int ww = q / (d * MM), n = (q / MM) % d, ij = q % MM;
          const int xo = x0g + ww;
          const int a = (ww * d + n) * 16 + (ij / MSL) * 4 + ij % MSL;
          if (xo < N0) {
            double* dst = partial + (long)n * pstride + ((long)xo * nslot + jw) * MM + ij;
            *dst = first ? acc[a] : *dst + acc[a];
          }
          acc[a] = 0.0;
        }
      }
    }  // start groups
  }  // pair blocks
  if (sweep) {
    // (sweep) a workgroup with no pair block: zero sums in its partial slots
    if (3L * jw >= nb)
      for (int q = threadIdx.x; q < N0 * d * MM; q += NT) {
        const int xo = q / (d * MM), n = (q / MM) % d, ij = q % MM;
        partial[(long)n * pstride + ((long)xo * nslot + jw) * MM + ij] = 0.0;
      }
  } else {
    __syncthreads();
    for (int q = threadIdx.x; q < W * d * MM; q += NT) {
      const int ww = q / (d * MM), n = (q / MM) % d, ij = q % MM;
      const int xo = x0g + ww;
      if (xo < N0)
        partial[(long)n * pstride + ((long)xo * nslot + jw) * MM + ij] =
            acc[(ww * d + n) * 16 + (ij / MSL) * 4 + ij % MSL];
    }
  }
}


// W_ij(R,T;x0) of one R from the partials of its (R,T) launches, stored in
// wl_idx order.  partial layout: [t][x0][slot][i*msl+j] (nwg slots per x0).
// One wave per (t, x0): lane l takes entry ij = l % mm of slot group
// g = l / mm (G groups, the largest power of two with G mm <= 64) and sums the
// slots g, g+G, ... with
// Neumaier compensation (consecutive lanes read consecutive doubles); the G
// compensated sums of an entry are then combined in a fixed butterfly order and
// normalised.  Deterministic: the order depends only on nwg and msl.
__device__ __forceinline__ void neumaier_add(double& s, double& c, double v) {
  const double t = s + v;
  c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
  s = t;
}

__global__ __launch_bounds__(256) void finish_wloop(int N0, int msl, int mwlt, int mwlr, int r,
                                                    int nwg, const double* __restrict__ partial,
                                                    double norm, double* __restrict__ wl) {
  const int mm = msl * msl;
  int G = 1;  // slot groups: the largest power of two with G * mm <= 64
  while (2 * G * mm <= 64) G *= 2;
  const int lane = threadIdx.x & 63;
  const long tx = (long)blockIdx.x * 4 + (threadIdx.x >> 6);  // (t, x0) = t * N0 + x0
  if (tx >= (long)mwlt * N0) return;
  const int g = lane / mm, ij = lane % mm;
  double sum = 0.0, comp = 0.0;
  if (g < G) {
    const double* p = partial + tx * nwg * mm + ij;
    for (int w = g; w < nwg; w += G) neumaier_add(sum, comp, p[(long)w * mm]);
  }
  // combine the groups of an entry: partner lane + mm*h for h = G/2, G/4, ...
  // (G is 2, 4, 4, 16, 64 for msl = 5..1)
  for (int h = G / 2; h >= 1; h /= 2) {
    const double os = __shfl_down(sum, h * mm, 64);
    const double oc = __shfl_down(comp, h * mm, 64);
    if (g < h) {
      neumaier_add(sum, comp, os);
      comp += oc;
    }
  }
  if (g == 0) {
    const int t = (int)(tx / N0), x0 = (int)(tx % N0);
    const int i = ij / msl, j = ij % msl;
    // wl_idx order: i + msl*(j + msl*(t + mwlt*(r + mwlr*x0)))
    wl[i + (long)msl * (j + (long)msl * (t + (long)mwlt * (r + (long)mwlr * x0)))] = (sum + comp) / norm;
  }
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

int contract_slots(const Lat& L, int kernel, int wg) {
  if (kernel >= 1) return wg;
  long maxc = (L.st[0] + 255) / 256;
  return 3 * (int)(wg < maxc ? wg : maxc);
}

template <int MSL, int MSLP, class TS, class TC>
static void chain_launch(const Lat& L, int R, int T, int g, const v2<TS>* LT,
                         const LevelPtrs<TS>& sp, double* partial, int nwg, hipStream_t st) {
  constexpr int MM = MSL * MSL;
  size_t lds4 = (size_t)4 * L.n[0] * MM * sizeof(double);
  if (lds4 <= 65536) {
    hipLaunchKernelGGL((contract_chain<MSL, MSLP, 4, TS, TC>), dim3(nwg), dim3(256), lds4, st, L, R,
                       T, g, LT, sp, partial);
  } else {
    size_t lds1 = (size_t)L.n[0] * MM * sizeof(double);
    hipLaunchKernelGGL((contract_chain<MSL, MSLP, 1, TS, TC>), dim3(nwg), dim3(64), lds1, st, L, R,
                       T, g, LT, sp, partial);
  }
}

template <int MSL, class TS, class TC>
static void stream_launch(const Lat& L, int R, int T, int g, const v2<TS>* LT,
                          const LevelPtrs<TS>& sp, double* partial, int nchunk, hipStream_t st) {
  long maxc = (L.st[0] + 255) / 256;
  if (nchunk > maxc) nchunk = (int)maxc;
  hipLaunchKernelGGL((contract_stream<MSL, TS, TC>), dim3(nchunk, 3, L.n[0]), dim3(256), 0, st, L,
                     R, T, g, LT, sp, partial);
}

// Chains of a per-(R,T) chain launch: the gb = gcd(T, N0) chains x0 -> x0+T
// -> ... of N0/gb slices each, split into the number of equal segments (each
// restarted from its own first slice, which costs one more slice of loads)
// that keeps the persistent workgroups busiest: work units (blocks of one
// segment) are taken in rounds of nwg, so too few units leave workgroups idle
// and a last partial round leaves most of them idle at the end.
static int chain_count(int N0, int gb, long blocks, int nwg) {
  const int lb = N0 / gb;
  int best = 1;
  double beff = 0.0;
  for (int s = 1; s <= lb; s++) {
    if (lb % s) continue;
    const int len = lb / s;
    if (len < 2 && s > 1) break;
    const long u = blocks * gb * s;
    const long rounds = (u + nwg - 1) / nwg;
    const double eff = (double)u / ((double)rounds * nwg) * (double)len / (len + 1);
    if (eff > beff * 1.01) {  // a split must gain at least 1%
      beff = eff;
      best = s;
    }
  }
  return gb * best;
}

template <int MSL, class TS, class TC>
static void chain1_launch(const Lat& L, int R, int T, int gb, const v2<TS>* LT,
                          const LevelPtrs<TS>& sp, double* partial, int nwg, hipStream_t st,
                          const BuildPtrs<TS>* build = nullptr) {
  size_t lds = (size_t)L.n[0] * MSL * MSL * sizeof(double) + 2 * 4 * 16 * sizeof(double) +
               4 * 64 * 16 * sizeof(TC);
  const int g = chain_count(L.n[0], gb, 3 * ((L.st[0] + 255) / 256), nwg);
  if (build)
    hipLaunchKernelGGL((contract_chain1<MSL, TS, TC, true>), dim3(nwg), dim3(256), lds, st, L, R, T, g,
                       gb, LT, sp, *build, partial);
  else
    hipLaunchKernelGGL((contract_chain1<MSL, TS, TC, false>), dim3(nwg), dim3(256), lds, st, L, R, T, g,
                       gb, LT, sp, BuildPtrs<TS>{}, partial);
}

template <int MSL, class TS, class TC>
static void chain2_launch(const Lat& L, int R, int T, int gb, const v2<TS>* LT,
                          const LevelPtrs<TS>& sp, double* partial, int nwg, hipStream_t st,
                          const BuildPtrs<TS>* build = nullptr) {
  // per-wave accumulators where 4 of them fit into LDS (N0 <= 300 at msl 4)
  const size_t rest = 2 * 4 * 16 * sizeof(double) + 4 * 32 * 16 * sizeof(TC);
  const size_t acc1 = (size_t)L.n[0] * MSL * MSL * sizeof(double);
  const BuildPtrs<TS> b = build ? *build : BuildPtrs<TS>{};
  const int g = chain_count(L.n[0], gb, 3 * ((L.st[0] + 127) / 128), nwg);
  if (4 * acc1 + rest <= 163840) {
    if (build)
      hipLaunchKernelGGL((contract_chain2<MSL, TS, TC, true, true>), dim3(nwg), dim3(256), 4 * acc1 + rest,
                         st, L, R, T, g, gb, LT, sp, b, partial);
    else
      hipLaunchKernelGGL((contract_chain2<MSL, TS, TC, true, false>), dim3(nwg), dim3(256), 4 * acc1 + rest,
                         st, L, R, T, g, gb, LT, sp, b, partial);
  } else {
    if (build)
      hipLaunchKernelGGL((contract_chain2<MSL, TS, TC, false, true>), dim3(nwg), dim3(256), acc1 + rest, st,
                         L, R, T, g, gb, LT, sp, b, partial);
    else
      hipLaunchKernelGGL((contract_chain2<MSL, TS, TC, false, false>), dim3(nwg), dim3(256), acc1 + rest, st,
                         L, R, T, g, gb, LT, sp, b, partial);
  }
}

template <class TS, class TC>
void launch_contract_build(const Lat& L, int msl, int R, int T, const v2<TS>* LT,
                           const v2<TS>* const* F, const v2<TS>* const* prev, v2<TS>* out,
                           double* partial, int wg, int two_lane, hipStream_t st) {
  LevelPtrs<TS> sp;
  BuildPtrs<TS> b;
  for (int i = 0; i < 5; i++) {
    sp.p[i] = i < msl ? out + (size_t)i * 27 * L.V : nullptr;
    b.f[i] = i < msl ? F[i] : nullptr;
    b.prev[i] = i < msl ? prev[i] : nullptr;
    b.out[i] = i < msl ? out + (size_t)i * 27 * L.V : nullptr;
  }
  int g = L.n[0], t = T % L.n[0];
  while (t) {  // g = gcd(T, N0)
    int r = g % t;
    g = t;
    t = r;
  }
  if (two_lane && msl == 2) {
    chain2_launch<2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b);
    return;
  }
  if (two_lane && msl == 4) {
    chain2_launch<4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b);
    return;
  }
  switch (msl) {
    case 1: chain1_launch<1, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b); break;
    case 2: chain1_launch<2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b); break;
    case 3: chain1_launch<3, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b); break;
    default: chain1_launch<4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st, &b); break;
  }
}

size_t contract_depth_lds(int msl, int W, int ts_bytes, int tc_bytes, int stored, int d) {
  const size_t nu = (size_t)2 * 9 * 16 * 2 * tc_bytes;  // time links (or L_T) of a slice
  const size_t slot = (size_t)msl * sink_stride(2 * ts_bytes) * 2 * ts_bytes + (stored ? 0 : nu);
  return (size_t)W * d * 16 * sizeof(double) + 16 * sizeof(long) + (size_t)(W + 1) * slot +
         (stored ? (size_t)W * nu : 0);
}

// the pointers of a fused launch besides its time links
template <class TS>
struct DepthArgs {
  LevelPtrs<TS> sp, fp;  // S_R and the operator levels F
  v2<TS>* snext;         // S_{R+1} formed by the launch, or null
  double* partial;
  long pstride;
};

template <int MSL, int W, class TS, class TC, bool STORED, bool MX>
static void depth_launch(const Lat& L, int R, int Ta, int d, int nslot, const v2<TC>* LTa,
                         const v2<TC>* U0, const DepthArgs<TS>& a, hipStream_t st) {
  const size_t lds = contract_depth_lds(MSL, W, (int)sizeof(TS), (int)sizeof(TC), STORED, d);
  if constexpr (!STORED && sizeof(TC) == 4) {
    if (depth_sweep(STORED, true, Ta, d, W)) {
      hipLaunchKernelGGL((contract_depth<MSL, W, TS, TC, STORED, MX, true>), dim3((unsigned)nslot), dim3(64 * W),
                         lds, st, L, R, Ta, d, nslot, LTa, U0, a.sp, a.fp, a.snext, a.partial, a.pstride);
      return;
    }
  }
  const dim3 grid((unsigned)(((L.n[0] + W - 1) / W) * nslot));
  hipLaunchKernelGGL((contract_depth<MSL, W, TS, TC, STORED, MX, false>), grid, dim3(64 * W), lds, st, L,
                     R, Ta, d, nslot, LTa, U0, a.sp, a.fp, a.snext, a.partial, a.pstride);
}

template <int MSL, class TS, class TC, bool STORED, bool MX>
static void depth_launch_w(const Lat& L, int R, int Ta, int d, int W, int nslot, const v2<TC>* LTa,
                           const v2<TC>* U0, const DepthArgs<TS>& a, hipStream_t st) {
  if (W >= 8)
    depth_launch<MSL, 8, TS, TC, STORED, MX>(L, R, Ta, d, nslot, LTa, U0, a, st);
  else
    depth_launch<MSL, 4, TS, TC, STORED, MX>(L, R, Ta, d, nslot, LTa, U0, a, st);
}

template <class TS, class TC, bool STORED, bool MX>
static void depth_launch_msl(const Lat& L, int msl, int R, int Ta, int d, int W, int nslot,
                             const v2<TC>* LTa, const v2<TC>* U0, const DepthArgs<TS>& a,
                             hipStream_t st) {
  switch (msl) {
    case 1: depth_launch_w<1, TS, TC, STORED, MX>(L, R, Ta, d, W, nslot, LTa, U0, a, st); break;
    case 2: depth_launch_w<2, TS, TC, STORED, MX>(L, R, Ta, d, W, nslot, LTa, U0, a, st); break;
    case 3: depth_launch_w<3, TS, TC, STORED, MX>(L, R, Ta, d, W, nslot, LTa, U0, a, st); break;
    default: depth_launch_w<4, TS, TC, STORED, MX>(L, R, Ta, d, W, nslot, LTa, U0, a, st); break;
  }
}

template <class TS>
static DepthArgs<TS> depth_args(int msl, const v2<TS>* const* S, const v2<TS>* const* F,
                                v2<TS>* snext, double* partial, long pstride) {
  DepthArgs<TS> a;
  for (int i = 0; i < 5; i++) {
    a.sp.p[i] = i < msl ? S[i] : nullptr;
    a.fp.p[i] = (snext && i < msl) ? F[i] : nullptr;
  }
  a.snext = snext;
  a.partial = partial;
  a.pstride = pstride;
  return a;
}

template <class TS, class TC>
void launch_contract_depth(const Lat& L, int msl, int R, int Ta, int d, int W, int nslot,
                           const v2<TC>* LTa, const v2<TC>* U0, const v2<TS>* const* S,
                           const v2<TS>* const* F, v2<TS>* snext, double* partial, long pstride,
                           bool mx, hipStream_t st) {
  const DepthArgs<TS> a = depth_args<TS>(msl, S, F, snext, partial, pstride);
  if (mx)
    depth_launch_msl<TS, TC, false, true>(L, msl, R, Ta, d, W, nslot, LTa, U0, a, st);
  else
    depth_launch_msl<TS, TC, false, false>(L, msl, R, Ta, d, W, nslot, LTa, U0, a, st);
}

void launch_contract_stored(const Lat& L, int msl, int R, int Ta, int d, int W, int nslot,
                            const float2* LTa, const float2* const* S, const float2* const* F,
                            float2* snext, double* partial, long pstride, bool mx, hipStream_t st) {
  const DepthArgs<float> a = depth_args<float>(msl, S, F, snext, partial, pstride);
  if (mx)
    depth_launch_msl<float, float, true, true>(L, msl, R, Ta, d, W, nslot, LTa, nullptr, a, st);
  else
    depth_launch_msl<float, float, true, false>(L, msl, R, Ta, d, W, nslot, LTa, nullptr, a, st);
}

// kernel 0: streaming form with wg chunks per (x0,k); 1: chain form (msl
// lanes per pair) with wg persistent workgroups; 2: one-lane chain form with
// wg persistent workgroups (msl <= 4; otherwise the streaming form); 4:
// two-lane chain form with wg persistent workgroups (msl = 2 or 4; the
// one-lane form for msl = 1 or 3, the streaming form for msl > 4).
// partial receives contract_slots() sums per x0.
template <class TS, class TC>
void launch_contract(const Lat& L, int msl, int R, int T, const v2<TS>* LT,
                     const v2<TS>* const* S, double* partial, int kernel, int wg, hipStream_t st) {
  LevelPtrs<TS> sp;
  for (int i = 0; i < 5; i++) sp.p[i] = i < msl ? S[i] : nullptr;
  int g = L.n[0], t = T % L.n[0];
  while (t) {  // g = gcd(T, N0)
    int r = g % t;
    g = t;
    t = r;
  }
  if (kernel == 4 && (msl == 2 || msl == 4)) {
    if (msl == 2) chain2_launch<2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st);
    else chain2_launch<4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st);
  } else if ((kernel == 2 || kernel == 4) && msl <= 4) {
    switch (msl) {
      case 1: chain1_launch<1, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 2: chain1_launch<2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 3: chain1_launch<3, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      default: chain1_launch<4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
    }
  } else if (kernel == 1) {
    switch (msl) {
      case 1: chain_launch<1, 1, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 2: chain_launch<2, 2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 3: chain_launch<3, 4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 4: chain_launch<4, 4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      default: chain_launch<5, 8, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
    }
  } else {
    switch (msl) {
      case 1: stream_launch<1, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 2: stream_launch<2, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 3: stream_launch<3, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      case 4: stream_launch<4, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
      default: stream_launch<5, TS, TC>(L, R, T, g, LT, sp, partial, wg, st); break;
    }
  }
}

void launch_finish_wloop(int N0, int msl, int mwlt, int mwlr, int r, int nwg, const double* partial,
                         double norm, double* wl, hipStream_t st) {
  const long waves = (long)mwlt * N0;  // one per (t, x0), four per workgroup
  hipLaunchKernelGGL(finish_wloop, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, st, N0, msl, mwlt,
                     mwlr, r, nwg, partial, norm, wl);
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
                                              const double2* const*, double*, int, int,
                                              hipStream_t);
template void launch_contract<float, float>(const Lat&, int, int, int, const float2*,
                                            const float2* const*, double*, int, int, hipStream_t);
template void launch_contract_build<float, float>(const Lat&, int, int, int, const float2*,
                                                  const float2* const*, const float2* const*, float2*,
                                                  double*, int, int, hipStream_t);
template void launch_contract_build<double, double>(const Lat&, int, int, int, const double2*,
                                                    const double2* const*, const double2* const*,
                                                    double2*, double*, int, int, hipStream_t);
template void launch_contract_depth<double, double>(const Lat&, int, int, int, int, int, int,
                                                    const double2*, const double2*,
                                                    const double2* const*, const double2* const*,
                                                    double2*, double*, long, bool, hipStream_t);
template void launch_contract_depth<float, float>(const Lat&, int, int, int, int, int, int,
                                                  const float2*, const float2*,
                                                  const float2* const*, const float2* const*,
                                                  float2*, double*, long, bool, hipStream_t);
template void launch_contract_depth<float, double>(const Lat&, int, int, int, int, int, int,
                                                   const double2*, const double2*,
                                                   const float2* const*, const float2* const*,
                                                   float2*, double*, long, bool, hipStream_t);

}  // namespace wl
