// su3.hpp -- 3x3 complex algebra and lattice indexing for the gfx950 kernels.
//
// Link fields live in HBM as structure-of-arrays "planes": a field with nl
// links per site is an array of 2-vectors (re,im) indexed
//     f[(l*9 + e)*V + s]
// with l the link slot, e = 3*row + col the matrix element and s the
// lexicographic site (x3 fastest).  A wavefront reading element e of 64
// consecutive sites touches one contiguous 1 KiB (FP64) or 512 B (FP32) run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wl {

template <class R> struct V2;
template <> struct V2<double> { using type = double2; };
template <> struct V2<float> { using type = float2; };
template <class R> using v2 = typename V2<R>::type;

// A 3x3 complex matrix held in registers, element e = 3*row + col.
template <class R> struct M3 {
  R re[9], im[9];
};

template <class R> __device__ __forceinline__ void set_zero(M3<R>& m) {
#pragma unroll
  for (int e = 0; e < 9; e++) { m.re[e] = R(0); m.im[e] = R(0); }
}

template <class D, class S> __device__ __forceinline__ M3<D> convert(const M3<S>& a) {
  M3<D> m;
#pragma unroll
  for (int e = 0; e < 9; e++) { m.re[e] = D(a.re[e]); m.im[e] = D(a.im[e]); }
  return m;
}

// ---- HBM access ----------------------------------------------------------
// Site indices are 32-bit and so are the byte offsets of a site within a plane
// (the volume is below 2^28, checked when a context is created): the address of
// element e of slot l is the plane base f + (9 l + e) V, uniform across a
// wavefront when l is, plus a 32-bit per-lane byte offset.
using site_t = uint32_t;

template <class T> __device__ __forceinline__ const T* at(const T* plane, site_t s) {
  return (const T*)((const char*)plane + (uint32_t)(s * (uint32_t)sizeof(T)));
}
template <class T> __device__ __forceinline__ T* at(T* plane, site_t s) {
  return (T*)((char*)plane + (uint32_t)(s * (uint32_t)sizeof(T)));
}

template <class R>
__device__ __forceinline__ void load(M3<R>& m, const v2<R>* __restrict__ f, long V, int l, site_t s) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    v2<R> v = *at(f + (long)(9 * l + e) * V, s);
    m.re[e] = v.x;
    m.im[e] = v.y;
  }
}

// Loads and stores through a base pointer every lane of the wavefront holds
// the same value of: the base goes to scalar registers (readfirstlane) as a
// global-memory pointer, so each access is the scalar-base form (base +
// 32-bit lane offset) instead of a 64-bit vector add per access.  Only for
// bases that are uniform across the wavefront.
template <class T> using gptr = __attribute__((address_space(1))) T*;

template <class T> __device__ __forceinline__ gptr<T> uniform_ptr(const T* p) {
  const uint64_t u = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)u);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(u >> 32));
  return (gptr<T>)(((uint64_t)hi << 32) | lo);
}

template <class R> __device__ __forceinline__ v2<R> ld_u(const v2<R>* base, site_t s) {
  typedef R pair __attribute__((ext_vector_type(2)));
  gptr<const char> b = (gptr<const char>)uniform_ptr(base);
  const pair t = *(gptr<const pair>)(b + (uint32_t)(s * (uint32_t)sizeof(pair)));
  v2<R> v;
  v.x = t.x;
  v.y = t.y;
  return v;
}

template <class R> __device__ __forceinline__ void st_u(v2<R>* base, site_t s, v2<R> v) {
  typedef R pair __attribute__((ext_vector_type(2)));
  gptr<char> b = (gptr<char>)uniform_ptr(base);
  pair t;
  t.x = v.x;
  t.y = v.y;
  *(gptr<pair>)(b + (uint32_t)(s * (uint32_t)sizeof(pair))) = t;
}

// load / store of slot l, the same slot on every lane of the wavefront
template <class R>
__device__ __forceinline__ void load_u(M3<R>& m, const v2<R>* __restrict__ f, long V, int l, site_t s) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    v2<R> v = ld_u<R>(f + (long)(9 * l + e) * V, s);
    m.re[e] = v.x;
    m.im[e] = v.y;
  }
}

template <class R>
__device__ __forceinline__ void store_u(const M3<R>& m, v2<R>* __restrict__ f, long V, int l, site_t s) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    v2<R> v;
    v.x = m.re[e];
    v.y = m.im[e];
    st_u<R>(f + (long)(9 * l + e) * V, s, v);
  }
}

// Packed FP32 link storage (the mixed path's APE steps): elements 2q and 2q+1
// of slot l in one float4 of plane 5l + q (element 8 with a zero pair), so a
// link is 5 loads of 16 B per lane instead of 9 of 8 B.  The vector memory
// address unit, which bounds the APE step, spends a fixed cost per load
// instruction besides its cost per byte.
__device__ __forceinline__ float4 ld4_u(const float4* base, site_t s) {
  typedef float quad __attribute__((ext_vector_type(4)));
  gptr<const char> b = (gptr<const char>)uniform_ptr(base);
  const quad t = *(gptr<const quad>)(b + (uint32_t)(s * (uint32_t)sizeof(quad)));
  return make_float4(t.x, t.y, t.z, t.w);
}

__device__ __forceinline__ void st4_u(float4* base, site_t s, float4 v) {
  typedef float quad __attribute__((ext_vector_type(4)));
  gptr<char> b = (gptr<char>)uniform_ptr(base);
  quad t;
  t.x = v.x;
  t.y = v.y;
  t.z = v.z;
  t.w = v.w;
  *(gptr<quad>)(b + (uint32_t)(s * (uint32_t)sizeof(quad))) = t;
}

__device__ __forceinline__ void load_pk_u(M3<float>& m, const float4* __restrict__ f, long V, int l,
                                          site_t s) {
#pragma unroll
  for (int q = 0; q < 5; q++) {
    const float4 v = ld4_u(f + (long)(5 * l + q) * V, s);
    m.re[2 * q] = v.x;
    m.im[2 * q] = v.y;
    if (q < 4) {
      m.re[2 * q + 1] = v.z;
      m.im[2 * q + 1] = v.w;
    }
  }
}

__device__ __forceinline__ void store_pk_u(const M3<float>& m, float4* __restrict__ f, long V, int l,
                                           site_t s) {
#pragma unroll
  for (int q = 0; q < 5; q++)
    st4_u(f + (long)(5 * l + q) * V, s,
          make_float4(m.re[2 * q], m.im[2 * q], q < 4 ? m.re[2 * q + 1] : 0.f,
                      q < 4 ? m.im[2 * q + 1] : 0.f));
}

// link load / store of a field in the standard layout (PK = false) or, for
// FP32 links, in the packed layout (PK = true; the pointer then addresses
// float4 planes)
template <bool PK, class R>
__device__ __forceinline__ void ld_link(M3<R>& m, const v2<R>* __restrict__ f, long V, int l, site_t s) {
  if constexpr (PK)
    load_pk_u(m, (const float4*)f, V, l, s);
  else
    load_u(m, f, V, l, s);
}

template <bool PK, class R>
__device__ __forceinline__ void st_link(const M3<R>& m, v2<R>* __restrict__ f, long V, int l, site_t s) {
  if constexpr (PK)
    store_pk_u(m, (float4*)f, V, l, s);
  else
    store_u(m, f, V, l, s);
}

template <class R>
__device__ __forceinline__ void store(const M3<R>& m, v2<R>* __restrict__ f, long V, int l, site_t s) {
#pragma unroll
  for (int e = 0; e < 9; e++) {
    v2<R> v;
    v.x = m.re[e];
    v.y = m.im[e];
    *at(f + (long)(9 * l + e) * V, s) = v;
  }
}

// s, made to depend on v: loads whose addresses are computed from the result
// cannot be issued before v is formed.  Where a thread works through several
// links in turn, this keeps the compiler from issuing the loads of all of them
// at once, which spills the live matrices.  (An empty asm statement; it moves
// no data.)
template <class R> __device__ __forceinline__ site_t after(site_t s, const M3<R>& v) {
  asm volatile(""
               : "+v"(s)
               : "v"(v.re[0]), "v"(v.re[1]), "v"(v.re[2]), "v"(v.re[3]), "v"(v.re[4]),
                 "v"(v.re[5]), "v"(v.re[6]), "v"(v.re[7]), "v"(v.re[8]), "v"(v.im[0]),
                 "v"(v.im[1]), "v"(v.im[2]), "v"(v.im[3]), "v"(v.im[4]), "v"(v.im[5]),
                 "v"(v.im[6]), "v"(v.im[7]), "v"(v.im[8]));
  return s;
}

// ---- products (element order as the textbook sum over j = 0,1,2) ----------
// Each complex multiply-add is two fused multiply-adds per real component,
// accumulated in the order of the sum over j (x*y - z*w written as one
// expression costs a multiply, a fused multiply-add and an add).

template <class R> __device__ __forceinline__ R fmadd(R a, R b, R c) { return __builtin_fma(a, b, c); }
template <> __device__ __forceinline__ float fmadd<float>(float a, float b, float c) {
  return __builtin_fmaf(a, b, c);
}

// w = a * b
template <class R> __device__ __forceinline__ M3<R> mul(const M3<R>& a, const M3<R>& b) {
  M3<R> w;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = a.re[3 * i] * b.re[k];
      R im = a.re[3 * i] * b.im[k];
      re = fmadd(-a.im[3 * i], b.im[k], re);
      im = fmadd(a.im[3 * i], b.re[k], im);
#pragma unroll
      for (int j = 1; j < 3; j++) {
        re = fmadd(a.re[3 * i + j], b.re[3 * j + k], re);
        im = fmadd(a.re[3 * i + j], b.im[3 * j + k], im);
        re = fmadd(-a.im[3 * i + j], b.im[3 * j + k], re);
        im = fmadd(a.im[3 * i + j], b.re[3 * j + k], im);
      }
      w.re[3 * i + k] = re;
      w.im[3 * i + k] = im;
    }
  return w;
}

// w = a * b^dag
template <class R> __device__ __forceinline__ M3<R> mul_dag(const M3<R>& a, const M3<R>& b) {
  M3<R> w;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = a.re[3 * i] * b.re[3 * k];
      R im = a.im[3 * i] * b.re[3 * k];
      re = fmadd(a.im[3 * i], b.im[3 * k], re);
      im = fmadd(-a.re[3 * i], b.im[3 * k], im);
#pragma unroll
      for (int j = 1; j < 3; j++) {
        re = fmadd(a.re[3 * i + j], b.re[3 * k + j], re);
        im = fmadd(a.im[3 * i + j], b.re[3 * k + j], im);
        re = fmadd(a.im[3 * i + j], b.im[3 * k + j], re);
        im = fmadd(-a.re[3 * i + j], b.im[3 * k + j], im);
      }
      w.re[3 * i + k] = re;
      w.im[3 * i + k] = im;
    }
  return w;
}

// w = a^dag * b
template <class R> __device__ __forceinline__ M3<R> dag_mul(const M3<R>& a, const M3<R>& b) {
  M3<R> w;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = a.re[i] * b.re[k];
      R im = a.re[i] * b.im[k];
      re = fmadd(a.im[i], b.im[k], re);
      im = fmadd(-a.im[i], b.re[k], im);
#pragma unroll
      for (int j = 1; j < 3; j++) {
        re = fmadd(a.re[3 * j + i], b.re[3 * j + k], re);
        im = fmadd(a.re[3 * j + i], b.im[3 * j + k], im);
        re = fmadd(a.im[3 * j + i], b.im[3 * j + k], re);
        im = fmadd(-a.im[3 * j + i], b.re[3 * j + k], im);
      }
      w.re[3 * i + k] = re;
      w.im[3 * i + k] = im;
    }
  return w;
}

// v += a * b
template <class R> __device__ __forceinline__ void mul_acc(M3<R>& v, const M3<R>& a, const M3<R>& b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = v.re[3 * i + k], im = v.im[3 * i + k];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        re = fmadd(a.re[3 * i + j], b.re[3 * j + k], re);
        im = fmadd(a.re[3 * i + j], b.im[3 * j + k], im);
        re = fmadd(-a.im[3 * i + j], b.im[3 * j + k], re);
        im = fmadd(a.im[3 * i + j], b.re[3 * j + k], im);
      }
      v.re[3 * i + k] = re;
      v.im[3 * i + k] = im;
    }
}

// v += a^dag * b
template <class R>
__device__ __forceinline__ void dag_mul_acc(M3<R>& v, const M3<R>& a, const M3<R>& b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = v.re[3 * i + k], im = v.im[3 * i + k];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        re = fmadd(a.re[3 * j + i], b.re[3 * j + k], re);
        im = fmadd(a.re[3 * j + i], b.im[3 * j + k], im);
        re = fmadd(a.im[3 * j + i], b.im[3 * j + k], re);
        im = fmadd(-a.im[3 * j + i], b.re[3 * j + k], im);
      }
      v.re[3 * i + k] = re;
      v.im[3 * i + k] = im;
    }
}

// v += a * b^dag
template <class R>
__device__ __forceinline__ void mul_dag_acc(M3<R>& v, const M3<R>& a, const M3<R>& b) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      R re = v.re[3 * i + k], im = v.im[3 * i + k];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        re = fmadd(a.re[3 * i + j], b.re[3 * k + j], re);
        im = fmadd(a.im[3 * i + j], b.re[3 * k + j], im);
        re = fmadd(a.im[3 * i + j], b.im[3 * k + j], re);
        im = fmadd(-a.re[3 * i + j], b.im[3 * k + j], im);
      }
      v.re[3 * i + k] = re;
      v.im[3 * i + k] = im;
    }
}

template <class R> __device__ __forceinline__ void acc(M3<R>& v, const M3<R>& a) {
#pragma unroll
  for (int e = 0; e < 9; e++) { v.re[e] += a.re[e]; v.im[e] += a.im[e]; }
}

// Re tr(a^dag b), row by row
template <class R> __device__ __forceinline__ R retr_dag_mul(const M3<R>& a, const M3<R>& b) {
  R r[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    R t = a.re[3 * i] * b.re[3 * i];
    t = fmadd(a.im[3 * i], b.im[3 * i], t);
#pragma unroll
    for (int j = 1; j < 3; j++) {
      t = fmadd(a.re[3 * i + j], b.re[3 * i + j], t);
      t = fmadd(a.im[3 * i + j], b.im[3 * i + j], t);
    }
    r[i] = t;
  }
  return r[0] + r[1] + r[2];
}

// Re tr(a^dag b) as one fused multiply-add chain over the elements in order
template <class R> __device__ __forceinline__ R retr_dag_mul_chain(const M3<R>& a, const M3<R>& b) {
  R t = a.re[0] * b.re[0];
  t = fmadd(a.im[0], b.im[0], t);
#pragma unroll
  for (int e = 1; e < 9; e++) {
    t = fmadd(a.re[e], b.re[e], t);
    t = fmadd(a.im[e], b.im[e], t);
  }
  return t;
}

// ---- approximate SU(3) projection of hep-lat/0506008, always FP64 ---------
// Scale to unit Frobenius norm/3, then p times
//     X = u (3 - u^dag u)/2 ;  u = (1 - i Im det X / 3) X .
// Matrices with sum |u_ab|^2 < sqrt(18)*DBL_EPSILON are returned untouched
// (this keeps the zero links of open boundaries zero).
__device__ __forceinline__ double det_im(const M3<double>& u) {
  const double* r = u.re;
  const double* i = u.im;
  // minors d1 = u4 u8 - u5 u7, d2 = u3 u8 - u5 u6, d3 = u3 u7 - u4 u6
  double d1r = r[4] * r[8], d1i = r[4] * i[8];
  d1r = fmadd(-i[4], i[8], d1r);
  d1i = fmadd(i[4], r[8], d1i);
  d1r = fmadd(-r[5], r[7], d1r);
  d1i = fmadd(-r[5], i[7], d1i);
  d1r = fmadd(i[5], i[7], d1r);
  d1i = fmadd(-i[5], r[7], d1i);
  double d2r = r[3] * r[8], d2i = r[3] * i[8];
  d2r = fmadd(-i[3], i[8], d2r);
  d2i = fmadd(i[3], r[8], d2i);
  d2r = fmadd(-r[5], r[6], d2r);
  d2i = fmadd(-r[5], i[6], d2i);
  d2r = fmadd(i[5], i[6], d2r);
  d2i = fmadd(-i[5], r[6], d2i);
  double d3r = r[3] * r[7], d3i = r[3] * i[7];
  d3r = fmadd(-i[3], i[7], d3r);
  d3i = fmadd(i[3], r[7], d3i);
  d3r = fmadd(-r[4], r[6], d3r);
  d3i = fmadd(-r[4], i[6], d3i);
  d3r = fmadd(i[4], i[6], d3r);
  d3i = fmadd(-i[4], r[6], d3i);
  // Im (u0 d1 - u1 d2 + u2 d3)
  double t = r[0] * d1i;
  t = fmadd(i[0], d1r, t);
  t = fmadd(-r[1], d2i, t);
  t = fmadd(-i[1], d2r, t);
  t = fmadd(r[2], d3i, t);
  return fmadd(i[2], d3r, t);
}

__device__ __forceinline__ void project(M3<double>& u, int iter) {
  double d = 0.0;
#pragma unroll
  for (int e = 0; e < 9; e++) {
    d = fmadd(u.re[e], u.re[e], d);
    d = fmadd(u.im[e], u.im[e], d);
  }
  if (d < 4.242640687119285 * 2.220446049250313e-16) return;
  d = 1.0 / sqrt(d / 3.0);
#pragma unroll
  for (int e = 0; e < 9; e++) { u.re[e] *= d; u.im[e] *= d; }
  for (int it = 0; it < iter; it++) {
    // x = (3 - u^dag u)/2 is Hermitian: real diagonal xd[a], upper entries
    // xr/xi[ab] (ab = 01, 02, 12), x_ba = conj(x_ab)
    double xd[3], xr[3], xi[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      double re = u.re[a] * u.re[a];
      re = fmadd(u.im[a], u.im[a], re);
      re = fmadd(u.re[3 + a], u.re[3 + a], re);
      re = fmadd(u.im[3 + a], u.im[3 + a], re);
      re = fmadd(u.re[6 + a], u.re[6 + a], re);
      re = fmadd(u.im[6 + a], u.im[6 + a], re);
      xd[a] = fmadd(-0.5, re, 1.5);
    }
#pragma unroll
    for (int q = 0; q < 3; q++) {
      const int a = q == 2 ? 1 : 0, b = q == 0 ? 1 : 2;
      double re = u.re[a] * u.re[b];
      re = fmadd(u.im[a], u.im[b], re);
      re = fmadd(u.re[3 + a], u.re[3 + b], re);
      re = fmadd(u.im[3 + a], u.im[3 + b], re);
      re = fmadd(u.re[6 + a], u.re[6 + b], re);
      re = fmadd(u.im[6 + a], u.im[6 + b], re);
      double im = u.re[a] * u.im[b];
      im = fmadd(-u.im[a], u.re[b], im);
      im = fmadd(u.re[3 + a], u.im[3 + b], im);
      im = fmadd(-u.im[3 + a], u.re[3 + b], im);
      im = fmadd(u.re[6 + a], u.im[6 + b], im);
      im = fmadd(-u.im[6 + a], u.re[6 + b], im);
      xr[q] = -0.5 * re;
      xi[q] = -0.5 * im;
    }
    // X = u x, element (i,k) = sum_j u_ij x_jk in the order j = 0,1,2, the
    // real diagonal entry of x taken as a real factor
    M3<double> X;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) {
          const double ur = u.re[3 * i + j], ui = u.im[3 * i + j];
          if (j == k) {
            if (j == 0) {
              re = ur * xd[k];
              im = ui * xd[k];
            } else {
              re = fmadd(ur, xd[k], re);
              im = fmadd(ui, xd[k], im);
            }
          } else {
            // x_jk: upper entry if j < k, else the conjugate of x_kj
            const int q = (j + k == 1) ? 0 : ((j + k == 2) ? 1 : 2);
            const double br = xr[q], bi = j < k ? xi[q] : -xi[q];
            if (j == 0) {
              re = ur * br;
              im = ur * bi;
            } else {
              re = fmadd(ur, br, re);
              im = fmadd(ur, bi, im);
            }
            re = fmadd(-ui, bi, re);
            im = fmadd(ui, br, im);
          }
        }
        X.re[3 * i + k] = re;
        X.im[3 * i + k] = im;
      }
    const double zi = det_im(X) * (-1.0 / 3.0);
#pragma unroll
    for (int e = 0; e < 9; e++) {
      u.re[e] = fmadd(-zi, X.im[e], X.re[e]);
      u.im[e] = fmadd(zi, X.re[e], X.im[e]);
    }
  }
}

// ---- workgroup order ---------------------------------------------------------
// Workgroups are dispatched to the 8 XCDs of the chip round robin (workgroup
// b to XCD b % 8), each XCD with its own L2.  xcd_order maps b to a position
// such that every XCD works through one contiguous range of positions.
__device__ __forceinline__ int xcd_order(int b, int n) {
  const int q = n / 8, r = n % 8, x = b % 8, l = b / 8;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + l;
}

// ---- lattice geometry ------------------------------------------------------
// Lexicographic site s = x3 + N3*(x2 + N2*(x1 + N1*x0)); every direction wraps
// periodically in index space (open boundaries are zero time links).
struct Lat {
  int n[4];
  long st[4];  // strides: N1N2N3, N2N3, N3, 1
  long V;
  // division by n[m] of a site index r < 2^31: floor(r / n) =
  // mulhi(r, mg) >> sh, or r >> sh where n is a power of two (mg = 0)
  uint32_t mg[4], sh[4];
};

// extents, strides and division constants of a lattice.  For n not a power of
// two, sh = floor(log2 n) and mg = ceil(2^(32+sh) / n) < 2^32; with
// mg n = 2^(32+sh) + e, 0 <= e < n < 2^(sh+1), r mg / 2^(32+sh) differs from
// r / n by r e / (n 2^(32+sh)) < 1/n for r < 2^31, so the floor is exact.
inline void lat_set(Lat& L, const int dims[4]) {
  for (int m = 0; m < 4; m++) {
    L.n[m] = dims[m];
    const uint32_t d = (uint32_t)dims[m];
    uint32_t l = 0;
    while (l < 31 && (2u << l) <= d) l++;
    L.sh[m] = l;
    L.mg[m] = (d & (d - 1)) == 0 ? 0u : (uint32_t)((((uint64_t)1 << (32 + l)) + d - 1) / d);
  }
  L.st[3] = 1;
  L.st[2] = dims[3];
  L.st[1] = (long)dims[2] * dims[3];
  L.st[0] = (long)dims[1] * dims[2] * dims[3];
  L.V = (long)dims[0] * L.st[0];
}

__device__ __forceinline__ uint32_t div_n(const Lat& L, int m, uint32_t r) {
  return L.mg[m] ? (__umulhi(r, L.mg[m]) >> L.sh[m]) : (r >> L.sh[m]);
}

__device__ __forceinline__ void coords(const Lat& L, site_t s, int x[4]) {
  uint32_t r = s, q;
  q = div_n(L, 3, r); x[3] = (int)(r - q * (uint32_t)L.n[3]); r = q;
  q = div_n(L, 2, r); x[2] = (int)(r - q * (uint32_t)L.n[2]); r = q;
  q = div_n(L, 1, r); x[1] = (int)(r - q * (uint32_t)L.n[1]); r = q;
  x[0] = (int)r;
}

// site reached from s (coordinates x) by 0 <= d < N_mu steps in direction mu
// coordinate mu of x by selects (no run-time index into the array)
__device__ __forceinline__ int coord(const int x[4], int mu) {
  return mu == 0 ? x[0] : (mu == 1 ? x[1] : (mu == 2 ? x[2] : x[3]));
}
__device__ __forceinline__ site_t fwd(const Lat& L, site_t s, const int x[4], int mu, int d) {
  int y = coord(x, mu) + d;
  return (y >= L.n[mu]) ? s + (site_t)((long)(d - L.n[mu]) * L.st[mu]) : s + (site_t)((long)d * L.st[mu]);
}

__device__ __forceinline__ site_t bwd(const Lat& L, site_t s, const int x[4], int mu) {
  return (coord(x, mu) == 0) ? s + (site_t)((long)(L.n[mu] - 1) * L.st[mu]) : s - (site_t)L.st[mu];
}

}  // namespace wl
