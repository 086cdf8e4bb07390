// gen.hip -- synthetic gauge configurations on the device, plus layout
// conversion and boundary kernels.
//
// Generator (no reference counterpart; openQCD's random_ud is not part of the
// path): counter-based Philox4x32-10 (Salmon et al., SC'11) keyed by the
// 64-bit seed, counter (site, stream, draw, tag); Box-Muller normals;
// U = exp(i eps H) with H Hermitian traceless (N(0,1) real components),
// scaled Taylor series + squaring, Gram-Schmidt reunitarisation; then a
// Haar-random SU(3) gauge transformation U_mu(x) <- g(x) U_mu(x) g(x+mu)^dag.
// Every site is independent, so the result does not depend on the launch
// shape; it agrees with the CPU restatement to rounding of the libm calls.
#include "su3.hpp"
#include "kernels.h"

namespace wl {

__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ void gauss2(uint64_t seed, uint32_t site, uint32_t stream, uint32_t draw,
                                       double* g) {
  uint32_t c[4] = {site, stream, draw, 0x574c4f4fu};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const double two53 = 1.0 / 9007199254740992.0;
  double u1 = ((double)(c[0] >> 5) * 67108864.0 + (double)(c[1] >> 6)) * two53;
  double u2 = ((double)(c[2] >> 5) * 67108864.0 + (double)(c[3] >> 6)) * two53;
  double rad = sqrt(-2.0 * log(1.0 - u1));
  double th = 6.283185307179586 * u2;
  g[0] = rad * cos(th);
  g[1] = rad * sin(th);
}

// rows 1,2 Gram-Schmidt, row 3 = conj(r1 x r2)
__device__ __forceinline__ void reunitarize(M3<double>& m) {
  double n = 0.0;
  for (int b = 0; b < 3; b++) n += m.re[b] * m.re[b] + m.im[b] * m.im[b];
  n = 1.0 / sqrt(n);
  for (int b = 0; b < 3; b++) { m.re[b] *= n; m.im[b] *= n; }
  double pr = 0.0, pi = 0.0;
  for (int b = 0; b < 3; b++) {
    pr += m.re[b] * m.re[3 + b] + m.im[b] * m.im[3 + b];
    pi += m.re[b] * m.im[3 + b] - m.im[b] * m.re[3 + b];
  }
  for (int b = 0; b < 3; b++) {
    m.re[3 + b] -= pr * m.re[b] - pi * m.im[b];
    m.im[3 + b] -= pr * m.im[b] + pi * m.re[b];
  }
  n = 0.0;
  for (int b = 0; b < 3; b++) n += m.re[3 + b] * m.re[3 + b] + m.im[3 + b] * m.im[3 + b];
  n = 1.0 / sqrt(n);
  for (int b = 0; b < 3; b++) { m.re[3 + b] *= n; m.im[3 + b] *= n; }
  for (int b = 0; b < 3; b++) {
    int b1 = (b + 1) % 3, b2 = (b + 2) % 3;
    double cr = (m.re[b1] * m.re[3 + b2] - m.im[b1] * m.im[3 + b2]) -
                (m.re[b2] * m.re[3 + b1] - m.im[b2] * m.im[3 + b1]);
    double ci = (m.re[b1] * m.im[3 + b2] + m.im[b1] * m.re[3 + b2]) -
                (m.re[b2] * m.im[3 + b1] + m.im[b2] * m.re[3 + b1]);
    m.re[6 + b] = cr;
    m.im[6 + b] = -ci;
  }
}

__device__ M3<double> random_link(uint64_t seed, uint32_t site, uint32_t mu, double eps) {
  double g[10];
  for (int d = 0; d < 5; d++) gauss2(seed, site, mu, (uint32_t)d, g + 2 * d);
  double tr = (g[0] + g[1] + g[2]) / 3.0;
  // Q = i eps H  (H_aa = g_a - tr, H_ab = g + i g', H_ba = conj)
  M3<double> q;
  double hd[3] = {g[0] - tr, g[1] - tr, g[2] - tr};
  for (int a = 0; a < 3; a++) { q.re[4 * a] = -eps * 0.0; q.im[4 * a] = eps * hd[a]; }
  const int ab[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  for (int p = 0; p < 3; p++) {
    int a = ab[p][0], b = ab[p][1];
    double hr = g[3 + 2 * p], hi = g[4 + 2 * p];
    q.re[3 * a + b] = -eps * hi;
    q.im[3 * a + b] = eps * hr;
    q.re[3 * b + a] = -eps * (-hi);
    q.im[3 * b + a] = eps * hr;
  }
  double nrm = 0.0;
  for (int e = 0; e < 9; e++) nrm += q.re[e] * q.re[e] + q.im[e] * q.im[e];
  nrm = sqrt(nrm);
  int s = 0;
  while (nrm > 0.25) { nrm *= 0.5; s++; }
  double sc = ldexp(1.0, -s);
  for (int e = 0; e < 9; e++) { q.re[e] *= sc; q.im[e] *= sc; }
  M3<double> e;
  set_zero(e);
  for (int a = 0; a < 3; a++) e.re[4 * a] = 1.0;
  for (int k = 12; k >= 1; k--) {
    M3<double> t = mul(q, e);
    double inv = 1.0 / (double)k;
    for (int x = 0; x < 9; x++) { t.re[x] *= inv; t.im[x] *= inv; }
    for (int a = 0; a < 3; a++) t.re[4 * a] += 1.0;
    e = t;
  }
  for (int k = 0; k < s; k++) e = mul(e, e);
  reunitarize(e);
  return e;
}

__device__ M3<double> random_su3(uint64_t seed, uint32_t site, uint32_t stream) {
  double g[12];
  for (int d = 0; d < 6; d++) gauss2(seed, site, stream, (uint32_t)d, g + 2 * d);
  M3<double> m;
  set_zero(m);
  for (int b = 0; b < 3; b++) {
    m.re[b] = g[2 * b];
    m.im[b] = g[2 * b + 1];
    m.re[3 + b] = g[6 + 2 * b];
    m.im[3 + b] = g[7 + 2 * b];
  }
  reunitarize(m);
  return m;
}

// pass 1: U_mu(x) = exp(i eps H) for mu = blockIdx.y < 4; blockIdx.y == 4 -> g(x)
__global__ __launch_bounds__(256) void gen_links(Lat L, double2* __restrict__ U,
                                                 double2* __restrict__ G, uint64_t seed,
                                                 double eps) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int mu = blockIdx.y;
  if (mu < 4)
    store(random_link(seed, (uint32_t)s, (uint32_t)mu, eps), U, L.V, mu, s);
  else
    store(random_su3(seed, (uint32_t)s, 4u), G, L.V, 0, s);
}

// pass 2: U_mu(x) <- g(x) U_mu(x) g(x+mu)^dag, then open bc zeroing
__global__ __launch_bounds__(256) void gen_transform(Lat L, double2* __restrict__ U,
                                                     const double2* __restrict__ G, int open_bc) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= L.V) return;
  int mu = blockIdx.y;
  int x[4];
  coords(L, s, x);
  M3<double> u, g0, g1;
  load(u, U, L.V, mu, s);
  load(g0, G, L.V, 0, s);
  load(g1, G, L.V, 0, fwd(L, s, x, mu, 1));
  u = mul(g0, mul_dag(u, g1));
  if (open_bc && mu == 0 && x[0] == L.n[0] - 1) set_zero(u);
  store(u, U, L.V, mu, s);
}

// AoS lexicographic host image [V][nl][18] <-> SoA planes [nl][9][V]
__global__ __launch_bounds__(256) void aos_to_soa(long V, int nl, const double* __restrict__ a,
                                                  double2* __restrict__ f) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  int l = blockIdx.y;
  const double2* p = reinterpret_cast<const double2*>(a + 18 * ((long)s * nl + l));
  for (int e = 0; e < 9; e++) f[((long)l * 9 + e) * V + s] = p[e];
}

__global__ __launch_bounds__(256) void soa_to_aos(long V, int nl, const double2* __restrict__ f,
                                                  double* __restrict__ a) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  int l = blockIdx.y;
  double2* p = reinterpret_cast<double2*>(a + 18 * ((long)s * nl + l));
  for (int e = 0; e < 9; e++) p[e] = f[((long)l * 9 + e) * V + s];
}

// U_0(x0 = N0-1) = 0 (openQCD set_bc for open boundary conditions)
__global__ __launch_bounds__(256) void zero_last_time_links(Lat L, double2* __restrict__ U) {
  long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long slice = L.st[0];
  if (s >= slice) return;
  long t = (long)(L.n[0] - 1) * slice + s;
  M3<double> z;
  set_zero(z);
  store(z, U, L.V, 0, t);
}

// element-wise precision conversion of n 2-vectors
__global__ __launch_bounds__(256) void d2f(long n, const double2* __restrict__ a,
                                           float2* __restrict__ b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

// Plaquette sum per workgroup (lnkplaq.c:48-128): per site the three spatial
// planes and, unless x0 = N0-1 with open bc, the three time planes;
// Re tr[(U_mu(x)U_nu(x+mu))^dag U_nu(x)U_mu(x+nu)].  Grid-stride loop over
// sites, fixed-order reduction to one partial per workgroup.
__global__ __launch_bounds__(256) void plaquette_kernel(Lat L, const double2* __restrict__ U,
                                                        int open_bc, double* __restrict__ partial) {
  double acc = 0.0;
  for (long s = (long)blockIdx.x * 256 + threadIdx.x; s < L.V; s += (long)gridDim.x * 256) {
    int x[4];
    coords(L, s, x);
    for (int mu = 1; mu < 4; mu++) {
      for (int nu = 0; nu < mu; nu++) {
        if (nu == 0 && open_bc && x[0] == L.n[0] - 1) continue;
        M3<double> a, b, c, d;
        load(a, U, L.V, mu, s);
        load(b, U, L.V, nu, fwd(L, s, x, mu, 1));
        load(c, U, L.V, nu, s);
        load(d, U, L.V, mu, fwd(L, s, x, nu, 1));
        acc += retr_dag_mul(mul(a, b), mul(c, d));
      }
    }
  }
  __shared__ double red[4];
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ __launch_bounds__(256) void copy2(long n, const double2* __restrict__ a,
                                             double2* __restrict__ b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void f2d(long n, const float2* __restrict__ a,
                                           double2* __restrict__ b) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}

static inline dim3 grid_for(long V, int ny) { return dim3((unsigned)((V + 255) / 256), ny, 1); }

void launch_generate(const Lat& L, double2* U, double2* G, uint64_t seed, double eps, int open_bc,
                     hipStream_t st) {
  hipLaunchKernelGGL(gen_links, grid_for(L.V, 5), dim3(256), 0, st, L, U, G, seed, eps);
  hipLaunchKernelGGL(gen_transform, grid_for(L.V, 4), dim3(256), 0, st, L, U,
                     (const double2*)G, open_bc);
}

void launch_aos_to_soa(long V, int nl, const double* a, double2* f, hipStream_t st) {
  hipLaunchKernelGGL(aos_to_soa, grid_for(V, nl), dim3(256), 0, st, V, nl, a, f);
}

void launch_soa_to_aos(long V, int nl, const double2* f, double* a, hipStream_t st) {
  hipLaunchKernelGGL(soa_to_aos, grid_for(V, nl), dim3(256), 0, st, V, nl, f, a);
}

void launch_zero_bc(const Lat& L, double2* U, hipStream_t st) {
  hipLaunchKernelGGL(zero_last_time_links, grid_for(L.st[0], 1), dim3(256), 0, st, L, U);
}

int plaquette_blocks(const Lat& L) {
  long b = (L.V + 255) / 256;
  return (int)(b < 1024 ? b : 1024);
}

void launch_plaquette(const Lat& L, const double2* U, int open_bc, double* partial, int nblk,
                      hipStream_t st) {
  hipLaunchKernelGGL(plaquette_kernel, dim3(nblk), dim3(256), 0, st, L, U, open_bc, partial);
}

void launch_copy(long n, const double2* a, double2* b, hipStream_t st) {
  hipLaunchKernelGGL(copy2, grid_for(n, 1), dim3(256), 0, st, n, a, b);
}

void launch_f2d(long n, const float2* a, double2* b, hipStream_t st) {
  hipLaunchKernelGGL(f2d, grid_for(n, 1), dim3(256), 0, st, n, a, b);
}

void launch_d2f(long n, const double2* a, float2* b, hipStream_t st) {
  hipLaunchKernelGGL(d2f, grid_for(n, 1), dim3(256), 0, st, n, a, b);
}

}  // namespace wl
