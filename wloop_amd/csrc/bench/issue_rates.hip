// Issue cost of the instruction forms the kernels of this library are made
// of, on one CU of the GPU: each kernel runs 16 waves (4 per SIMD) of 8
// independent chains of one instruction form (so the rate, not the latency of
// a chain, is measured), or one chain (the latency).  Cycles are shader-clock
// cycles from s_memtime (its rate against the fixed 100 MHz s_memrealtime is
// printed), the slowest wave's count over (iterations x instructions per
// iteration x waves per SIMD): the SIMD cycles one wave instruction of the form
// takes when the SIMD issues nothing else.  Waves are not spread evenly over
// the four SIMDs, so each SIMD's rate is taken from the waves that ran on it
// (HW_ID): its slowest wave's cycles over (iterations x instructions per
// iteration x its waves), averaged over the SIMDs.  The loop branch is
// included (a few percent).  Output: one line per form.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

typedef float f4 __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                     \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) {                                                          \
      std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));                   \
      std::exit(1);                                                                  \
    }                                                                                \
  } while (0)

// per wave: shader cycles and 100 MHz ticks of the timed loop, and its SIMD
// (bits 5:4 of the HW_ID register, hardware register 4)
__device__ void record(long long* t, long long c0, long long c1, long long r0, long long r1) {
  const unsigned hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);
  if ((threadIdx.x & 63) == 0) {
    const int w = threadIdx.x >> 6;
    t[3 * w] = c1 - c0;
    t[3 * w + 1] = r1 - r0;
    t[3 * w + 2] = (hw >> 4) & 3;
  }
}

// 8 chains of one VALU instruction form (inline assembly: the exact encoding)
#define VALU32(NAME, ASM)                                                                    \
  __global__ void NAME(float* out, long long* t, int n) {                                    \
    float c[8], a = 1e-3f * threadIdx.x, b = 0.5f;                                           \
    for (int j = 0; j < 8; j++) c[j] = threadIdx.x + j;                                      \
    const long long c0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();                   \
    for (int it = 0; it < n; it++) {                                                         \
      _Pragma("unroll") for (int j = 0; j < 8; j++) asm volatile(ASM : "+v"(c[j]) : "v"(a), "v"(b)); \
    }                                                                                        \
    const long long c1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();                   \
    float s = 0.f;                                                                           \
    for (int j = 0; j < 8; j++) s += c[j];                                                   \
    out[threadIdx.x] = s;                                                                    \
    record(t, c0, c1, r0, r1);                                                               \
  }
#define VALU64(NAME, ASM)                                                                    \
  __global__ void NAME(float* out, long long* t, int n) {                                    \
    double c[8], a = 1e-3 * threadIdx.x, b = 0.5;                                            \
    for (int j = 0; j < 8; j++) c[j] = threadIdx.x + j;                                      \
    const long long c0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();                   \
    for (int it = 0; it < n; it++) {                                                         \
      _Pragma("unroll") for (int j = 0; j < 8; j++) asm volatile(ASM : "+v"(c[j]) : "v"(a), "v"(b)); \
    }                                                                                        \
    const long long c1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();                   \
    double s = 0.0;                                                                          \
    for (int j = 0; j < 8; j++) s += c[j];                                                   \
    out[threadIdx.x] = (float)s;                                                             \
    record(t, c0, c1, r0, r1);                                                               \
  }

// v_fmac_f32 with the operands in fixed VGPRs (VGPR bank = register number mod
// 4): a in v8 (bank 0), b in v9 (bank 1), the 8 accumulators in banks 2 and 3
// (no bank shared by two operands of one instruction) or in banks 0 and 1
// (every instruction reads two operands from one bank)
#define FMAC_PINNED(NAME, R0, R1, R2, R3, R4, R5, R6, R7)                                   \
  __global__ void NAME(float* out, long long* t, int n) {                                  \
    float c0 = threadIdx.x, c1 = c0 + 1, c2 = c0 + 2, c3 = c0 + 3, c4 = c0 + 4, c5 = c0 + 5, \
          c6 = c0 + 6, c7 = c0 + 7, a = 1e-3f * threadIdx.x, b = 0.5f;                    \
    const long long t0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();                 \
    for (int it = 0; it < n; it++) {                                                       \
      asm volatile("v_fmac_f32 %0, %8, %9\n v_fmac_f32 %1, %8, %9\n"                     \
                   " v_fmac_f32 %2, %8, %9\n v_fmac_f32 %3, %8, %9\n"                    \
                   " v_fmac_f32 %4, %8, %9\n v_fmac_f32 %5, %8, %9\n"                    \
                   " v_fmac_f32 %6, %8, %9\n v_fmac_f32 %7, %8, %9"                       \
                   : "+{" R0 "}"(c0), "+{" R1 "}"(c1), "+{" R2 "}"(c2), "+{" R3 "}"(c3),    \
                     "+{" R4 "}"(c4), "+{" R5 "}"(c5), "+{" R6 "}"(c6), "+{" R7 "}"(c7)     \
                   : "{v8}"(a), "{v9}"(b));                                                \
    }                                                                                      \
    const long long t1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();                 \
    out[threadIdx.x] = c0 + c1 + c2 + c3 + c4 + c5 + c6 + c7;                              \
    record(t, t0, t1, r0, r1);                                                             \
  }
FMAC_PINNED(fmac_banks_free, "v10", "v11", "v14", "v15", "v18", "v19", "v22", "v23")
FMAC_PINNED(fmac_banks_shared, "v12", "v13", "v16", "v17", "v20", "v21", "v24", "v25")

VALU32(fmac_f32, "v_fmac_f32 %0, %1, %2")                  // 4-byte encoding
VALU32(mul_f32, "v_mul_f32 %0, %1, %0")
VALU32(add_u32, "v_add_u32 %0, %1, %0")
VALU32(fma_f32, "v_fma_f32 %0, %1, %2, %0")                // 8-byte encoding
VALU32(fma_f32_neg, "v_fma_f32 %0, -%1, %2, %0")
VALU32(mul_f32_e64, "v_mul_f32_e64 %0, %1, %0")
VALU32(cvt_f32_i32, "v_cvt_f32_i32 %0, %1")
VALU32(mov_dpp, "v_mov_b32_dpp %0, %1 row_ror:8 row_mask:0xf bank_mask:0xf")
VALU64(fma_f64, "v_fma_f64 %0, %1, %2, %0")
VALU64(add_f64, "v_add_f64 %0, %1, %0")

// CH chains of v_mfma_f32_4x4x1f32 (16 blocks of 4x4 outer products)
template <int CH>
__global__ void mfma4(float* out, long long* t, int n) {
  f4 c[CH];
  for (int j = 0; j < CH; j++) c[j] = f4{0.f, 0.f, 0.f, 0.f};
  const float a = 1e-3f * threadIdx.x, b = 1.0f;
  const long long c0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < n; it++) {
#pragma unroll
    for (int j = 0; j < CH; j++) c[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c[j], 0, 0, 0);
  }
  const long long c1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();
  float s = 0.f;
  for (int j = 0; j < CH; j++) s += c[j][0] + c[j][1] + c[j][2] + c[j][3];
  out[threadIdx.x] = s;
  record(t, c0, c1, r0, r1);
}

// 8 chains of v_mfma_f64_4x4x4f64 (4 blocks of 4x4x4)
__global__ void mfma64(float* out, long long* t, int n) {
  double c[8];
  for (int j = 0; j < 8; j++) c[j] = 0.0;
  const double a = 1e-3 * threadIdx.x, b = 1.0;
  const long long c0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < n; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) c[j] = __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c[j], 0, 0, 0);
  }
  const long long c1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();
  double s = 0.0;
  for (int j = 0; j < 8; j++) s += c[j];
  out[threadIdx.x] = (float)s;
  record(t, c0, c1, r0, r1);
}

// 8 chains of v_mfma_f32_4x4x1f32 and 16 of v_fmac_f32 interleaved in one
// wave: the SIMD time of the mix against the sum of its parts
__global__ void mix(float* out, long long* t, int n) {
  f4 c[8];
  float v[16];
  for (int j = 0; j < 8; j++) c[j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int j = 0; j < 16; j++) v[j] = 0.f;
  const float a = 1e-3f * threadIdx.x, b = 1.0f;
  const long long c0 = clock64(), r0 = __builtin_amdgcn_s_memrealtime();
  for (int it = 0; it < n; it++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      c[j] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c[j], 0, 0, 0);
      v[2 * j] = __builtin_fmaf(a, b, v[2 * j]);
      v[2 * j + 1] = __builtin_fmaf(a, b, v[2 * j + 1]);
    }
  }
  const long long c1 = clock64(), r1 = __builtin_amdgcn_s_memrealtime();
  float s = 0.f;
  for (int j = 0; j < 8; j++) s += c[j][0] + c[j][3];
  for (int j = 0; j < 16; j++) s += v[j];
  out[threadIdx.x] = s;
  record(t, c0, c1, r0, r1);
}

constexpr int kThreads = 1024;  // 16 waves: 4 per SIMD
constexpr int kWaves = kThreads / 64;

static void run(const char* name, void (*kern)(float*, long long*, int), int ops, int wps, int n,
                float* out, long long* t) {
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  hipLaunchKernelGGL(kern, dim3(1), dim3(wps * 256), 0, 0, out, t, n / 8);  // warm-up
  CHECK(hipEventRecord(e0, 0));
  hipLaunchKernelGGL(kern, dim3(1), dim3(wps * 256), 0, 0, out, t, n);
  CHECK(hipEventRecord(e1, 0));
  CHECK(hipGetLastError());
  CHECK(hipDeviceSynchronize());
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  long long h[3 * kWaves];
  CHECK(hipMemcpy(h, t, sizeof(h), hipMemcpyDeviceToHost));
  long long cmax[4] = {0, 0, 0, 0}, rmax = 0, cyc = 0;
  int cnt[4] = {0, 0, 0, 0};
  for (int w = 0; w < 4 * wps; w++) {
    const int sd = (int)h[3 * w + 2];
    cnt[sd]++;
    if (h[3 * w] > cmax[sd]) cmax[sd] = h[3 * w];
    if (h[3 * w] > cyc) cyc = h[3 * w];
    if (h[3 * w + 1] > rmax) rmax = h[3 * w + 1];
  }
  double rate = 0.0;
  int used = 0;
  for (int sd = 0; sd < 4; sd++)
    if (cnt[sd]) {
      rate += (double)cmax[sd] / ((double)n * ops * cnt[sd]);
      used++;
    }
  const double ghz = (double)cyc / (rmax * 10.0);
  std::printf("%-46s waves per SIMD %d/%d/%d/%d  %6.2f cycles per wave instruction  (clock %.2f GHz; "
              "kernel %.3f ms = %.2f cycles per instruction at that clock)\n",
              name, cnt[0], cnt[1], cnt[2], cnt[3], rate / used, ghz, ms,
              ms * 1e6 * ghz / ((double)n * ops * wps));
}

int main() {
  float* out;
  long long* t;
  CHECK(hipMalloc(&out, kThreads * sizeof(float)));
  CHECK(hipMalloc(&t, 3 * kWaves * sizeof(long long)));
  const int n = 1 << 15;
  run("v_fmac_f32 (VOP2)", fmac_f32, 8, 4, n, out, t);
  run("v_fmac_f32 (VOP2), no shared VGPR bank", fmac_banks_free, 8, 4, n, out, t);
  run("v_fmac_f32 (VOP2), two operands in one bank", fmac_banks_shared, 8, 4, n, out, t);
  run("v_mul_f32 (VOP2)", mul_f32, 8, 4, n, out, t);
  run("v_add_u32 (VOP2)", add_u32, 8, 4, n, out, t);
  run("v_fma_f32 (VOP3)", fma_f32, 8, 4, n, out, t);
  run("v_fma_f32 -a (VOP3)", fma_f32_neg, 8, 4, n, out, t);
  run("v_mul_f32_e64 (VOP3)", mul_f32_e64, 8, 4, n, out, t);
  run("v_cvt_f32_i32", cvt_f32_i32, 8, 4, n, out, t);
  run("v_mov_b32_dpp row_ror", mov_dpp, 8, 4, n, out, t);
  run("v_fma_f64", fma_f64, 8, 4, n, out, t);
  run("v_add_f64", add_f64, 8, 4, n, out, t);
  run("v_mfma_f32_4x4x1f32, 8 chains", mfma4<8>, 8, 4, n, out, t);
  run("v_mfma_f32_4x4x1f32, 8 chains", mfma4<8>, 8, 1, n, out, t);
  run("v_mfma_f32_4x4x1f32, 1 chain (latency)", mfma4<1>, 1, 1, n, out, t);
  run("v_mfma_f64_4x4x4f64, 8 chains", mfma64, 8, 4, n, out, t);
  run("8 v_mfma_f32_4x4x1f32 + 16 v_fmac_f32 (per 8)", mix, 8, 4, n, out, t);
  CHECK(hipFree(out));
  CHECK(hipFree(t));
  return 0;
}
