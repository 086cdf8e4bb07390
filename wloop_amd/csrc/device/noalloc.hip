// noalloc.hip -- entry points that allocate no device memory.
//
// For callers that cannot call hipMalloc (code under HIP graph capture, tight
// serving loops, harnesses that forbid it), the scratch these entry points
// need lives in a file-scope device array of fixed size, one copy per GPU,
// held until the process ends.  Its device address is looked up once per
// device with hipGetSymbolAddress; host code uses that address only (the
// array's name in host code is a stand-in, not a device pointer).
//
// Every call on a GPU shares that GPU's array and leaves its contents there;
// calls that overlap on one GPU (from several host threads, or on several
// streams) race on it and must be serialised by the caller.
#include <hip/hip_runtime.h>
#include <mutex>

#include "kernels.h"
#include "wloop_gpu.h"

using namespace wl;

namespace {

// per-workgroup partial sums of the plaquette (plaquette_blocks() <= 1024)
// and the final sum
constexpr int kScratchDoubles = 2048;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices];  // device address of `scratch` on each device

// the device address of this device's scratch array (current device = dev)
int scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) return wl_set_error("device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return wl_set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                          hipGetErrorString(e));
    g_ptr[dev] = (double*)p;
  }
  *out = g_ptr[dev];
  return 0;
}

// sum of the partials in index order (the order wlg_plaquette sums them on
// the host: the same additions, so the same bits)
__global__ void sum_in_order(const double* __restrict__ part, int n, double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += part[i];
  *out = s;
}

}  // namespace

extern "C" int wlg_plaquette_noalloc(wlg_ctx* c, double* sum) {
  int dev = 0;
  const Lat* L = nullptr;
  const double2* U = nullptr;
  hipStream_t st = nullptr;
  int open_bc = 0;
  if (wl_resident_field(c, &dev, &L, &U, &st, &open_bc)) return -1;
  hipError_t e = hipSetDevice(dev);
  if (e != hipSuccess) return wl_set_error("hipSetDevice(%d) failed: %s", dev, hipGetErrorString(e));
  double* ptr = nullptr;
  if (scratch_ptr(dev, &ptr)) return -1;
  const int nblk = plaquette_blocks(*L);
  if (nblk + 1 > kScratchDoubles)
    return wl_set_error("plaquette needs %d partial sums and the total, more than the %d doubles "
                        "of the static scratch", nblk, kScratchDoubles);
  launch_plaquette(*L, U, open_bc, ptr, nblk, st);
  hipLaunchKernelGGL(sum_in_order, dim3(1), dim3(64), 0, st, ptr, nblk, ptr + nblk);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(sum, ptr + nblk, sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return wl_set_error("plaquette failed: %s", hipGetErrorString(e));
  return 0;
}
