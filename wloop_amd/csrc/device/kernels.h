// kernels.h -- host-side launchers of the gfx950 kernels (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "su3.hpp"

namespace wl {

// smear.hip
// TD: storage of the decorated links of levels 1 and 2 (D1, D2)
template <class TC, class TS, class TD>
void launch_hyp(const Lat& L, const v2<TS>* U, v2<TS>* out, v2<TD>* D1, v2<TD>* D2,
                const double a[3], int proj, int piter, int time, int space, int open_bc,
                hipStream_t st);
template <class TC, class TS>
void launch_hyp3(const Lat& L, const v2<TS>* F, v2<TS>* D, v2<TS>* out, double a1, double a2,
                 int proj, int piter, hipStream_t st);
template <class TC, class TS>
void launch_ape(const Lat& L, const v2<TS>* in, v2<TS>* out, double alpha, int proj, int piter,
                hipStream_t st);
// APE step with one thread per site for its three spatial links
template <class TC, class TS>
void launch_ape_site(const Lat& L, const v2<TS>* in, v2<TS>* out, double alpha, int proj, int piter,
                     hipStream_t st);
// APE step as a staple-sum kernel (sums into S, 3 slots of TC; per link, or
// per site when site != 0) and a lincomb/projection kernel
template <class TC, class TS>
void launch_ape_split(const Lat& L, const v2<TS>* in, v2<TC>* S, v2<TS>* out, double alpha, int proj,
                      int piter, int site, hipStream_t st);
void launch_ape4(const Lat& L, const double2* in, double2* out, double alpha, int proj, int piter,
                 hipStream_t st);
void launch_staple_sum(const Lat& L, const double2* A, const double2* B, double2* v, int mu, int nu,
                       hipStream_t st);

// gen.hip
void launch_generate(const Lat& L, double2* U, double2* G, uint64_t seed, double eps, int open_bc,
                     hipStream_t st);
void launch_aos_to_soa(long V, int nl, const double* a, double2* f, hipStream_t st);
void launch_soa_to_aos(long V, int nl, const double2* f, double* a, hipStream_t st);
void launch_zero_bc(const Lat& L, double2* U, hipStream_t st);
void launch_d2f(long n, const double2* a, float2* b, hipStream_t st);
void launch_f2d(long n, const float2* a, double2* b, hipStream_t st);
void launch_copy(long n, const double2* a, double2* b, hipStream_t st);
int plaquette_blocks(const Lat& L);
void launch_plaquette(const Lat& L, const double2* U, int open_bc, double* partial, int nblk,
                      hipStream_t st);

// wloop.hip
template <class TC, class TU, class TP, class TO>
void launch_time_transporter(const Lat& L, const v2<TU>* U, int lu, const v2<TP>* prev, v2<TO>* out,
                             float2* out32, hipStream_t st);
template <class TS>
void launch_space_transporter(const Lat& L, int msl, const v2<TS>* const* F,
                              const v2<TS>* const* prev, v2<TS>* out, hipStream_t st);
int contract_slots(const Lat& L, int kernel, int wg);
template <class TS, class TC>
void launch_contract(const Lat& L, int msl, int R, int T, const v2<TS>* LT,
                     const v2<TS>* const* S, double* partial, int kernel, int wg,
                     hipStream_t st);
// chain form for one (R,T) that also forms the spatial transporters of R
// (out: msl 3-slot fields, S_R = F S_{R-1}(x+k)), which it contracts; the
// two-lane form if two_lane and msl is 2 or 4, else the one-lane form; msl <= 4
template <class TS, class TC>
void launch_contract_build(const Lat& L, int msl, int R, int T, const v2<TS>* LT,
                           const v2<TS>* const* F, const v2<TS>* const* prev, v2<TS>* out,
                           double* partial, int wg, int two_lane, hipStream_t st);
// fused time-slice form: T = Ta .. Ta+d-1 of one R, W start slices per
// workgroup (one wave each), nslot workgroups per start group; partial points
// at the block of T = Ta, pstride apart per T
// APE step on links in the packed FP32 layout (5 float4 planes per slot:
// elements 2q, 2q+1 in plane 5l+q), and the conversions to and from it
void launch_ape_packed(const Lat& L, const float4* in, float4* out, double alpha, int proj, int piter,
                       hipStream_t st);
void launch_pack_links(long V, int nslot, const float2* src, float4* dst, hipStream_t st);
void launch_unpack_links(long V, int nslot, const float4* src, float2* dst, hipStream_t st);
// LDS bytes of a fused launch over d T (stored: the form over stored FP32
// transporters)
size_t contract_depth_lds(int msl, int W, int ts_bytes, int tc_bytes, int stored, int d);
template <class TS, class TC>
void launch_contract_depth(const Lat& L, int msl, int R, int Ta, int d, int W, int nslot,
                           const v2<TC>* LTa, const v2<TC>* U0, const v2<TS>* const* S,
                           const v2<TS>* const* F, v2<TS>* snext, double* partial, long pstride,
                           bool mx, hipStream_t st);
// fused launch over T = Ta .. Ta+d-1 with D_i(T) formed from the stored FP32
// transporters LTa + (T-Ta) 9V (mixed path, FP32 level links)
// (mx: the traces of every step on the matrix cores)
void launch_contract_stored(const Lat& L, int msl, int R, int Ta, int d, int W, int nslot,
                            const float2* LTa, const float2* const* S, const float2* const* F,
                            float2* snext, double* partial, long pstride, bool mx, hipStream_t st);
// W of one R (index r) from the partials of its (R,T) launches
void launch_finish_wloop(int N0, int msl, int mwlt, int mwlr, int r, int nwg, const double* partial,
                         double norm, double* wl, hipStream_t st);

}  // namespace wl

// ---- internal services of api.hip for the other files of the library ----
struct wlg_ctx;
// record an error message for wlg_error(); returns -1
__attribute__((visibility("hidden"))) int wl_set_error(const char* fmt, ...);
// the resident gauge field of a context and what launching on it needs
__attribute__((visibility("hidden"))) int wl_resident_field(wlg_ctx* c, int* dev,
                                                            const wl::Lat** L,
                                                            const double2** U, hipStream_t* st,
                                                            int* open_bc);

