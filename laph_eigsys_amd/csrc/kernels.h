// Launchers of the liblapheig device kernels (all asynchronous on `st`).
#pragma once

#include <hip/hip_runtime.h>

namespace laph {

// --- geometry --------------------------------------------------------------
// site s = x*L*L + y*L + z; dir 0 = x (stride L*L), 1 = y (stride L), 2 = z.

// --- HYP (timeslice.cpp:243-369) --------------------------------------------
// U, Unew: AoS links [(s*3+dir)*9 + 3a+b]; dec: 6 decorated links per site.
hipError_t launch_hyp_level1(const double2* U, double2* dec, int L, double alpha2, hipStream_t st);
hipError_t launch_hyp_level2(const double2* U, const double2* dec, double2* Unew, int L,
                             double alpha1, hipStream_t st);

// AoS links -> SoA stencil layouts: Us64[(dir*9+e)*V3 + s] (FP64) and
// Us32[(dir*9+e)*V3 + s] (FP32 complex).
hipError_t launch_pack_links64(const double2* U, double2* Us, int L, hipStream_t st);
hipError_t launch_pack_links32(const double2* U, float2* Us, int L, hipStream_t st);

// --- FP64 single-vector stencil + Chebyshev step (shell_matop.cpp:16-168) --
// out = alpha * (c*(H in - 6 in) + a*in) + beta * old, vectors [3*s + colour].
// old may be null when beta == 0.
hipError_t launch_cheb_step64(const double2* Us, const double2* in, const double2* old,
                              double2* out, int L, double c, double a, double alpha,
                              double beta, hipStream_t st);
// The same step reading the links in the AoS host layout U[(s*3+dir)*9 + e]
// (the arithmetic is identical, so the results are bitwise equal).
hipError_t launch_cheb_step64_aos(const double2* U, const double2* in, const double2* old,
                                  double2* out, int L, double c, double a, double alpha,
                                  double beta, hipStream_t st);

}  // namespace laph
