// The per-timeslice context behind laph_ctx* (the MatShell context of the
// plain FP64 drop-in): device links and the scratch vectors of the filtered
// matvec.  Everything the reference keeps in process singletons
// (Tslice, Nav, IO: shell_matop.cpp:9-11) lives here instead.
#pragma once

#include "common.h"
#include "kernels.h"

struct laph_ctx {
  laph_params p{};
  int V3 = 0;
  size_t n = 0;           // 3*V3 complex entries per vector
  double c = 0, a = 0;    // B = c*D + a   (shell_matop.cpp:49-50)
  hipStream_t stream = nullptr;
  double2* U = nullptr;   // AoS links [(s*3+dir)*9 + e]
  double2* Us = nullptr;  // SoA links for the stencil [(dir*9+e)*V3 + s]
  double2* dec = nullptr; // HYP decorated links (6 per site)
  double2* Unew = nullptr;
  double2* vx = nullptr;  // host-boundary staging: x
  double2* vy = nullptr;  //                        y
  double2* t[3] = {nullptr, nullptr, nullptr};  // Chebyshev recurrence buffers
  bool links_dirty = true;
};

namespace laph {
// y = T_deg(B) x on the device (x, y distinct device vectors of ctx->n).
int cheb_apply(laph_ctx* ctx, const double2* x, double2* y);
// y = D x on the device.
int laplacian_apply(laph_ctx* ctx, const double2* x, double2* y);
// Refresh the SoA copy of the links if the AoS links changed.
int ensure_packed(laph_ctx* ctx);
}  // namespace laph
