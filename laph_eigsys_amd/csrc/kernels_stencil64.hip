// FP64 single-vector covariant 3D Laplacian fused with one Chebyshev
// recurrence step -- the plain drop-in of tv2/tv_iter
// (shell_matop.cpp:16-103, 136-168):
//   (B x)(s) = c * (sum_mu [U_mu(s) x(s+mu) + U_mu(s-mu)^H x(s-mu)] - 6 x(s)) + a x(s)
//   out      = alpha * (B in) + beta * old          (alpha,beta) = (1,0) or (2,-1)
// One thread per (site, colour row): 3*V3 threads, each row reads its three
// link rows (SoA, coalesced over sites) and the six neighbour colour vectors.
// Algorithmic traffic per site and step: links 27*16 B + in/old/out 3*48 B
// = 576 B (links and neighbour vectors re-used through L2/MALL).
#include "kernels.h"

namespace laph {

__device__ __forceinline__ double2 fma_c(double2 u, double2 v, double2 acc) {
  acc.x = fma(u.x, v.x, fma(-u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(u.y, v.x, acc.y));
  return acc;
}
// acc += conj(u) * v
__device__ __forceinline__ double2 fma_cc(double2 u, double2 v, double2 acc) {
  acc.x = fma(u.x, v.x, fma(u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(-u.y, v.x, acc.y));
  return acc;
}

__global__ __launch_bounds__(256) void k_cheb_step64(const double2* __restrict__ Us,
                                                      const double2* __restrict__ in,
                                                      const double2* __restrict__ old,
                                                      double2* __restrict__ out, int L, double c,
                                                      double a, double alpha, double beta) {
  const int V3 = L * L * L;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= 3 * V3) return;
  const int s = tid / 3, r = tid - 3 * s;
  const int L2 = L * L;
  const int x = s / L2, y = (s / L) % L, z = s % L;
  const int up[3] = {x == L - 1 ? s - (L - 1) * L2 : s + L2, y == L - 1 ? s - (L - 1) * L : s + L,
                     z == L - 1 ? s - (L - 1) : s + 1};
  const int dn[3] = {x == 0 ? s + (L - 1) * L2 : s - L2, y == 0 ? s + (L - 1) * L : s - L,
                     z == 0 ? s + (L - 1) : s - 1};
  double2 acc = make_double2(0.0, 0.0);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    const double2* Um = Us + (size_t)mu * 9 * V3;
    const double2* xu = in + 3 * (size_t)up[mu];
    const double2* xm = in + 3 * (size_t)dn[mu];
    // U_mu(s) row r . x(s+mu)
#pragma unroll
    for (int b = 0; b < 3; ++b) acc = fma_c(Um[(size_t)(3 * r + b) * V3 + s], xu[b], acc);
    // (U_mu(s-mu)^H)(r, b) = conj(U_mu(s-mu)(b, r))
#pragma unroll
    for (int b = 0; b < 3; ++b) acc = fma_cc(Um[(size_t)(3 * b + r) * V3 + dn[mu]], xm[b], acc);
  }
  const double2 xs = in[tid];
  double2 v;
  v.x = c * (acc.x - 6.0 * xs.x) + a * xs.x;
  v.y = c * (acc.y - 6.0 * xs.y) + a * xs.y;
  if (beta != 0.0) {
    const double2 o = old[tid];
    v.x = alpha * v.x + beta * o.x;
    v.y = alpha * v.y + beta * o.y;
  } else {
    v.x *= alpha;
    v.y *= alpha;
  }
  out[tid] = v;
}

hipError_t launch_cheb_step64(const double2* Us, const double2* in, const double2* old,
                              double2* out, int L, double c, double a, double alpha, double beta,
                              hipStream_t st) {
  const int n = 3 * L * L * L;
  k_cheb_step64<<<(n + 255) / 256, 256, 0, st>>>(Us, in, old, out, L, c, a, alpha, beta);
  return hipGetLastError();
}

}  // namespace laph
