// FP64 single-vector covariant 3D Laplacian fused with one Chebyshev
// recurrence step -- the plain drop-in of tv2/tv_iter
// (shell_matop.cpp:16-103, 136-168):
//   (B x)(s) = c * (sum_mu [U_mu(s) x(s+mu) + U_mu(s-mu)^H x(s-mu)] - 6 x(s)) + a x(s)
//   out      = alpha * (B in) + beta * old          (alpha,beta) = (1,0) or (2,-1)
// One thread per (site, colour row): 3*V3 threads, each row reads its three
// link rows and the six neighbour colour vectors.  Links either SoA
// Us[(dir*9+e)*V3 + s] (coalesced over sites, the context's packed copy) or
// the AoS host layout U[(s*3+dir)*9 + e] (read in place, no packed copy).
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

// Element e = 3*row + col of U_mu(s) in either link layout.
template <bool AOS>
__device__ __forceinline__ double2 link(const double2* __restrict__ U, int mu, int e, size_t s, int V3) {
  return AOS ? U[(s * 3 + mu) * 9 + e] : U[(size_t)(mu * 9 + e) * V3 + s];
}

template <bool AOS>
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
    const double2* xu = in + 3 * (size_t)up[mu];
    const double2* xm = in + 3 * (size_t)dn[mu];
    // U_mu(s) row r . x(s+mu)
#pragma unroll
    for (int b = 0; b < 3; ++b) acc = fma_c(link<AOS>(Us, mu, 3 * r + b, s, V3), xu[b], acc);
    // (U_mu(s-mu)^H)(r, b) = conj(U_mu(s-mu)(b, r))
#pragma unroll
    for (int b = 0; b < 3; ++b) acc = fma_cc(link<AOS>(Us, mu, 3 * b + r, dn[mu], V3), xm[b], acc);
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
  k_cheb_step64<false><<<(n + 255) / 256, 256, 0, st>>>(Us, in, old, out, L, c, a, alpha, beta);
  return hipGetLastError();
}

hipError_t launch_cheb_step64_aos(const double2* U, const double2* in, const double2* old, double2* out, int L,
                                  double c, double a, double alpha, double beta, hipStream_t st) {
  const int n = 3 * L * L * L;
  k_cheb_step64<true><<<(n + 255) / 256, 256, 0, st>>>(U, in, old, out, L, c, a, alpha, beta);
  return hipGetLastError();
}

}  // namespace laph
