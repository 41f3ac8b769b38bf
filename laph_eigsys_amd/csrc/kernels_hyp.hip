// HYP smearing of the spatial links of one timeslice (3D HYP,
// timeslice.cpp:243-369), site-parallel in FP64, plus the link repacking
// into the SoA layouts the stencils read.
//
// Level 1 (timeslice.cpp:263-306): one thread per (site, dir, plane) builds
//   Vbar_{dir;rho}(x) = P[(1-a2) U_dir(x) + a2/2 (U_rho(x) U_dir(x+rho) U_rho^H(x+dir)
//                        + U_rho^H(x-rho) U_dir(x-rho) U_rho(x-rho+dir))]
// and stores it in slot (rho == first perpendicular ? dir : dir+3), the
// reference's decor() map.
// Level 2 (timeslice.cpp:311-351): one thread per (site, dir) sums the four
// outer staples of the decorated links and projects
//   Unew = P[(1-a1) U + a1/4 * sum].
// The update is Jacobi-style (all of an iteration reads the old links).
#include "kernels.h"
#include "su3.h"

namespace laph {

__device__ __forceinline__ int nb_up(int s, int dir, int L) {
  const int L2 = L * L;
  if (dir == 0) { int x = s / L2; return x == L - 1 ? s - (L - 1) * L2 : s + L2; }
  if (dir == 1) { int y = (s / L) % L; return y == L - 1 ? s - (L - 1) * L : s + L; }
  int z = s % L;
  return z == L - 1 ? s - (L - 1) : s + 1;
}
__device__ __forceinline__ int nb_dn(int s, int dir, int L) {
  const int L2 = L * L;
  if (dir == 0) { int x = s / L2; return x == 0 ? s + (L - 1) * L2 : s - L2; }
  if (dir == 1) { int y = (s / L) % L; return y == 0 ? s + (L - 1) * L : s - L; }
  int z = s % L;
  return z == 0 ? s + (L - 1) : s - 1;
}

// decor(dir, plane) of timeslice.cpp:125-140
__device__ __forceinline__ int decor_slot(int dir, int plane) {
  if (dir == 0) return plane == 1 ? 0 : 3;
  if (dir == 1) return plane == 0 ? 1 : 4;
  return plane == 0 ? 2 : 5;
}

__global__ __launch_bounds__(128) void k_hyp_level1(const double2* __restrict__ U,
                                                     double2* __restrict__ dec, int L, double a2) {
  const int V3 = L * L * L;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= V3 * 6) return;
  const int s = tid / 6, slot = tid % 6;
  const int dir = slot % 3, n_el = slot / 3;
  // get_dirs (timeslice.cpp:106-122)
  const int rho = (dir == 0) ? (n_el == 0 ? 1 : 2) : (dir == 1 ? (n_el == 0 ? 0 : 2) : (n_el == 0 ? 0 : 1));
  const int xr = nb_up(s, rho, L), xd = nb_up(s, dir, L);
  const int xmr = nb_dn(s, rho, L), xmrd = nb_up(xmr, dir, L);
  M3 st = mul(load_m3(U + ((size_t)s * 3 + rho) * 9),
              mul_dag(load_m3(U + ((size_t)xr * 3 + dir) * 9), load_m3(U + ((size_t)xd * 3 + rho) * 9)));
  M3 dn = dag_mul(load_m3(U + ((size_t)xmr * 3 + rho) * 9),
                  mul(load_m3(U + ((size_t)xmr * 3 + dir) * 9), load_m3(U + ((size_t)xmrd * 3 + rho) * 9)));
  M3 u = load_m3(U + ((size_t)s * 3 + dir) * 9);
  const double h = a2 / 2.0;
#pragma unroll
  for (int i = 0; i < 9; ++i)
    u.e[i] = make_double2(u.e[i].x * (1.0 - a2) + (st.e[i].x + dn.e[i].x) * h,
                          u.e[i].y * (1.0 - a2) + (st.e[i].y + dn.e[i].y) * h);
  proj_su3(u);
  store_m3(dec + ((size_t)s * 6 + (n_el == 0 ? dir : dir + 3)) * 9, u);
}

__global__ __launch_bounds__(128) void k_hyp_level2(const double2* __restrict__ U,
                                                     const double2* __restrict__ dec,
                                                     double2* __restrict__ Unew, int L, double a1) {
  const int V3 = L * L * L;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= V3 * 3) return;
  const int i = tid / 3, dir = tid % 3;
  M3 outer;
#pragma unroll
  for (int k = 0; k < 9; ++k) outer.e[k] = make_double2(0.0, 0.0);
  for (int nd = 0; nd < 3; ++nd) {
    if (nd == dir) continue;
    const int plane = ((dir + 1) ^ (nd + 1)) - 1;
    const int a = decor_slot(nd, plane), b = decor_slot(dir, plane);
    const int xn = nb_up(i, nd, L), xd = nb_up(i, dir, L);
    M3 t = mul(load_m3(dec + ((size_t)i * 6 + a) * 9),
               mul_dag(load_m3(dec + ((size_t)xn * 6 + b) * 9), load_m3(dec + ((size_t)xd * 6 + a) * 9)));
#pragma unroll
    for (int k = 0; k < 9; ++k) outer.e[k] = cadd(outer.e[k], t.e[k]);
    const int xm = nb_dn(i, nd, L), xmd = nb_up(xm, dir, L);
    t = dag_mul(load_m3(dec + ((size_t)xm * 6 + a) * 9),
                mul(load_m3(dec + ((size_t)xm * 6 + b) * 9), load_m3(dec + ((size_t)xmd * 6 + a) * 9)));
#pragma unroll
    for (int k = 0; k < 9; ++k) outer.e[k] = cadd(outer.e[k], t.e[k]);
  }
  M3 u = load_m3(U + ((size_t)i * 3 + dir) * 9);
  const double q = a1 / 4.0;
#pragma unroll
  for (int k = 0; k < 9; ++k)
    u.e[k] = make_double2(u.e[k].x * (1.0 - a1) + outer.e[k].x * q, u.e[k].y * (1.0 - a1) + outer.e[k].y * q);
  proj_su3(u);
  store_m3(Unew + ((size_t)i * 3 + dir) * 9, u);
}

__global__ void k_pack64(const double2* __restrict__ U, double2* __restrict__ Us, int V3) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V3) return;
  for (int de = 0; de < 27; ++de) Us[(size_t)de * V3 + s] = U[(size_t)s * 27 + de];
}

__global__ void k_pack32(const double2* __restrict__ U, float2* __restrict__ Us, int V3) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V3) return;
  for (int de = 0; de < 27; ++de) {
    double2 u = U[(size_t)s * 27 + de];
    Us[(size_t)de * V3 + s] = make_float2((float)u.x, (float)u.y);
  }
}

hipError_t launch_hyp_level1(const double2* U, double2* dec, int L, double alpha2, hipStream_t st) {
  const int n = L * L * L * 6;
  k_hyp_level1<<<(n + 127) / 128, 128, 0, st>>>(U, dec, L, alpha2);
  return hipGetLastError();
}

hipError_t launch_hyp_level2(const double2* U, const double2* dec, double2* Unew, int L,
                             double alpha1, hipStream_t st) {
  const int n = L * L * L * 3;
  k_hyp_level2<<<(n + 127) / 128, 128, 0, st>>>(U, dec, Unew, L, alpha1);
  return hipGetLastError();
}

hipError_t launch_pack_links64(const double2* U, double2* Us, int L, hipStream_t st) {
  const int V3 = L * L * L;
  k_pack64<<<(V3 + 255) / 256, 256, 0, st>>>(U, Us, V3);
  return hipGetLastError();
}

hipError_t launch_pack_links32(const double2* U, float2* Us, int L, hipStream_t st) {
  const int V3 = L * L * L;
  k_pack32<<<(V3 + 255) / 256, 256, 0, st>>>(U, Us, V3);
  return hipGetLastError();
}

}  // namespace laph
