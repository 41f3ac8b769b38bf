// Source shape psi(r) of a timeslice's eigensystem (sourceshape_funcs.cpp:31-81):
//   psi(r) = sum_x || V_{x+r} V_x^H ||_F ,
// V_x the 3 x nev block of the eigenvectors at site x, r on the grid
// {0, 2, 4, ..} ^ 3 below L (the loops of sourceshape_funcs.cpp:57-60), site
// index x*L^2 + y*L + z, periodic.  The reference's value is
// sqrt(tr(M M^H)) of M = V_{x+r} V_x^H, i.e. the Frobenius norm.
//
// One thread per (site, r): it accumulates the nine entries of M over the
// nev eigenvectors, reading V in the file layout (eigenvector k is the
// contiguous column V[k*n + 3*site + colour]) so that consecutive threads
// (consecutive sites) read consecutive 48-byte colour triplets.  The
// work-group sums its norms in LDS and adds them to psi[r] atomically.
#include <hip/hip_runtime.h>

#include <vector>

#include "batch_internal.h"
#include "common.h"
#include "kernels_block.h"

namespace laph {

__global__ __launch_bounds__(256) void k_source_shape(const double2* __restrict__ V, int L, int nev, int half,
                                                      double* __restrict__ psi) {
  const int V3 = L * L * L;
  const size_t n = 3 * (size_t)V3;
  const int ir = blockIdx.y;
  const int r1 = 2 * (ir / (half * half)), r2 = 2 * ((ir / half) % half), r3 = 2 * (ir % half);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0.0;
  if (s < V3) {
    const int x = s / (L * L), y = (s / L) % L, z = s % L;
    const int d = (((x + r1) % L) * L + (y + r2) % L) * L + (z + r3) % L;
    double2 m[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) m[i] = make_double2(0.0, 0.0);
    for (int k = 0; k < nev; ++k) {
      const double2* col = V + (size_t)k * n;
      double2 a[3], b[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        a[c] = col[3 * (size_t)d + c];  // V_{x+r}
        b[c] = col[3 * (size_t)s + c];  // V_x
      }
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {  // m_ij += a_i conj(b_j)
          m[3 * i + j].x = fma(a[i].x, b[j].x, fma(a[i].y, b[j].y, m[3 * i + j].x));
          m[3 * i + j].y = fma(a[i].y, b[j].x, fma(-a[i].x, b[j].y, m[3 * i + j].y));
        }
    }
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) f += m[i].x * m[i].x + m[i].y * m[i].y;
    nrm = sqrt(f);
  }
  __shared__ double part[256];
  part[threadIdx.x] = nrm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(psi + ir, part[0]);
}

hipError_t launch_source_shape(const double2* V, int L, int nev, double* psi, hipStream_t st) {
  const int V3 = L * L * L;
  const int half = (L + 1) / 2;
  const int nr = half * half * half;
  hipError_t e = hipMemsetAsync(psi, 0, nr * sizeof(double), st);
  if (e != hipSuccess) return e;
  dim3 g((unsigned)((V3 + 255) / 256), (unsigned)nr);
  k_source_shape<<<g, 256, 0, st>>>(V, L, nev, half, psi);
  return hipGetLastError();
}

}  // namespace laph

extern "C" {

int laph_source_shape(int L, int nev, const double* V, int device, double* psi) {
  LAPH_CHECK(L >= 1 && L <= 128 && nev >= 1 && V && psi, LAPH_ERR_INVALID, "laph_source_shape: bad arguments");
  LAPH_HIP(hipSetDevice(device));
  const size_t n = 3 * (size_t)L * L * L;
  const int half = (L + 1) / 2, nr = half * half * half;
  double2* dV = nullptr;
  double* dpsi = nullptr;
  hipError_t e = hipMalloc(&dV, n * nev * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&dpsi, nr * sizeof(double));
  if (e == hipSuccess) e = hipMemcpy(dV, V, n * nev * sizeof(double2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = laph::launch_source_shape(dV, L, nev, dpsi, nullptr);
  if (e == hipSuccess) e = hipMemcpy(psi, dpsi, nr * sizeof(double), hipMemcpyDeviceToHost);
  if (dV) (void)hipFree(dV);
  if (dpsi) (void)hipFree(dpsi);
  LAPH_HIP(e);
  return LAPH_OK;
}

int laph_batch_source_shape(laph_batch* b, int ts, double* psi) {
  LAPH_CHECK(b && ts >= 0 && ts < b->nts && psi, LAPH_ERR_INVALID, "laph_batch_source_shape: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_source_shape: batch not solved");
  LAPH_HIP(hipSetDevice(b->p.device));
  return laph::guarded([&]() -> int {
    LAPH_TRY(laph::output_buffers(b));
    const int L = b->p.L, half = (L + 1) / 2, nr = half * half * half;
    LAPH_HIP(laph::launch_extract_cols(b->B[0] + (size_t)ts * b->per_ts, b->wV, b->n, b->K, b->nev, b->st));
    double* dpsi = nullptr;
    LAPH_HIP(hipMalloc(&dpsi, nr * sizeof(double)));
    hipError_t e = laph::launch_source_shape(b->wV, L, b->nev, dpsi, b->st);
    if (e == hipSuccess) e = hipMemcpyAsync(psi, dpsi, nr * sizeof(double), hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    (void)hipFree(dpsi);
    LAPH_HIP(e);
    return LAPH_OK;
  });
}

}  // extern "C"
