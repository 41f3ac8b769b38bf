// Hermitian Gram products of the batched solver's FP64 blocks on the FP64
// matrix cores: for every timeslice and site chunk,
//     C(i, j) = alpha * sum_s A[s][l + i] * conj(B[s][l + j])
// over the 64 x 64 tiles ti <= tj of the leading a x a block (the upper
// triangle and the diagonal tiles whole; C column-major, lda K).  The blocks
// are site-major [n][K] (vectors fastest), so a row of a 64-column panel is a
// contiguous 1 KiB.  This is the Gram Y^H Y and the Rayleigh quotient
// (D Y)^H Y of the Rayleigh-Ritz step with 10 of the 16 tiles of a 256-wide
// block computed, against 3/4 of the full product for two row blocks of
// zgemm.
//
// Work-group = 4 wavefronts on one 64 x 64 output tile of one (timeslice,
// site chunk); wavefront w owns the 32 x 32 quadrant (w / 2, w % 2) as 2 x 2
// tiles of v_mfma_f64_16x16x4f64, real and imaginary parts in separate
// accumulators: with a = A_re + i A_im, b = B_re + i B_im,
//     re += A_re B_re + A_im B_im,   im += A_im B_re - A_re B_im
// (four MFMAs per tile and k-step).  The two 16-row x 64-column panels of an
// iteration pass through one LDS buffer (rows padded by one element; 33 KB,
// so that three work-groups share a CU, as many as the 138 VGPRs allow); the
// next iteration's panels are fetched into registers while the current ones
// feed the matrix cores, and written to LDS between two barriers.  Operand layout of the instruction (measured):
// lane l holds A[i = l % 16][k = l / 16] and B[k = l / 16][j = l % 16];
// output register r of lane l is D[4 r + l / 16][l % 16].
#include <hip/hip_runtime.h>

#include "kernels_block.h"

namespace laph {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int kGramRows = 16;  // site rows per iteration (4 k-steps of 4)
constexpr int kGramPad = 65;   // padded panel row (double2 entries)

__device__ __forceinline__ void tile_of(int p, int T, int& ti, int& tj) {
  ti = 0;
  while (p >= T - ti) {
    p -= T - ti;
    ++ti;
  }
  tj = ti + p;
}

// The 1D grid holds nts * S units (timeslice, site chunk) of NT = T(T+1)/2
// tiles each.  Work-groups go to the 8 XCDs round-robin by linear id; when the
// unit count is a multiple of 8 the id is decoded so that all tiles of a unit
// run on one XCD, side by side, and share their panels through its L2.
__device__ __forceinline__ void gram_unit(int NT, int units, int& unit, int& tile) {
  const int id = blockIdx.x;
  if ((units & 7) == 0) {
    const int j = id >> 3;
    unit = (j / NT) * 8 + (id & 7);
    tile = j % NT;
  } else {
    unit = id / NT;
    tile = id % NT;
  }
}

__global__ __launch_bounds__(256) void k_gram_mfma(const double2* __restrict__ A, const double2* __restrict__ B,
                                                    double2* __restrict__ C, int n, int K, int l, int a,
                                                    int rows_chunk, int S, int nts, size_t per_ts, double alpha) {
  __shared__ double2 pa[kGramRows][kGramPad];
  __shared__ double2 pb[kGramRows][kGramPad];
  const int T = a / 64;
  int unit, tile, ti, tj;
  gram_unit(T * (T + 1) / 2, nts * S, unit, tile);
  tile_of(tile, T, ti, tj);
  const int ts = unit / S, chunk = unit % S;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int qi = wave >> 1, qj = wave & 1;
  const size_t row0 = (size_t)chunk * rows_chunk;
  const double2* Ab = A + (size_t)ts * per_ts + row0 * K + l + ti * 64;
  const double2* Bb = B + (size_t)ts * per_ts + row0 * K + l + tj * 64;
  // panel fetch: 1024 entries per panel, 4 per thread, 64 threads per row
  const int fr = tid >> 6, fc = tid & 63;  // entry e of a thread: row fr + 4 e, column fc
  double2 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
#define LAPH_GRAM_FETCH(it)                                         \
  do {                                                              \
    const size_t o_ = ((size_t)(it) * kGramRows + fr) * K + fc;     \
    ra0 = Ab[o_];                                                   \
    ra1 = Ab[o_ + 4 * (size_t)K];                                   \
    ra2 = Ab[o_ + 8 * (size_t)K];                                   \
    ra3 = Ab[o_ + 12 * (size_t)K];                                  \
    rb0 = Bb[o_];                                                   \
    rb1 = Bb[o_ + 4 * (size_t)K];                                   \
    rb2 = Bb[o_ + 8 * (size_t)K];                                   \
    rb3 = Bb[o_ + 12 * (size_t)K];                                  \
  } while (0)
#define LAPH_GRAM_STASH()                                           \
  do {                                                              \
    pa[fr][fc] = ra0;                                               \
    pa[fr + 4][fc] = ra1;                                           \
    pa[fr + 8][fc] = ra2;                                           \
    pa[fr + 12][fc] = ra3;                                          \
    pb[fr][fc] = rb0;                                               \
    pb[fr + 4][fc] = rb1;                                           \
    pb[fr + 8][fc] = rb2;                                           \
    pb[fr + 12][fc] = rb3;                                          \
  } while (0)
  d4v re00 = {0, 0, 0, 0}, re01 = re00, re10 = re00, re11 = re00;
  d4v im00 = re00, im01 = re00, im10 = re00, im11 = re00;
  const int nit = rows_chunk / kGramRows;
  LAPH_GRAM_FETCH(0);
  LAPH_GRAM_STASH();
  __syncthreads();
  const int li = lane & 15, lk = lane >> 4;
  const int ci = qi * 32 + li, cj = qj * 32 + li;
  for (int it = 0; it < nit; ++it) {
    if (it + 1 < nit) LAPH_GRAM_FETCH(it + 1);
#pragma unroll
    for (int kk = 0; kk < kGramRows / 4; ++kk) {
      const int k = kk * 4 + lk;
      const double2 a0 = pa[k][ci], a1 = pa[k][ci + 16];
      const double2 b0 = pb[k][cj], b1 = pb[k][cj + 16];
      // the eight accumulators take their first product, then their second:
      // no MFMA waits on the one just issued
      re00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, b0.x, re00, 0, 0, 0);
      re01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, b1.x, re01, 0, 0, 0);
      re10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, b0.x, re10, 0, 0, 0);
      re11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, b1.x, re11, 0, 0, 0);
      im00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b0.x, im00, 0, 0, 0);
      im01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b1.x, im01, 0, 0, 0);
      im10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b0.x, im10, 0, 0, 0);
      im11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b1.x, im11, 0, 0, 0);
      re00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b0.y, re00, 0, 0, 0);
      re01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, b1.y, re01, 0, 0, 0);
      re10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b0.y, re10, 0, 0, 0);
      re11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, b1.y, re11, 0, 0, 0);
      im00 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a0.x, b0.y, im00, 0, 0, 0);
      im01 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a0.x, b1.y, im01, 0, 0, 0);
      im10 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a1.x, b0.y, im10, 0, 0, 0);
      im11 = __builtin_amdgcn_mfma_f64_16x16x4f64(-a1.x, b1.y, im11, 0, 0, 0);
    }
    __syncthreads();  // every wavefront has read the panels
    if (it + 1 < nit) LAPH_GRAM_STASH();
    __syncthreads();
  }
#undef LAPH_GRAM_FETCH
#undef LAPH_GRAM_STASH
  const d4v acc_re[2][2] = {{re00, re01}, {re10, re11}}, acc_im[2][2] = {{im00, im01}, {im10, im11}};
  double2* Cb = C + (size_t)unit * K * K;
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = ti * 64 + qi * 32 + x * 16 + 4 * r + lk;
        const int j = tj * 64 + qj * 32 + y * 16 + li;
        Cb[(size_t)j * K + i] = make_double2(alpha * acc_re[x][y][r], alpha * acc_im[x][y][r]);
      }
}

bool gram_mfma_fits(int n, int a, int S) {
  return a % 64 == 0 && S >= 1 && n % S == 0 && (n / S) % kGramRows == 0;
}

hipError_t launch_gram_mfma(const double2* A, const double2* B, double2* C, int n, int K, int l, int a, int nts,
                            int S, size_t per_ts, double alpha, hipStream_t st) {
  if (!gram_mfma_fits(n, a, S)) return hipErrorInvalidValue;
  const int T = a / 64;
  const unsigned g = (unsigned)(T * (T + 1) / 2) * (unsigned)(nts * S);
  k_gram_mfma<<<g, 256, 0, st>>>(A, B, C, n, K, l, a, n / S, S, nts, per_ts, alpha);
  return hipGetLastError();
}

}  // namespace laph
