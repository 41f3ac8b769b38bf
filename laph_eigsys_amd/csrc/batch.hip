// Mixed-precision batched eigensolver: every timeslice of a configuration on
// one device, in lock-step, each with its own window and lock count.
//
// Method: Chebyshev-filtered subspace iteration (Zhou & Saad, J. Comput.
// Phys. 219 (2006)) on D with a K-column block per timeslice (K >= nev+guard,
// multiple of 64), in residual-correction form so the FP32 filter acts on a
// shrinking quantity.  For a Ritz pair (x, theta) of B = c D + a with
// r = (B - theta) x, the degree-m Chebyshev filter satisfies exactly
//     T_m(B) x = T_m(theta) x + e_m,   e_m = q(B) r,
//     q(y) = (T_m(y) - T_m(theta)) / (y - theta),
// so only e (of size |r|) is computed in FP32; x stays FP64 and the FP32
// rounding error scales with the residual.  Both polynomials are applied in
// product form, one factor a1 (B - z) per step (a complex shift for q when
// theta > 1), so a step reads one vector and writes one; the columns are
// carried scaled to stay O(|r|).  T_m(B) is applied P times per cycle (a
// correction pass, then plain passes that add gamma^(p-1) f_m in their last
// step), with P cut so that T_m(theta_max)^P stays below the amplification cap.
//
// Per cycle (every timeslice of a lane batched; one lane by default, more
// host threads and streams with LAPH_LANES):
//   1. FP32 filter (hipGraphs, one launch per step covering every active
//      vector block of every timeslice)
//   2. Y = gamma^P X + W (FP64), projection against the hard-locked columns
//   3. AY = D Y (FP64 blocked stencil); Gram G = Y^H Y and H = Y^H AY
//   4. Rayleigh-Ritz in FP64 on the non-orthonormal block: Cholesky of the
//      equilibrated Gram, its explicit inverse, zheevd, X = Y U^-1 Q
//      (CholQR2 and shifted CholQR as fallbacks)
//   5. r = c_t (D X + X Lambda) from the FP64 stencil of the rotated block and
//      its column norms; lock the leading columns with |R_j| <= tol; window
//      lower bound = largest Ritz value of -D.
// The first cycle starts from the seeded Gaussian block itself (every column
// at the window edge), without a Rayleigh-Ritz step.
// The window map is B = c_t D + a_t with c_t = 2/(lambda_l - lambda_c,t),
// a_t = 1 + c_t lambda_c,t (shell_matop.cpp:49-50) and lambda_c,t adapted per
// timeslice; eigenvalues come straight from the FP64 Rayleigh-Ritz, so
// recover_spectrum's inversion is not needed on this path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "kernels_block.h"

#include "batch_internal.h"

#include <cstdio>
#include <cstdlib>
#include <thread>

namespace laph {

// Candidate block widths of the filter buffers: {32, 64, 128} for FP32
// storage, {128, 256} for BF16 (the narrowest divides every K).
static std::vector<int> block_widths(const laph_batch* b) {
  if (b->bf16) return {128, 256};
  return {32, 64, 128};
}

// Filter buffer `base` at entry offset `off` (entries of esz bytes).
static inline void* fb(const laph_batch* b, void* base, size_t off) {
  return static_cast<char*>(base) + off * b->esz;
}

static int batch_alloc(laph_batch* b) {
  const size_t lk = (size_t)b->V3 * 27;
  const size_t nb = (size_t)b->nts * b->per_ts;
  LAPH_HIP(hipMalloc(&b->U64, lk * b->nts * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->U32, lk * b->nts * sizeof(float2)));
  LAPH_HIP(hipMalloc(&b->U32c, lk * b->nts * sizeof(float2)));
  LAPH_HIP(hipMemset(b->U32c, 0, lk * b->nts * sizeof(float2)));
  for (int i = 0; i < 3; ++i) LAPH_HIP(hipMalloc(&b->B[i], nb * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->R32, nb * b->esz));
  for (int i = 0; i < 3; ++i) LAPH_HIP(hipMalloc(&b->P[i], nb * b->esz));
  const size_t kk = (size_t)b->nts * b->K * b->K;
  LAPH_HIP(hipMalloc(&b->G, kk * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->G2, kk * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->P2, kk * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->G3, kk * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->Gpart, kk * b->gram_ksplit_max * sizeof(double2)));
  LAPH_HIP(hipMemset(b->G3, 0, kk * sizeof(double2)));
  const size_t nk = (size_t)b->nts * b->K;
  LAPH_HIP(hipMalloc(&b->mu, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->E, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->info, b->nts * sizeof(rocblas_int)));
  LAPH_HIP(hipMalloc(&b->nrm2, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->gam, nk * sizeof(double)));
  const int csteps = std::max(1, std::max(b->deg, b->fdeg_max) - 1);  // correction steps of the largest degree
  LAPH_HIP(hipMalloc(&b->coef, (size_t)csteps * nk * sizeof(float4)));
  LAPH_HIP(hipMalloc(&b->coef2, (size_t)(b->deg + LAPH_MAX_PASSES - 2) * nk * sizeof(float4)));
  LAPH_HIP(hipMalloc(&b->cab, b->nts * sizeof(float2)));
  LAPH_HIP(hipMalloc(&b->first_blk, b->nts * sizeof(int)));
  LAPH_HIP(hipMalloc(&b->first_blk_steps, (size_t)csteps * b->nts * sizeof(int)));
  LAPH_HIP(hipMalloc(&b->lock_blk, b->nts * sizeof(int)));
  LAPH_HIP(hipMalloc(&b->cn, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->red, (size_t)red_chunks(b->n) * nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->phase, (size_t)b->nts * b->nev * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->jsweeps, b->nts * sizeof(rocblas_int)));
  // defined contents for the width autotuner's first replays
  LAPH_HIP(hipMemset(b->coef, 0, (size_t)csteps * nk * sizeof(float4)));
  LAPH_HIP(hipMemset(b->coef2, 0, (size_t)(b->deg + LAPH_MAX_PASSES - 2) * nk * sizeof(float4)));
  LAPH_HIP(hipMemset(b->cab, 0, b->nts * sizeof(float2)));
  LAPH_HIP(hipMemset(b->first_blk, 0, b->nts * sizeof(int)));
  LAPH_HIP(hipMemset(b->first_blk_steps, 0, (size_t)csteps * b->nts * sizeof(int)));
  LAPH_HIP(hipMemset(b->R32, 0, nb * b->esz));
  for (int i = 0; i < 3; ++i) LAPH_HIP(hipMemset(b->P[i], 0, nb * b->esz));
  const std::vector<int> cand = block_widths(b);
  if (const char* e = std::getenv("LAPH_BW")) {
    const int w = std::atoi(e);
    LAPH_CHECK(std::find(cand.begin(), cand.end(), w) != cand.end() && b->K % w == 0, LAPH_ERR_INVALID,
               "LAPH_BW=%s: one of %s dividing K=%d", e, b->bf16 ? "128, 256" : "32, 64, 128", b->K);
    b->bw = w;
  }
  int nlanes = 1;
  if (const char* e = std::getenv("LAPH_LANES")) nlanes = std::max(1, std::min(4, std::atoi(e)));
  nlanes = std::min(nlanes, b->nts);
  b->lanes.resize(nlanes);
  for (int h = 0; h < nlanes; ++h) {
    laph_lane& ln = b->lanes[h];
    ln.t0 = (int)((long long)h * b->nts / nlanes);
    ln.nt = (int)((long long)(h + 1) * b->nts / nlanes) - ln.t0;
    ln.wcap = 0.75 * b->p.lambda_l;
    LAPH_HIP(hipStreamCreateWithFlags(&ln.st, hipStreamNonBlocking));
    LAPH_CHECK(rocblas_create_handle(&ln.rb) == rocblas_status_success &&
                   rocblas_set_stream(ln.rb, ln.st) == rocblas_status_success,
               LAPH_ERR_HIP, "laph_batch_create: rocBLAS handle");
    for (auto& e : ln.ev) LAPH_HIP(hipEventCreate(&e));
  }
  return LAPH_OK;
}

static void batch_free(laph_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->p.device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  for (auto& ln : b->lanes) {
    if (ln.st) (void)hipStreamSynchronize(ln.st);
    if (ln.graph) (void)hipGraphExecDestroy(ln.graph);
    for (auto& kv : ln.gdeg)
      if (kv.second) (void)hipGraphExecDestroy(kv.second);
    for (auto g : ln.gplain)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto e : ln.ev)
      if (e) (void)hipEventDestroy(e);
    if (ln.rb) rocblas_destroy_handle(ln.rb);
    if (ln.st) (void)hipStreamDestroy(ln.st);
  }
  for (auto e : b->wev)
    if (e) (void)hipEventDestroy(e);
  for (void* h : b->pin)
    if (h) (void)hipHostFree(h);
  if (b->wrb) rocblas_destroy_handle(b->wrb);
  for (void* q : {(void*)b->wV, (void*)b->wG})
    if (q) (void)hipFree(q);
  void* bufs[] = {b->U64, b->U32, b->U32c, b->dec,  b->Unew, b->B[0], b->B[1],   b->B[2], b->R32,     b->P[0],
                  b->P[1], b->P[2], b->G,  b->G2,   b->P2,   b->G3,   b->mu,     b->E,    b->info,    b->nrm2,
                  b->gam, b->coef, b->cab, b->first_blk, b->first_blk_steps, b->lock_blk, b->cn, b->coef2,
                  b->jsweeps,
                  b->red, b->phase, b->Gpart};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

using cplx = rocblas_double_complex;
static inline cplx* zc(double2* p) { return reinterpret_cast<cplx*>(p); }
static inline cplx* zc(const double2* p) { return reinterpret_cast<cplx*>(const_cast<double2*>(p)); }

// The dense steps act on the active block: columns [l, K) of every
// timeslice of the lane, l = ln.lock (a multiple of 64; columns below l are
// hard-locked converged eigenvectors).  In the column-major view the block of
// one timeslice is Yc = K x n with leading dimension K (the site-major
// [n][K] storage), so the active block is the sub-matrix starting at row l.
struct Act {
  int l, a;  // first active column, active width
};

// Lane-local views of the batch arrays.
struct View {
  laph_batch* b;
  laph_lane* ln;
  double2* blk(int i) const { return b->B[i] + (size_t)ln->t0 * b->per_ts; }
  double2* X() const { return blk(ln->iX); }
  double2* AX() const { return blk(ln->iAX); }
  double2* T() const { return blk(ln->iT); }
  size_t kk() const { return (size_t)ln->t0 * b->K * b->K; }
  size_t nk() const { return (size_t)ln->t0 * b->K; }
  size_t vo() const { return (size_t)ln->t0 * b->per_ts; }
  size_t lo() const { return (size_t)ln->t0 * b->V3 * 27; }
  double* red() const { return b->red + (size_t)ln->t0 * red_chunks(b->n) * b->K; }
  int m() const { return ln->m > 0 ? ln->m : b->deg; }  // filter degree of the current cycle
};

// LAPH_TRACE: smallest squared diagonal entry of the Cholesky factors
// (upper triangles in place, lda K) of the lane's timeslices -- with unit
// Gram diagonals a bound above each Gram's smallest eigenvalue -- and its
// timeslice, into ln->min_pivot2 / min_pivot_ts.
static int trace_pivots(const View& v, const double2* F, int a) {
  const int K = v.b->K, nt = v.ln->nt;
  std::vector<double2> d((size_t)a * nt);
  for (int t = 0; t < nt; ++t)
    LAPH_HIP(hipMemcpy2DAsync(d.data() + (size_t)t * a, sizeof(double2), F + (size_t)t * K * K,
                              (size_t)(K + 1) * sizeof(double2), sizeof(double2), a, hipMemcpyDeviceToHost,
                              v.ln->st));
  LAPH_HIP(hipStreamSynchronize(v.ln->st));
  double mp = INFINITY;
  int arg = 0;
  for (size_t i = 0; i < d.size(); ++i)
    if (d[i].x * d[i].x < mp) mp = d[i].x * d[i].x, arg = (int)(i / a);
  if (mp < v.ln->min_pivot2) {  // smallest over the factorizations of the cycle
    v.ln->min_pivot2 = mp;
    v.ln->min_pivot_ts = v.ln->t0 + arg;
  }
  return LAPH_OK;
}

static int dense_sync_info(const View& v, const char* what) {
  std::vector<rocblas_int> info(v.ln->nt);
  LAPH_HIP(hipMemcpyAsync(info.data(), v.b->info + v.ln->t0, v.ln->nt * sizeof(rocblas_int), hipMemcpyDeviceToHost,
                          v.ln->st));
  LAPH_HIP(hipStreamSynchronize(v.ln->st));
  for (int t = 0; t < v.ln->nt; ++t)
    if (info[t] != 0)
      return set_error(LAPH_ERR_NOCONV, "%s failed on timeslice %d (info %d)", what, v.ln->t0 + t, (int)info[t]);
  return LAPH_OK;
}

// Row blocks of the upper triangle of an a x a Hermitian product: `split`
// blocks of whole 64-column groups (the last one may be narrower).
static int gram_rows(int a, int split) {
  const int g = a / 64;
  return 64 * ((g + split - 1) / split);
}

// Gt = alpha * A_c B_c^H (a x a, column-major, lda K) for the active blocks,
// when the product is Hermitian (A = B, or B^H A Hermitian as for a Rayleigh
// quotient): its upper triangle only, as `gram_split` block rows
// [r0, r0+bs) x [r0, a) of strided-batched zgemm (2 row blocks of 256
// columns: 3/4 of the full product's flops at the same rate, 35.0 against
// 47.0 ms per Gram of 64 timeslices at 32^3; 4 row blocks of 64 rows fall to
// 56% of the FP64 matrix peak).  gram_split = 1 computes the full product.
// Entries strictly below the diagonal blocks are left as they were.
// With gram_ksplit = S > 1 the site sum is cut into S chunks that enter the
// same strided-batched call as extra batch entries (timeslice t, chunk c ->
// entry t*S + c: the chunks of n/S rows are (n/S)*K apart, the timeslices
// n*K = S*(n/S)*K apart, one stride for both), each writing its partial
// product into Gpart; k_sum_gram_chunks adds the S partials in chunk order.
// That multiplies the output tiles of a call by S when a lane holds few
// timeslices (24 per group at 48^3).
__global__ void k_sum_gram_chunks(const double2* __restrict__ P, double2* __restrict__ G, int a, int K, int S,
                                  int bs) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a * a) return;
  const int i = idx % a, j = idx / a;
  if (j < (i / bs) * bs) return;  // below the computed row blocks
  const double2* p = P + (size_t)blockIdx.y * S * K * K + (size_t)j * K + i;
  double2 acc = make_double2(0.0, 0.0);
  for (int c = 0; c < S; ++c) {
    const double2 u = p[(size_t)c * K * K];
    acc.x += u.x;
    acc.y += u.y;
  }
  G[(size_t)blockIdx.y * K * K + (size_t)j * K + i] = acc;
}

static int gram(const View& v, Act q, const double2* A, const double2* Bm, double2* Gt, double alpha) {
  const int K = v.b->K;
  const cplx al(alpha, 0.0), zero(0.0, 0.0);
  const int S = v.b->gram_ksplit;
  double2* C = S > 1 ? v.b->Gpart + (size_t)v.ln->t0 * S * K * K : Gt;
  if (v.b->gram_mfma && gram_mfma_fits(v.b->n, q.a, S)) {  // hand-written: the 64 x 64 upper tiles
    LAPH_HIP(launch_gram_mfma(A, Bm, C, v.b->n, K, q.l, q.a, v.ln->nt, S, v.b->per_ts, alpha, v.ln->st));
    if (S > 1) {
      k_sum_gram_chunks<<<dim3((unsigned)((q.a * q.a + 255) / 256), v.ln->nt), 256, 0, v.ln->st>>>(C, Gt, q.a, K, S,
                                                                                                     64);
      LAPH_HIP(hipGetLastError());
    }
    return LAPH_OK;
  }
  const int bs = gram_rows(q.a, std::max(1, v.b->gram_split));
  const int rows = v.b->n / S;  // S divides n
  for (int r0 = 0; r0 < q.a; r0 += bs) {
    const int m = std::min(bs, q.a - r0);
    LAPH_RB(rocblas_zgemm_strided_batched(v.ln->rb, rocblas_operation_none, rocblas_operation_conjugate_transpose, m,
                                          q.a - r0, rows, &al, zc(A + q.l + r0), K, (rocblas_stride)rows * K,
                                          zc(Bm + q.l + r0), K, (rocblas_stride)rows * K, &zero,
                                          zc(C + r0 + (size_t)r0 * K), K, (rocblas_stride)K * K, v.ln->nt * S));
  }
  if (S > 1) {
    k_sum_gram_chunks<<<dim3((unsigned)((q.a * q.a + 255) / 256), v.ln->nt), 256, 0, v.ln->st>>>(C, Gt, q.a, K, S,
                                                                                                   bs);
    LAPH_HIP(hipGetLastError());
  }
  return LAPH_OK;
}

// Lower triangle of every timeslice's a x a block from its upper triangle:
// G(i, j) = conj(G(j, i)) for i > j (column-major, lda K).
__global__ void k_herm_fill(double2* __restrict__ G, int a, int K) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a * a) return;
  const int i = idx % a, j = idx / a;
  if (i <= j) return;
  double2* M = G + (size_t)blockIdx.y * K * K;
  const double2 u = M[(size_t)i * K + j];
  M[(size_t)j * K + i] = make_double2(u.x, -u.y);
}

// Floating-point operations the Gram of one lane at active width a executes
// (complex multiply-adds of the row blocks computed, 8 flops each).
static double gram_flops(const laph_batch* b, int a, int nt) {
  if (b->gram_mfma && gram_mfma_fits(b->n, a, b->gram_ksplit)) {
    const int T = a / 64;
    return 8.0 * b->n * 64.0 * 64.0 * (T * (T + 1) / 2) * nt;
  }
  const int bs = gram_rows(a, std::max(1, b->gram_split));
  double f = 0;
  for (int r0 = 0; r0 < a; r0 += bs) f += 8.0 * b->n * (double)std::min(bs, a - r0) * (a - r0);
  return f * nt;
}

// Y_A <- Y_A - X_L (X_L^H Y_A): one classical Gram-Schmidt pass against the
// hard-locked eigenvectors (the filtered block is near-orthogonal to them
// already: the residuals that drive the filter are orthogonal to X_L up to
// the locked residual).
static int project_locked(const View& v, Act q) {
  if (q.l == 0) return LAPH_OK;
  const int K = v.b->K;
  const cplx one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
  double2* X = v.X();
  double2* P2 = v.b->P2 + v.kk();
  LAPH_RB(rocblas_zgemm_strided_batched(v.ln->rb, rocblas_operation_none, rocblas_operation_conjugate_transpose, q.a,
                                        q.l, v.b->n, &one, zc(X + q.l), K, (rocblas_stride)v.b->per_ts, zc(X), K,
                                        (rocblas_stride)v.b->per_ts, &zero, zc(P2), K, (rocblas_stride)K * K,
                                        v.ln->nt));
  LAPH_RB(rocblas_zgemm_strided_batched(v.ln->rb, rocblas_operation_none, rocblas_operation_none, q.a, v.b->n, q.l,
                                        &mone, zc(P2), K, (rocblas_stride)K * K, zc(X), K,
                                        (rocblas_stride)v.b->per_ts, &one, zc(X + q.l), K,
                                        (rocblas_stride)v.b->per_ts, v.ln->nt));
  return LAPH_OK;
}

// Explicit CholQR of the active block: Yc_A <- U^{-H} Yc_A, Yc_A Yc_A^H = U^H U.
// G_A += shift * I for every timeslice of the lane (a x a block, lda K).
__global__ void k_shift_diag(double2* __restrict__ G, int a, int K, double shift) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a) G[(size_t)blockIdx.y * K * K + (size_t)i * K + i].x += shift;
}

// One Cholesky QR pass of the active block: Y_A <- Y_A R^{-1}, R^H R = Y_A^H Y_A.
// When the Gram is numerically singular (high-degree filtering can make
// buffer columns nearly parallel to the dominant eigenvectors), the pass is
// done on the shifted Gram Y_A^H Y_A + s I, s = 11 (n a + a (a+1)) u |Y_A|_2^2
// (shifted Cholesky QR: the Cholesky then exists and cond(Y_A R^{-1}) is
// O(s^{-1/2})), followed by a plain pass.
static int cholqr(const View& v, Act q, bool shifted = false) {
  const int K = v.b->K;
  const cplx one(1.0, 0.0);
  double2* Y = v.X();
  double2* G = v.b->G + v.kk();
  if (!shifted && v.b->force_shift) LAPH_TRY(cholqr(v, q, true));  // LAPH_SHIFTED_CHOLQR (tests)
  // unit columns first: the filtered columns differ in norm by many orders
  // of magnitude (the span is unchanged by the scaling); then |Y_A|_2^2 <= a
  LAPH_HIP(launch_normalize_columns(Y, v.b->cn + v.nk(), v.red(), v.b->n, K, v.ln->nt, q.l, v.ln->st));
  LAPH_TRY(gram(v, q, Y, Y, G, 1.0));
  if (shifted) {
    ++v.ln->shifted;
    const double a = q.a, u = 0x1p-53;
    const double sh = 11.0 * ((double)v.b->n * a + a * (a + 1.0)) * u * a;
    k_shift_diag<<<dim3((q.a + 255) / 256, v.ln->nt), 256, 0, v.ln->st>>>(G, q.a, K, sh);
    LAPH_HIP(hipGetLastError());
  }
  LAPH_RB(rocsolver_zpotrf_strided_batched(v.ln->rb, rocblas_fill_upper, q.a, zc(G), K, (rocblas_stride)K * K,
                                           v.b->info + v.ln->t0, v.ln->nt));
  if (dense_sync_info(v, "CholQR (zpotrf)") != LAPH_OK) {
    if (shifted) {  // the shifted Gram is definite for any finite block: report what the block holds
      const std::string msg = laph_last_error();
      std::vector<double> cn((size_t)v.ln->nt * K);
      LAPH_HIP(hipMemcpy(cn.data(), v.b->cn + v.nk(), cn.size() * sizeof(double), hipMemcpyDeviceToHost));
      int bad = 0;
      double lo = INFINITY, hi = 0.0;
      for (int t = 0; t < v.ln->nt; ++t)
        for (int j = q.l; j < K; ++j) {
          const double c = cn[(size_t)t * K + j];
          if (!std::isfinite(c)) ++bad;
          else {
            lo = std::min(lo, c);
            hi = std::max(hi, c);
          }
        }
      return set_error(LAPH_ERR_NOCONV, "%s; shifted pass failed too: %d non-finite column norms, finite ones in "
                       "[%.3g, %.3g]", msg.c_str(), bad, lo, hi);
    }
    LAPH_TRY(cholqr(v, q, true));
    return cholqr(v, q);
  }
  if (v.b->trace && !shifted) LAPH_TRY(trace_pivots(v, G, q.a));  // the unit-column block's Gram
  LAPH_RB(rocblas_ztrsm_strided_batched(v.ln->rb, rocblas_side_left, rocblas_fill_upper,
                                        rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, q.a, v.b->n,
                                        &one, zc(G), K, (rocblas_stride)K * K, zc(Y + q.l), K,
                                        (rocblas_stride)v.b->per_ts, v.ln->nt));
  return LAPH_OK;
}

// Rayleigh-Ritz of the active block Y_A (in X) with AY_A = D Y_A (in AX).
// With Mt = -(AYc_A Yc_A^H) = -conj(H) and, when the block is not
// orthonormalised explicitly, Gt = Yc_A Yc_A^H = conj(G) = U^H U, the
// generalized problem Mt z = mu Gt z becomes the standard Hermitian problem
// U^{-H} Mt U^{-1} w = mu w, z = U^{-1} w.  mu ascending = eigenvalues of -D
// nearest 0 first.  New block: Xc_A = Z^H Yc_A; AX_A = D X_A by the FP64
// blocked stencil (bandwidth-bound, far cheaper than rotating AY_A by Z).
// Equilibration of the generalized Rayleigh-Ritz without a pass over the
// block: s_i = 1 / sqrt(Gt_ii) (the inverse column norms), Gt and Mt scaled
// to S Gt S, S Mt S (the Gram and projection of the unit-column block Y S),
// and the rotation of Y S by Z applied to Y as S Z.
__global__ void k_equil_scale(const double2* __restrict__ Gt, double* __restrict__ sc, int a, int K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a) sc[(size_t)blockIdx.y * K + i] = 1.0 / sqrt(Gt[(size_t)blockIdx.y * K * K + (size_t)i * K + i].x);
}
__global__ void k_equil_apply(double2* __restrict__ A, double2* __restrict__ B, const double* __restrict__ sc, int a,
                              int K) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a * a) return;
  const int i = idx % a, j = idx / a;
  const double* s = sc + (size_t)blockIdx.y * K;
  const double f = s[i] * s[j];
  const size_t e = (size_t)blockIdx.y * K * K + (size_t)j * K + i;
  A[e].x *= f;
  A[e].y *= f;
  B[e].x *= f;
  B[e].y *= f;
}
__global__ void k_scale_rows(double2* __restrict__ Z, const double* __restrict__ sc, int a, int K) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a * a) return;
  const int i = idx % a, j = idx / a;
  const double f = sc[(size_t)blockIdx.y * K + i];
  const size_t e = (size_t)blockIdx.y * K * K + (size_t)j * K + i;
  Z[e].x *= f;
  Z[e].y *= f;
}

// Clears the strictly lower triangle of the a x a leading block (lda K) of
// every timeslice's matrix: the triangular inverse enters full GEMMs.
__global__ void k_zero_lower(double2* __restrict__ W, int a, int K) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a * a) return;
  const int i = idx % a, j = idx / a;
  if (i > j) W[(size_t)blockIdx.y * K * K + (size_t)j * K + i] = make_double2(0.0, 0.0);
}

static int rayleigh_ritz(const View& v, Act q, bool orthonormal, bool* chol_failed) {
  laph_batch* b = v.b;
  const int K = b->K, nt = v.ln->nt;
  const cplx one(1.0, 0.0), zero(0.0, 0.0);
  double2* Y = v.X();
  double2* AY = v.AX();
  double2* T = v.T();
  double2* G = b->G + v.kk();
  double2* G2 = b->G2 + v.kk();
  double2* W = b->G3 + v.kk();  // U^{-1}
  double2* T1 = b->P2 + v.kk();  // free during Rayleigh-Ritz
  rocblas_int* info = b->info + v.ln->t0;
  const rocblas_stride kk = (rocblas_stride)K * K;
  rocblas_handle rb = v.ln->rb;
  *chol_failed = false;
  LAPH_TRY(gram(v, q, AY, Y, G, -1.0));  // Mt (upper triangle)
  double* sc = b->cn + v.nk();
  const dim3 gs((unsigned)((q.a + 255) / 256), nt), ga((unsigned)((q.a * q.a + 255) / 256), nt);
  if (!orthonormal) {
    if (b->gram_split > 1 || b->gram_mfma) {  // Mt (upper triangle) enters full products below
      k_herm_fill<<<ga, 256, 0, v.ln->st>>>(G, q.a, K);
      LAPH_HIP(hipGetLastError());
    }
    LAPH_TRY(gram(v, q, Y, Y, G2, 1.0));  // Gt
    k_equil_scale<<<gs, 256, 0, v.ln->st>>>(G2, sc, q.a, K);
    k_equil_apply<<<ga, 256, 0, v.ln->st>>>(G2, G, sc, q.a, K);
    LAPH_HIP(hipGetLastError());
    LAPH_RB(rocsolver_zpotrf_strided_batched(rb, rocblas_fill_upper, q.a, zc(G2), K, kk, info, nt));
    if (dense_sync_info(v, "Rayleigh-Ritz (zpotrf)") != LAPH_OK) {
      ++v.ln->rr_fallbacks;
      *chol_failed = true;
      return LAPH_OK;
    }
    if (b->trace) LAPH_TRY(trace_pivots(v, G2, q.a));
    if (b->use_trsm) {
      LAPH_RB(rocblas_ztrsm_strided_batched(rb, rocblas_side_left, rocblas_fill_upper,
                                            rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, q.a,
                                            q.a, &one, zc(G2), K, kk, zc(G), K, kk, nt));
      LAPH_RB(rocblas_ztrsm_strided_batched(rb, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                                            rocblas_diagonal_non_unit, q.a, q.a, &one, zc(G2), K, kk, zc(G), K, kk,
                                            nt));
    } else {
      // W = U^{-1} explicitly (a few dispatches), then U^{-H} Mt U^{-1} as two
      // small GEMMs: the triangular solves cost hundreds of tiny dispatches
      LAPH_RB(rocblas_ztrtri_strided_batched(rb, rocblas_fill_upper, rocblas_diagonal_non_unit, q.a, zc(G2), K, kk,
                                             zc(W), K, kk, nt));
      k_zero_lower<<<ga, 256, 0, v.ln->st>>>(W, q.a, K);
      LAPH_HIP(hipGetLastError());
      LAPH_RB(rocblas_zgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_none, q.a, q.a, q.a, &one,
                                            zc(G), K, kk, zc(W), K, kk, &zero, zc(T1), K, kk, nt));
      LAPH_RB(rocblas_zgemm_strided_batched(rb, rocblas_operation_conjugate_transpose, rocblas_operation_none, q.a,
                                            q.a, q.a, &one, zc(W), K, kk, zc(T1), K, kk, &zero, zc(G), K, kk, nt));
    }
  }
  if (b->eig_mode == 1) {
    LAPH_RB(rocsolver_zheevdj_strided_batched(rb, rocblas_evect_original, rocblas_fill_upper, q.a, zc(G), K, kk,
                                              b->mu + v.nk() + q.l, K, info, nt));
  } else if (b->eig_mode == 2) {
    LAPH_RB(rocsolver_zheevj_strided_batched(rb, rocblas_esort_ascending, rocblas_evect_original, rocblas_fill_upper,
                                             q.a, zc(G), K, kk, 0.0, b->E + v.nk(), 100, b->jsweeps + v.ln->t0,
                                             b->mu + v.nk() + q.l, K, info, nt));
  } else {
    LAPH_RB(rocsolver_zheevd_strided_batched(rb, rocblas_evect_original, rocblas_fill_upper, q.a, zc(G), K, kk,
                                             b->mu + v.nk() + q.l, K, b->E + v.nk(), K, info, nt));
  }
  LAPH_TRY(dense_sync_info(v, "Rayleigh-Ritz (zheevd)"));
  double2* Z = G;  // rotation of the active block
  if (!orthonormal) {
    if (b->use_trsm) {
      LAPH_RB(rocblas_ztrsm_strided_batched(rb, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                            rocblas_diagonal_non_unit, q.a, q.a, &one, zc(G2), K, kk, zc(G), K, kk,
                                            nt));
    } else {
      LAPH_RB(rocblas_zgemm_strided_batched(rb, rocblas_operation_none, rocblas_operation_none, q.a, q.a, q.a, &one,
                                            zc(W), K, kk, zc(G), K, kk, &zero, zc(T1), K, kk, nt));
      Z = T1;
    }
    k_scale_rows<<<ga, 256, 0, v.ln->st>>>(Z, sc, q.a, K);
    LAPH_HIP(hipGetLastError());
  }
  LAPH_RB(rocblas_zgemm_strided_batched(rb, rocblas_operation_conjugate_transpose, rocblas_operation_none, q.a, b->n,
                                        q.a, &one, zc(Z), K, kk, zc(Y + q.l), K, (rocblas_stride)b->per_ts, &zero,
                                        zc(T + q.l), K, (rocblas_stride)b->per_ts, nt));
  // residuals of the new Ritz pairs for the next filter pass (D X is not kept:
  // the next Rayleigh-Ritz applies D to the filtered block)
  LAPH_HIP(launch_lap_residual64(b->U64 + v.lo(), T, b->mu + v.nk(), b->p.lambda_l, v.ln->wcap,
                                 fb(b, b->R32, v.vo()), b->p.L, K, nt, b->lock_blk + v.ln->t0, b->strip, b->bw,
                                 b->bf16, stencil_zrun(b), v.ln->st));
  if (v.ln->keep_roles) {
    // dense steps on a run of the lane's timeslices: the others keep their
    // columns in X, so the rotated active columns go back into X and the
    // buffer roles stay
    LAPH_HIP(hipMemcpy2DAsync(Y + q.l, (size_t)K * sizeof(double2), T + q.l, (size_t)K * sizeof(double2),
                              (size_t)q.a * sizeof(double2), (size_t)nt * b->n, hipMemcpyDeviceToDevice, v.ln->st));
  } else {
    // the hard-locked columns of X live in the old X buffer: carry them over
    if (q.l > 0)
      LAPH_HIP(hipMemcpy2DAsync(T, (size_t)K * sizeof(double2), Y, (size_t)K * sizeof(double2),
                                (size_t)q.l * sizeof(double2), (size_t)nt * b->n, hipMemcpyDeviceToDevice,
                                v.ln->st));
    // roles: X <- T, T <- old X buffer
    std::swap(v.ln->iX, v.ln->iT);
  }
  LAPH_HIP(hipMemcpyAsync(b->h_mu.data() + v.nk(), b->mu + v.nk(), (size_t)nt * K * sizeof(double),
                          hipMemcpyDeviceToHost, v.ln->st));
  LAPH_HIP(hipStreamSynchronize(v.ln->st));
  return LAPH_OK;
}

// Orthonormalise (explicitly when asked or when the Gram Cholesky fails),
// apply D and run Rayleigh-Ritz on the active block.
static int orthonormalise_and_rr(const View& v, Act q, bool explicit_qr) {
  laph_batch* b = v.b;
  for (int pass = 0; pass < 2; ++pass) {
    if (explicit_qr || pass == 1) {
      int rc = cholqr(v, q);
      if (rc == LAPH_ERR_NOCONV) {
        // even the shifted pass failed (a block that is not finite): the
        // failing timeslices (non-zero info) restart their active columns
        // from Gaussian vectors; the others keep theirs
        if (b->trace)
          std::fprintf(stderr, "[laph] lane %d: %s; active columns re-seeded\n", v.ln->t0, laph_last_error());
        LAPH_HIP(launch_reseed_cols(v.X(), b->n, b->K, v.ln->nt, q.l, b->info + v.ln->t0,
                                    b->o.seed + 0x5851F42Dull * (++v.ln->recoveries), v.ln->st));
        rc = cholqr(v, q);
      }
      LAPH_TRY(rc);
      LAPH_TRY(cholqr(v, q));
    }
    LAPH_HIP(launch_lap_block64(b->U64 + v.lo(), v.X(), v.AX(), b->p.L, b->K, v.ln->nt, b->lock_blk + v.ln->t0,
                                b->strip, stencil_zrun(b), v.ln->st));
    v.ln->sv64 += 2LL * v.ln->nt * q.a;  // D Y for Rayleigh-Ritz, D X after it
    bool failed = false;
    LAPH_TRY(rayleigh_ritz(v, q, explicit_qr || pass == 1, &failed));
    if (!failed) return LAPH_OK;
  }
  return set_error(LAPH_ERR_NOCONV, "Rayleigh-Ritz: Gram matrix not positive definite after CholQR2");
}

// Lane timeslices with an unconverged wanted pair (those the cycle's filter
// ran on), as runs [r0, r1) of consecutive lane timeslices.
static std::vector<std::pair<int, int>> unconverged_runs(const laph_batch* b, const laph_lane* ln) {
  std::vector<std::pair<int, int>> runs;
  for (int t = 0; t < ln->nt; ++t) {
    if (b->nlock[ln->t0 + t] >= b->nev) continue;
    if (!runs.empty() && runs.back().second == t) ++runs.back().second;
    else runs.emplace_back(t, t + 1);
  }
  return runs;
}

// A strided-batched dense step costs a part independent of the batch count
// (the latency-bound Cholesky, inverse and eigensolver chains, launches) and
// a part per timeslice.  Measured at 32^3: the dense steps of 64 timeslices
// take 76 ms at an active width of 64 columns (158 ms at 256), those of one
// timeslice 18 ms, of two runs of one 34-38 ms.  The fixed part is then worth
// about 8 (at 256 columns) to 18 (at 64) timeslices' share; the rule below
// uses 12.
constexpr int kDenseRunCost = 12;

// Projection against the locked columns and Rayleigh-Ritz of the active
// block.  Once some timeslices of the lane have converged (their filter no
// longer runs, their columns stay as they are), the dense steps run on the
// runs of unconverged timeslices alone when the runs are few enough that the
// fixed part repeated per run costs less than the timeslices skipped.  The
// lane's timeslice range is narrowed to each run in turn: every dense step
// addresses its timeslices through it.
static int dense_steps(const View& v, Act q, bool explicit_qr) {
  laph_batch* b = v.b;
  laph_lane* ln = v.ln;
  const auto runs = unconverged_runs(b, ln);
  int active = 0;
  for (const auto& r : runs) active += r.second - r.first;
  const bool subset = b->dense_subset && !runs.empty() && active < ln->nt &&
                      (int)(runs.size() - 1) * kDenseRunCost < ln->nt - active;
  if (!subset) {
    LAPH_TRY(project_locked(v, q));
    return orthonormalise_and_rr(v, q, explicit_qr);
  }
  ++ln->subset_cycles;
  struct Restore {  // the lane's range and roles mode, also on an exception
    laph_lane* ln;
    int t0, nt;
    ~Restore() {
      ln->t0 = t0;
      ln->nt = nt;
      ln->keep_roles = false;
    }
  } restore{ln, ln->t0, ln->nt};
  ln->keep_roles = true;
  for (const auto& r : runs) {
    ln->t0 = restore.t0 + r.first;
    ln->nt = r.second - r.first;
    LAPH_TRY(project_locked(v, q));
    LAPH_TRY(orthonormalise_and_rr(v, q, explicit_qr));
  }
  return LAPH_OK;
}

// K-th lowest eigenvalue of -D on the free field (every link 1): the levels
// 4 sum_i sin^2(pi k_i / L) of the L^3 momenta, each three-fold (colour).
// HYP-smeared fields lie close above it (at 32^3 x 256 vectors, eps 0.3: 0.304
// free, 0.31 for the smeared field).
double free_level(int L, int K) {
  std::vector<double> s(L), ev;
  ev.reserve((size_t)L * L * L);
  for (int k = 0; k < L; ++k) s[k] = 4.0 * std::pow(std::sin(M_PI * k / L), 2);
  for (int a = 0; a < L; ++a)
    for (int c = 0; c < L; ++c)
      for (int d = 0; d < L; ++d) ev.push_back(s[a] + s[c] + s[d]);
  const size_t idx = std::min((size_t)(K - 1) / 3, ev.size() - 1);
  std::nth_element(ev.begin(), ev.begin() + idx, ev.end());
  return ev[idx];
}

// Start of a solve from the Gaussian block without a Rayleigh-Ritz step: a
// rotation of the block does not change its span and the filter is linear,
// so the first (single-pass) filter cycle can take the columns as they are,
// each with theta = 1, the window edge (T_m(1) = 1, gamma = 1: the filtered
// block is X + e).  Needs D X for the residuals.  The window lower end of
// this first cycle is laph_batch::window0, three times the K-th level of the
// free Laplacian (the largest Ritz value of a random block would sit near the
// top of the spectrum and make the first filter a weak low-pass).
static int start_from_block(const View& v) {
  laph_batch* b = v.b;
  const size_t nk = (size_t)v.ln->nt * b->K;
  std::fill(b->h_mu.begin() + v.nk(), b->h_mu.begin() + v.nk() + nk, b->window0);
  LAPH_HIP(hipMemcpyAsync(b->mu + v.nk(), b->h_mu.data() + v.nk(), nk * sizeof(double), hipMemcpyHostToDevice,
                          v.ln->st));
  LAPH_HIP(launch_lap_residual64(b->U64 + v.lo(), v.X(), b->mu + v.nk(), b->p.lambda_l, v.ln->wcap,
                                 fb(b, b->R32, v.vo()), b->p.L, b->K, v.ln->nt, b->lock_blk + v.ln->t0, b->strip,
                                 b->bw, b->bf16, stencil_zrun(b), v.ln->st));
  v.ln->sv64 += (long long)v.ln->nt * b->K;
  return LAPH_OK;
}

// Window per timeslice from the Ritz values, residuals of the active block,
// lock counts.  Sets *n_active to the number of unconverged timeslices.
static int residuals_and_window(const View& v, std::vector<double>& lc, int* n_active) {
  laph_batch* b = v.b;
  laph_lane* ln = v.ln;
  const int K = b->K, nt = ln->nt;
  const double lamL = b->p.lambda_l;
  for (int t = 0; t < nt; ++t) {
    // lower end of the damped interval: largest Ritz value of -D, kept
    // well below lambda_l so the amplification stays inside FP32 range (the
    // residual kernels derive the same window from the device copy of mu)
    lc[t] = std::min(b->h_mu[v.nk() + (size_t)t * K + K - 1], ln->wcap);
  }
  LAPH_HIP(launch_residual_norms(fb(b, b->R32, v.vo()), b->mu + v.nk(), lamL, ln->wcap, b->nrm2 + v.nk(), v.red(),
                                 b->n, K, nt, ln->lock, b->bw, b->bf16, ln->st));
  LAPH_HIP(hipMemcpyAsync(b->h_res.data() + v.nk(), b->nrm2 + v.nk(), (size_t)nt * K * sizeof(double),
                          hipMemcpyDeviceToHost, ln->st));
  LAPH_HIP(hipStreamSynchronize(ln->st));
  int active = 0, lmin = K;
  for (int t = 0; t < nt; ++t) {
    const double* res = &b->h_res[v.nk() + (size_t)t * K];
    int nl = ln->lock;
    while (nl < K && std::sqrt(res[nl]) <= b->o.tol) ++nl;
    b->nlock[ln->t0 + t] = nl;
    if (nl < b->nev) ++active;
    lmin = std::min(lmin, nl);
  }
  *n_active = active;
  // hard-lock the common converged prefix (64-aligned), never past nev
  if (active > 0) {
    const int l = std::min(lmin, b->nev) / 64 * 64;
    if (l > ln->lock) {
      ln->lock = l;
      std::vector<int> lb(nt, l / 64);
      LAPH_HIP(hipMemcpyAsync(b->lock_blk + ln->t0, lb.data(), nt * sizeof(int), hipMemcpyHostToDevice, ln->st));
      LAPH_HIP(hipStreamSynchronize(ln->st));
    }
  }
  return LAPH_OK;
}

// Per-column coefficients of the filter steps of a cycle of `passes`
// applications of T_m(B), and the combine factor gamma^passes.
// Roots of T_m, z_j = cos((2j-1) pi / 2m), in Leja order: each next root
// maximises the product of its distances to those already taken, starting
// from the largest.  Applied in this order the partial products
// prod_{j<=k} beta_j (x - z_j) stay within a small factor of 1 on [-1, 1].
static std::vector<double> leja_roots(int m) {
  std::vector<double> z(m), o;
  for (int j = 0; j < m; ++j) z[j] = std::cos((2.0 * j + 1.0) * M_PI / (2.0 * m));
  std::vector<bool> used(m, false);
  for (int k = 0; k < m; ++k) {
    int best = -1;
    double bv = -INFINITY;
    for (int i = 0; i < m; ++i) {
      if (used[i]) continue;
      double lp = 0.0;
      for (double q : o) lp += std::log(std::fabs(z[i] - q));
      if (k == 0) lp = std::fabs(z[i]);
      if (lp > bv) bv = lp, best = i;
    }
    used[best] = true;
    o.push_back(z[best]);
  }
  return o;
}

// Product-form coefficients of one column with Ritz value th (in B):
// T_m(x) = prod_k beta_k (x - z_k), beta_1 = 1, beta_k = 2, and p_k the
// partial products, s_k = max(|p_k(th)|, 1).
//   correction: e_k = beta_k (B - z_k) e_{k-1} + beta_k p_{k-1}(th) r, e_1 = r;
//               step st makes e_{st+1} / s_{st+1} from e_st / s_st (st = 1
//               reads e_1 = r unscaled);
//   plain pass: h_k = (beta_k s_{k-1} / s_k) (B - z_k) h_{k-1}, h_0 = w, and
//               the last step adds gamma^(p-1) f_m.
// co / co2 point at this column's entry of the first step; steps are
// `stride` apart.  *g = gamma^passes with gamma = T_m(th) / s_m.
// pk and sk are scratch arrays of m + 1 entries.
static void product_coefficients(const laph_batch* b, double th, int m, float4* co, float4* co2, size_t stride,
                                 double* g, int passes, double* pk, double* sk) {
  const std::vector<double>& z = b->roots;
  pk[0] = 1.0;
  sk[0] = 1.0;
  for (int k = 1; k <= m; ++k) {
    pk[k] = (k == 1 ? 1.0 : 2.0) * (th - z[k - 1]) * pk[k - 1];
    sk[k] = std::max(std::fabs(pk[k]), 1.0);
  }
  auto step = [&](int k, double in_scale, double rcoef) {  // makes index k from k-1
    const double bk = k == 1 ? 1.0 : 2.0;
    const double a1 = bk * in_scale / sk[k];
    return make_float4((float)a1, (float)rcoef, 0.f, (float)(-a1 * z[k - 1]));
  };
  for (int st = 1; st < m; ++st) {
    const int k = st + 1;
    co[(size_t)(st - 1) * stride] = step(k, st == 1 ? 1.0 : sk[k - 1], 2.0 * pk[k - 1] / sk[k]);
  }
  const double gam = pk[m] / sk[m];
  for (int st = 1; st < m; ++st) co2[(size_t)(st - 1) * stride] = step(st, sk[st - 1], 0.0);
  double gp = gam;
  for (int pp = 2; pp <= LAPH_MAX_PASSES; ++pp, gp *= gam)
    co2[(size_t)(m - 1 + pp - 2) * stride] = step(m, sk[m - 1], gp);
  *g = std::pow(gam, passes);
}

// Order in which the correction pass applies the roots x_k = cos(phi +
// 2 pi k / m), k = 1 .. m-1, of q_theta (below): Leja order of those points
// for a theta just above the window edge (phi = 0.1 i), where the pairs k,
// m - k nearly coincide; one order serves every column.
static std::vector<int> correction_root_order(int m) {
  std::vector<std::complex<double>> z;
  for (int k = 1; k < m; ++k) z.push_back(std::cos(std::complex<double>(2.0 * M_PI * k / m, 0.1)));
  std::vector<int> order;
  std::vector<bool> used(z.size(), false);
  for (size_t n = 0; n < z.size(); ++n) {
    int best = -1;
    double bv = -INFINITY;
    for (size_t i = 0; i < z.size(); ++i) {
      if (used[i]) continue;
      double lp = 0.0;
      for (int j : order) lp += std::log(std::abs(z[i] - z[j - 1]));
      if (n == 0) lp = std::abs(z[i]);
      if (lp > bv) bv = lp, best = (int)i;
    }
    used[best] = true;
    order.push_back(best + 1);
  }
  return order;
}

// Correction pass as a plain product (the default): with r = (B - theta) x,
// T_m(B) x - T_m(theta) x = q(B) r, q(y) = (T_m(y) - T_m(theta)) / (y - theta)
// = 2^(m-1) prod_{k=1}^{m-1} (y - x_k), x_k = cos(phi + 2 pi k / m), theta =
// cos phi (complex phi for theta > 1: complex shifts).  Step st makes
// h_st = q_st(B) r / sigma_st from h_{st-1} (h_0 = r), sigma_st = max(|q_st(theta)|, 1)
// and sigma_{m-1} = sm, so the pass ends with e_m / sm without reading r again.
// cos and sin of 2 pi k / m for the roots of q, in their application order.
static std::vector<std::pair<double, double>> correction_root_rotations(int m) {
  std::vector<std::pair<double, double>> r;
  for (int k : correction_root_order(m)) r.emplace_back(std::cos(2.0 * M_PI * k / m), std::sin(2.0 * M_PI * k / m));
  return r;
}

static void correction_q_coefficients(const std::vector<std::pair<double, double>>& rot, double th, int m, double sm,
                                      float4* co, size_t stride) {
  const std::complex<double> phi = std::acos(std::complex<double>(th, 0.0));
  const std::complex<double> sphi = std::sin(phi);
  std::complex<double> qk(1.0, 0.0);
  double sig = 1.0;
  for (int st = 1; st < m; ++st) {
    // x = cos(phi + 2 pi k / m), with cos(phi) = theta
    const std::complex<double> x = th * rot[st - 1].first - sphi * rot[st - 1].second;
    qk *= 2.0 * (th - x);
    const double sn = st == m - 1 ? sm : std::max(std::abs(qk), 1.0);
    const double a1 = 2.0 * sig / sn;
    const std::complex<double> a3 = -a1 * x;
    co[(size_t)(st - 1) * stride] = make_float4((float)a1, 0.f, (float)a3.imag(), (float)a3.real());
    sig = sn;
  }
}

// Filter degree of each 64-column block of lane timeslice t in a cycle of
// degree m (per-cycle degree mode): m, lowered to what the block's unconverged
// wanted columns need by the margin rule of cycle_degree (a multiple of 4, at
// least 4; a block whose wanted columns have all converged takes 4), kept at m
// for blocks of guard columns only and where a wanted Ritz value sits at the
// window edge, and made non-decreasing over the blocks.  The lower blocks hold
// the lower eigenvalues, which converge first, so the blocks still filtered at
// step s are a suffix: the filter kernels' first-active-block argument of that
// step, with no change to the kernels.  A block of degree d < m ends its
// d - 1 steps in the buffer the pass ends in (d and m are both multiples of 4).
static void block_degrees(const laph_batch* b, const laph_lane* ln, int t, double lct, int m, int* d) {
  const int nb = b->K / 64, T = ln->t0 + t;
  const double c = 2.0 / (b->p.lambda_l - lct);
  const double* mu = &b->h_mu[(size_t)T * b->K];
  const double* res = &b->h_res[(size_t)T * b->K];
  int prev = 4;
  for (int blk = 0; blk < nb; ++blk) {
    int db = m;
    if (blk * 64 < b->nev) {
      double need = 0.0;
      bool est = true;
      for (int j = std::max(blk * 64, b->nlock[T]); j < std::min((blk + 1) * 64, b->nev) && est; ++j) {
        const double r = std::sqrt(res[j]);
        if (r <= b->o.tol) continue;
        const double th = 1.0 + c * (lct - mu[j]);
        if (!(th > 1.0 + 1e-12)) est = false;
        else need = std::max(need, std::acosh(b->deg_safety * r / b->o.tol) / std::acosh(th));
      }
      if (est) db = std::min(m, std::max(4, (int)std::ceil(need / 4.0) * 4));
    }
    db = std::max(db, prev);
    prev = db;
    d[blk] = db;
  }
}

static int upload_coefficients(const View& v, const std::vector<double>& lc, int passes) {
  laph_batch* b = v.b;
  laph_lane* ln = v.ln;
  const int K = b->K, nt = ln->nt, m = v.m();
  const int n2 = m + LAPH_MAX_PASSES - 2;  // plain-pass tables: m-1 shared steps + one last step per pass
  const double lamL = b->p.lambda_l;
  std::vector<float4> co((size_t)(m - 1) * nt * K), co2(b->fdeg_max > 0 ? 0 : (size_t)n2 * nt * K);
  std::vector<double> ga((size_t)nt * K);
  std::vector<float2> cab(nt);
  std::vector<int> fb(nt);
  const int nb64 = K / 64;
  ln->dblk.assign((size_t)nt * nb64, m);
  // per-cycle degree mode: first filtered 64-column block of every correction step
  std::vector<int>& fbs = ln->fbs;
  fbs.assign(b->fdeg_max > 0 ? (size_t)(m - 1) * nt : 0, 0);
  // the timeslices' tables are independent: filled by a few host threads
  auto fill = [&](int tb, int te) {
    std::vector<double> tau(m + 1), s(m + 1);
    for (int t = tb; t < te; ++t) {
      const double c = 2.0 / (lamL - lc[t]);
      const double a = 1.0 + c * lc[t];
      cab[t] = make_float2((float)c, (float)(a - 6.0 * c));
      const int nl = b->nlock[ln->t0 + t];
      fb[t] = nl >= b->nev ? K / 64 : nl / 64;
      int* dt = &ln->dblk[(size_t)t * nb64];
      if (b->fdeg_max > 0) {
        if (ln->blk_cycle) block_degrees(b, ln, t, lc[t], m, dt);
        for (int st = 1; st < m; ++st) {  // blocks of degree d are filtered in steps 1 .. d-1
          int f = fb[t];
          while (f < nb64 && dt[f] - 1 < st) ++f;
          fbs[(size_t)(st - 1) * nt + t] = f;
        }
      }
      const double* mu = &b->h_mu[v.nk() + (size_t)t * K];
      for (int j = 0; j < K; ++j) {
        const double th = 1.0 + c * (lc[t] - mu[j]);  // Ritz value in B
        if (b->fdeg_max > 0) {  // one correction pass of its block's degree (this cycle's degree or lower)
          const int d = dt[j / 64];
          const double tm = std::real(std::cos((double)d * std::acos(std::complex<double>(th, 0.0))));
          const double sm = std::max(std::fabs(tm), 1.0);
          correction_q_coefficients(b->qrot[d], th, d, sm, &co[(size_t)t * K + j], (size_t)nt * K);
          ga[(size_t)t * K + j] = tm / sm;
          continue;
        }
        if (b->prod) {
          product_coefficients(b, th, m, &co[(size_t)t * K + j], &co2[(size_t)t * K + j], (size_t)nt * K,
                               &ga[(size_t)t * K + j], passes, tau.data(), s.data());
          if (b->qform) correction_q_coefficients(b->qrot[m], th, m, s[m], &co[(size_t)t * K + j], (size_t)nt * K);
          continue;
        }
        tau[0] = 1.0;
        tau[1] = th;
        for (int k = 1; k < m; ++k) tau[k + 1] = 2.0 * th * tau[k] - tau[k - 1];
        for (int k = 0; k <= m; ++k) s[k] = std::max(std::fabs(tau[k]), 1.0);
        for (int st = 1; st < m; ++st) {
          float4 q;
          if (st == 1) {
            q = make_float4((float)(2.0 / s[2]), (float)(2.0 * tau[1] / s[2]), 0.f, 0.f);
          } else if (st == 2) {
            q = make_float4((float)(2.0 * s[2] / s[3]), (float)((2.0 * tau[2] - 1.0) / s[3]), 0.f, 0.f);
          } else {
            q = make_float4((float)(2.0 * s[st] / s[st + 1]), (float)(2.0 * tau[st] / s[st + 1]),
                            (float)(-s[st - 1] / s[st + 1]), 0.f);
          }
          co[((size_t)(st - 1) * nt + t) * K + j] = q;
        }
        // plain pass p on w = w_{p-1}: h_k = T_k(B) w / s_k with h_0 = w,
        // h_1 = B w / s_1, h_{k+1} = (2 s_k/s_{k+1}) B h_k - (s_{k-1}/s_{k+1}) h_{k-1};
        // its last step also adds gamma^(p-1) f_m (source slot), giving w_p
        const double g = tau[m] / s[m];
        co2[((size_t)0 * nt + t) * K + j] = make_float4((float)(1.0 / s[1]), 0.f, 0.f, 0.f);
        for (int st = 2; st < m; ++st)
          co2[((size_t)(st - 1) * nt + t) * K + j] =
              make_float4((float)(2.0 * s[st - 1] / s[st]), 0.f, (float)(-s[st - 2] / s[st]), 0.f);
        double gp = g;
        for (int pp = 2; pp <= LAPH_MAX_PASSES; ++pp, gp *= g)
          co2[((size_t)(m - 1 + pp - 2) * nt + t) * K + j] =
              make_float4((float)(2.0 * s[m - 1] / s[m]), (float)gp, (float)(-s[m - 2] / s[m]), 0.f);
        ga[(size_t)t * K + j] = std::pow(g, passes);
      }
    }
  };
  // the kernels read w = a1 (a_t - 6 c_t) + a3 in the last slot (their links
  // are prescaled by c_t): folded per timeslice after its entries are made
  auto fold = [&](int tb, int te) {
    for (int t = tb; t < te; ++t) {
      const double cy = cab[t].y;
      for (int st = 0; st + 1 < m; ++st)
        for (int j = 0; j < K; ++j) {
          float4& e = co[((size_t)st * nt + t) * K + j];
          e.w = (float)((double)e.x * cy + (double)e.w);
        }
      for (size_t st = 0; st < co2.size() / ((size_t)nt * K); ++st)
        for (int j = 0; j < K; ++j) {
          float4& e = co2[(st * nt + t) * K + j];
          e.w = (float)((double)e.x * cy + (double)e.w);
        }
    }
  };
  auto fill_fold = [&](int tb, int te) {
    fill(tb, te);
    fold(tb, te);
  };
  const int nth = std::max(1, std::min({8, nt, (int)std::thread::hardware_concurrency()}));
  std::vector<std::thread> pool;
  for (int h = 1; h < nth; ++h) pool.emplace_back(fill_fold, h * nt / nth, (h + 1) * nt / nth);
  fill_fold(0, nt / nth);
  for (auto& th : pool) th.join();
  // coefficient table of the lane: [step][lane timeslice][K] at offset step*nts*K + t0*K
  for (int st = 1; st < m; ++st)
    LAPH_HIP(hipMemcpyAsync(b->coef + (size_t)(st - 1) * b->nts * K + v.nk(), co.data() + (size_t)(st - 1) * nt * K,
                            (size_t)nt * K * sizeof(float4), hipMemcpyHostToDevice, ln->st));
  for (int st = 1; b->fdeg_max == 0 && st <= n2; ++st)  // plain passes exist only at the batch's degree
    LAPH_HIP(hipMemcpyAsync(b->coef2 + (size_t)(st - 1) * b->nts * K + v.nk(), co2.data() + (size_t)(st - 1) * nt * K,
                            (size_t)nt * K * sizeof(float4), hipMemcpyHostToDevice, ln->st));
  LAPH_HIP(hipMemcpyAsync(b->gam + v.nk(), ga.data(), ga.size() * sizeof(double), hipMemcpyHostToDevice, ln->st));
  LAPH_HIP(hipMemcpyAsync(b->cab + ln->t0, cab.data(), cab.size() * sizeof(float2), hipMemcpyHostToDevice, ln->st));
  LAPH_HIP(hipMemcpyAsync(b->first_blk + ln->t0, fb.data(), fb.size() * sizeof(int), hipMemcpyHostToDevice, ln->st));
  for (int st = 1; !fbs.empty() && st < m; ++st)  // one row per step, as the coefficient table
    LAPH_HIP(hipMemcpyAsync(b->first_blk_steps + (size_t)(st - 1) * b->nts + ln->t0, fbs.data() + (size_t)(st - 1) * nt,
                            nt * sizeof(int), hipMemcpyHostToDevice, ln->st));
  return LAPH_OK;
}

// The m-1 FP32 filter steps of a lane; buffers rotate R32 -> P0 -> P1 -> P2 -> P0 ...
static int enqueue_filter(const View& v) {
  laph_batch* b = v.b;
  const int m = v.m(), K = b->K, nt = v.ln->nt, L = b->p.L;
  const size_t step = (size_t)b->nts * K;
  const size_t vo = v.vo();
  void* R = fb(b, b->R32, vo);
  const int nb = b->prod ? 2 : 3;  // the product form needs no previous vector
  for (int st = 1; st < m; ++st) {
    const void* in = st == 1 ? R : fb(b, b->P[(st - 2) % nb], vo);
    const void* old = st >= 3 && !b->prod ? fb(b, b->P[(st - 3) % 3], vo) : nullptr;
    void* out = fb(b, b->P[(st - 1) % nb], vo);
    // per-cycle degree mode: the step's own first filtered block (block degrees)
    const int* fbl = b->fdeg_max > 0 ? b->first_blk_steps + (size_t)(st - 1) * b->nts : b->first_blk;
    LAPH_HIP(launch_filter_step(b->U32c + v.lo(), in, b->qform ? nullptr : R, old, out,
                                b->coef + (st - 1) * step + v.nk(), b->cab + v.ln->t0, fbl + v.ln->t0, L, K, nt,
                                b->bw, b->bf16, b->strip, stencil_zrun(b), v.ln->st));
  }
  return LAPH_OK;
}

// Buffers of the plain passes: f = f_m stays where the correction pass left
// it (every plain pass adds gamma^(p-1) f in its last step); the recurrences
// rotate over the other three buffers q[] (R32 is free by then).
static void plain_buffers(const View& v, void** pf, void* q[3]) {
  laph_batch* b = v.b;
  const int m = v.m();
  void* all[4] = {b->R32, b->P[0], b->P[1], b->P[2]};
  void* f = b->P[(m - 2) % (b->prod ? 2 : 3)];
  int k = 0;
  for (void* a : all)
    if (a != f) q[k++] = fb(b, a, v.vo());
  *pf = fb(b, f, v.vo());
}

// Product form: the two slots of q[] pass p >= 2 writes, alternately (pass 2
// reads f, a later pass the slot where the previous one ended).
static void product_outputs(int m, int p, int o[2]) {
  int a = -1;
  for (int k = 2; k <= p; ++k) {
    o[0] = k == 2 ? 0 : (a + 1) % 3;
    o[1] = k == 2 ? 1 : (a + 2) % 3;
    a = o[(m - 1) % 2];
  }
}

// Slot of q[] holding w_p after plain pass p >= 2.
static int plain_result(const laph_batch* b, int p) {
  const int m = b->deg;
  if (b->prod) {
    int o[2];
    product_outputs(m, p, o);
    return o[(m - 1) % 2];
  }
  int r = (m - 1) % 3;
  for (int k = 3; k <= p; ++k) r = (r + m) % 3;
  return r;
}

// Plain pass p >= 2: w_p = T_m(B) w_{p-1} / s_m + gamma^(p-1) f (w_1 = f).
// Pass 2 reads f as h_0; pass p > 2 starts from q[a], a = where pass p-1
// ended, and rotates q[a], q[a+1], q[a+2] so h_0 is overwritten only after
// its last use (step 2).
static int enqueue_plain_pass(const View& v, int p) {
  laph_batch* b = v.b;
  const int m = b->deg, K = b->K, nt = v.ln->nt, L = b->p.L;
  const size_t step = (size_t)b->nts * K;
  void *f, *q[3];
  plain_buffers(v, &f, q);
  const int a = p == 2 ? 0 : plain_result(b, p - 1);
  int po[2] = {0, 1};
  if (b->prod) product_outputs(m, p, po);
  for (int st = 1; st <= m; ++st) {
    const void *in, *old;
    void* out;
    if (b->prod) {
      in = st == 1 ? (p == 2 ? f : q[a]) : q[po[(st - 2) % 2]];
      old = nullptr;
      out = q[po[(st - 1) % 2]];
    } else if (p == 2) {
      in = st == 1 ? f : q[(st - 2) % 3];
      old = st == 1 ? nullptr : (st == 2 ? f : q[(st - 3) % 3]);
      out = q[(st - 1) % 3];
    } else {
      in = q[(a + st - 1) % 3];
      old = st >= 2 ? q[(a + st - 2) % 3] : nullptr;
      out = q[(a + st) % 3];
    }
    const bool last = st == m;
    const float4* co = b->coef2 + (size_t)(last ? m - 1 + p - 2 : st - 1) * step + v.nk();
    LAPH_HIP(launch_filter_step(b->U32c + v.lo(), in, last ? f : nullptr, old, out, co, b->cab + v.ln->t0,
                                b->first_blk + v.ln->t0, L, K, nt, b->bw, b->bf16, b->strip, stencil_zrun(b),
                                v.ln->st));
  }
  return LAPH_OK;
}

static int capture(const View& v, hipGraphExec_t* exec, const std::function<int()>& enqueue) {
  laph_lane* ln = v.ln;
  hipGraph_t g;
  LAPH_HIP(hipStreamBeginCapture(ln->st, hipStreamCaptureModeThreadLocal));
  int rc = enqueue();
  hipError_t e = hipStreamEndCapture(ln->st, &g);
  LAPH_TRY(rc);
  LAPH_HIP(e);
  LAPH_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  LAPH_HIP(hipGraphDestroy(g));
  return LAPH_OK;
}

// Replays the captured correction pass of the lane and passes 2..`passes`
// (each captured on first use).  *w <- w_passes (f_m for one pass), with
// T_m(B)^passes x / s_m^passes = gamma^passes x + w_passes.
static int run_filter(const View& v, int passes, void** w) {
  laph_lane* ln = v.ln;
  // the batch's own degree uses ln->graph; other degrees (per-cycle degree
  // mode) a graph each, captured on first use
  hipGraphExec_t* g = v.m() == v.b->deg ? &ln->graph : &ln->gdeg[v.m()];
  const auto c0 = std::chrono::steady_clock::now();  // host time of first-use captures (warm-up degree timing)
  if (!*g) LAPH_TRY(capture(v, g, [&] { return enqueue_filter(v); }));
  for (int p = 2; p <= passes; ++p)
    if (!ln->gplain[p - 2]) LAPH_TRY(capture(v, &ln->gplain[p - 2], [&] { return enqueue_plain_pass(v, p); }));
  ln->s_capture += std::chrono::duration<double>(std::chrono::steady_clock::now() - c0).count();
  // the steps' links: c_t U for this cycle's window of every timeslice
  LAPH_HIP(launch_scale_links32(v.b->U32 + v.lo(), v.b->cab + ln->t0, v.b->U32c + v.lo(), (size_t)v.b->V3 * 27,
                                ln->nt, ln->st));
  LAPH_HIP(hipGraphLaunch(*g, ln->st));
  void *f, *q[3];
  plain_buffers(v, &f, q);
  *w = f;
  for (int p = 2; p <= passes; ++p) {
    LAPH_HIP(hipGraphLaunch(ln->gplain[p - 2], ln->st));
    *w = q[plain_result(v.b, p)];
  }
  return LAPH_OK;
}

template <class E>
__device__ __forceinline__ float2 load_entry(const E* p, size_t i);
template <>
__device__ __forceinline__ float2 load_entry<float2>(const float2* p, size_t i) {
  return p[i];
}
template <>
__device__ __forceinline__ float2 load_entry<unsigned>(const unsigned* p, size_t i) {
  const unsigned u = p[i];
  return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}

// X <- g X + w  (g = gamma^passes per column, w block-major in the filter
// storage: E = float2 for FP32, unsigned for a BF16 complex pair)
template <class E>
__global__ void k_combine_gamma(double2* __restrict__ X, const E* __restrict__ W, const double* __restrict__ gam,
                                size_t per_ts, int K, const int* __restrict__ first_blk, int bw) {
  const int ts = blockIdx.y;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts) return;
  const int col = (int)(i % K);
  if (col < first_blk[ts] * 64) return;
  const size_t idx = (size_t)ts * per_ts + i;
  const double g = gam[(size_t)ts * K + col];
  const size_t row = i / K;  // W is block-major: [ts][col/bw][row][bw]
  const size_t wi = (size_t)ts * per_ts + ((size_t)(col / bw) * (per_ts / K) + row) * bw + (col % bw);
  const float2 w = load_entry(W, wi);
  const double2 x = X[idx];
  X[idx] = make_double2(g * x.x + (double)w.x, g * x.y + (double)w.y);
}

// Applications of T_m(B) in this cycle of a lane: the requested count, cut
// so that the largest amplification T_m(theta_max)^P stays <= amp_cap (1e10),
// with theta_max = 1 + c lambda_c >= the image of the lowest eigenvalue of
// -D (-D >= 0).  A column whose Ritz value sits at the window edge (theta =
// 1, no scaling) grows its components along the lowest eigenvectors by that
// factor: far beyond the cap such buffer columns become numerically parallel
// in FP64 (their FP32 rounding then swamps their own direction), and a wide
// early window would overflow FP32.
static int filter_passes(const laph_batch* b, const std::vector<double>& lc) {
  const int want = std::min(b->o.filter_passes > 0 ? b->o.filter_passes : 5, LAPH_MAX_PASSES);
  double amax = 1.0;
  for (double l : lc) {
    const double th = 1.0 + 2.0 * l / (b->p.lambda_l - l);
    amax = std::max(amax, std::cosh(b->deg * std::acosh(th)));
  }
  const int pmax = amax > 1.0 ? (int)std::floor(std::log(b->amp_cap) / std::log(amax)) : want;
  return std::max(1, std::min(want, pmax));
}

// Per-cycle degree mode: the filter degree of this cycle of a lane, the
// largest multiple of 4 up to fdeg_max with T_d(theta_max) <= amp_cap (one
// application per cycle); the wide window of the first cycle gets a low
// degree, the settled windows a high one.
// Degree of a cycle: the highest multiple of 4 (at most fdeg_max) whose
// amplification T_d(theta_max) stays under amp_cap, lowered to what the
// unconverged wanted columns need when that is less: a Ritz pair (theta, r)
// comes out of a filter of degree d with its residual about r / T_d(theta)
// (within 1.3x in the last cycles of a 32^3 solve, 7x in the early ones), so
// the cycle takes the lowest multiple of 4 with T_d(theta_j) >= S r_j / tol
// for every such column j of the lane (S = deg_safety; 0 switches this off).
static int cycle_degree(const laph_batch* b, const laph_lane* ln, const std::vector<double>& lc) {
  double pmax = 0.0;  // largest acosh(theta_max) over the lane's timeslices
  for (double l : lc) pmax = std::max(pmax, std::acosh(1.0 + 2.0 * l / (b->p.lambda_l - l)));
  int d = pmax > 0.0 ? (int)std::floor(std::acosh(b->amp_cap) / pmax) : b->fdeg_max;
  d = std::max(std::min(d, b->fdeg_max) / 4 * 4, 4);
  if (b->deg_safety <= 0.0) return d;
  double need = 0.0;
  for (int t = 0; t < ln->nt; ++t) {
    const int T = ln->t0 + t;
    const double c = 2.0 / (b->p.lambda_l - lc[t]);
    const double* mu = &b->h_mu[(size_t)T * b->K];
    const double* res = &b->h_res[(size_t)T * b->K];
    for (int j = b->nlock[T]; j < b->nev; ++j) {
      const double r = std::sqrt(res[j]);
      if (r <= b->o.tol) continue;
      const double th = 1.0 + c * (lc[t] - mu[j]);
      if (!(th > 1.0 + 1e-12)) return d;  // a wanted value at or above the window edge: no estimate
      need = std::max(need, std::acosh(b->deg_safety * r / b->o.tol) / std::acosh(th));
    }
  }
  const int dn = std::max(4, (int)std::ceil(need / 4.0) * 4);
  return std::min(d, dn);
}

// Solve loop of one lane (runs on its own host thread).
// Largest residual of the wanted pairs not yet converged over the lane's
// timeslices (0 when all are).
static double worst_residual(const laph_batch* b, const laph_lane* ln) {
  double w = 0.0;
  for (int t = 0; t < ln->nt; ++t) {
    const double* res = &b->h_res[(size_t)(ln->t0 + t) * b->K];
    for (int j = 0; j < b->nev; ++j) {
      const double r = std::sqrt(res[j]);
      if (r > b->o.tol) w = std::max(w, r);
    }
  }
  return w;
}

static int lane_solve(laph_batch* b, laph_lane* ln, bool warm) {
  const View v{b, ln};
  const int K = b->K, nt = ln->nt;
  ln->iX = 0;
  ln->iAX = 1;
  ln->iT = 2;
  ln->lock = 0;
  ln->cycles = 0;
  ln->sv32 = ln->sv64 = ln->flaunch = 0;
  ln->recoveries = 0;
  ln->rr_fallbacks = ln->shifted = 0;
  ln->subset_cycles = 0;
  ln->fbytes = 0;
  ln->ms_filter = ln->ms_dense = 0;
  ln->s_capture = 0;
  {
    std::vector<int> zero(nt, 0);
    LAPH_HIP(hipMemcpyAsync(b->lock_blk + ln->t0, zero.data(), nt * sizeof(int), hipMemcpyHostToDevice, ln->st));
  }
  // start block (network proposal or seeded Gaussian), CholQR2, first Rayleigh-Ritz
  if (!warm) LAPH_HIP(launch_random_block(v.X(), (size_t)nt * b->per_ts, b->o.seed + 0x1000193ull * ln->t0, ln->st));
  const bool start_rr = warm || b->start_rr || b->qr_cycles > 0 || b->warm_cycles == 0;
  const bool warmup0 = !(warm && b->o.warm_start == 2) && b->warm_cycles > 0;
  // The warm-up cycle after a Rayleigh-Ritz start keeps its window's lower end
  // at most at the start window: the largest Ritz value of a start block that
  // is not yet filtered (a random block, or a proposal only partly inside the
  // wanted space) sits high in the spectrum, and a window that high makes the
  // warm-up filter a weak low-pass.  The cap holds for the residuals written
  // by the start's Rayleigh-Ritz step and for the filter of that cycle.
  ln->wcap = start_rr && warmup0 && b->warmup_cap ? std::min(b->window0, 0.75 * b->p.lambda_l)
                                                 : 0.75 * b->p.lambda_l;
  if (start_rr)
    LAPH_TRY(orthonormalise_and_rr(v, Act{0, K}, b->qr_cycles > 0));
  else
    LAPH_TRY(start_from_block(v));
  std::vector<double> lc(nt);
  int active = 0;
  LAPH_TRY(residuals_and_window(v, lc, &active));
  const int max_iter = b->o.max_iter > 0 ? b->o.max_iter : 60;
  int it = 0;
  for (it = 0; it < max_iter && active > 0; ++it) {
    const auto tc0 = std::chrono::steady_clock::now();
    // one filter pass while the window settles, then T_m(B)^P per cycle; in
    // the per-cycle degree mode one pass of the degree the window allows
    // a network proposal (warm_start = 2) replaces the warm-up cycles: its
    // Rayleigh-Ritz values give the window of the first cycle
    const bool warmup = it < (warm && b->o.warm_start == 2 ? 0 : b->warm_cycles);
    const bool wdeg = b->fdeg_max >= b->warm_deg && b->warm_deg > 0;  // warm-up degree of the per-cycle mode
    ln->m = b->fdeg_max > 0 && !warmup ? cycle_degree(b, ln, lc) : (warmup && wdeg ? b->warm_deg : b->deg);
    const int passes = b->fdeg_max > 0 || warmup ? 1 : filter_passes(b, lc);
    const int m = ln->m;
    ln->blk_cycle = b->block_degrees && b->fdeg_max > 0 && !warmup && b->deg_safety > 0.0;
    LAPH_TRY(upload_coefficients(v, lc, passes));
    for (int t = 0; t < nt; ++t) {
      const int nl = b->nlock[ln->t0 + t];
      if (nl >= b->nev) continue;
      if (b->fdeg_max > 0) {  // correction pass of block degrees: step s filters the blocks of degree > s
        const int nb64 = K / 64;
        const int* dt = &ln->dblk[(size_t)t * nb64];
        for (int st = 1; st < m; ++st) {
          int f = nl / 64;
          while (f < nb64 && dt[f] - 1 < st) ++f;
          const int cols = K - f * 64;
          if (cols == 0) break;
          const int nblk = (cols + b->bw - 1) / b->bw;
          ln->sv32 += cols;
          ln->fbytes += (double)b->V3 * ((double)cols * 2 * 3.0 * b->esz + nblk * 216.0);
        }
        continue;
      }
      const int cols = K - (nl / 64) * 64;
      const int nblk = (cols + b->bw - 1) / b->bw;  // blocks the kernel visits (links once each)
      ln->sv32 += (long long)(m - 1) * cols;
      // algorithmic bytes (vector touches per step): product form, correction
      // step 1 reads r and writes e_2 (2), later steps read e_k and r and
      // write (3); plain-pass steps read and write (2), the last one also
      // reads f (3).  Three-term form: later correction steps also read
      // e_{k-1} (4), plain-pass steps after the first the previous vector.
      // A vector touch is 3 colours x esz bytes per site; links 27 x 8 B per
      // site per vector block
      for (int s = 1; s < m; ++s) {
        const int touches = s == 1 || b->qform ? 2 : (s == 2 || b->prod ? 3 : 4);
        ln->fbytes += (double)b->V3 * ((double)cols * touches * 3.0 * b->esz + nblk * 216.0);
      }
      for (int p = 2; p <= passes; ++p) {
        ln->sv32 += (long long)m * cols;
        for (int s = 1; s <= m; ++s) {
          const int touches = (s == 1 || b->prod ? 2 : 3) + (s == m ? 1 : 0);
          ln->fbytes += (double)b->V3 * ((double)cols * touches * 3.0 * b->esz + nblk * 216.0);
        }
      }
    }
    ln->flaunch += (m - 1) + (long long)(passes - 1) * m;
    ln->min_pivot2 = INFINITY;
    void* w = nullptr;
    LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
    LAPH_TRY(run_filter(v, passes, &w));
    LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
    ln->wcap = 0.75 * b->p.lambda_l;  // later windows: the Ritz values (the next residuals use this cap)
    dim3 g((unsigned)((b->per_ts + 255) / 256), nt);
    if (b->bf16)
      k_combine_gamma<unsigned><<<g, 256, 0, ln->st>>>(v.X(), static_cast<const unsigned*>(w), b->gam + v.nk(),
                                                      b->per_ts, K, b->first_blk + ln->t0, b->bw);
    else
      k_combine_gamma<float2><<<g, 256, 0, ln->st>>>(v.X(), static_cast<const float2*>(w), b->gam + v.nk(),
                                                    b->per_ts, K, b->first_blk + ln->t0, b->bw);
    LAPH_HIP(hipGetLastError());
    const Act q{ln->lock, K - ln->lock};
    if (it == 2 && b->poison_ts >= ln->t0 && b->poison_ts < ln->t0 + nt)  // LAPH_TEST_POISON (tests)
      LAPH_HIP(hipMemsetAsync(v.X() + (size_t)(b->poison_ts - ln->t0) * b->per_ts + 7 * (size_t)K + q.l, 0xFF,
                              (size_t)q.a * sizeof(double2), ln->st));
    // explicit CholQR2 while the block is far from orthonormal (first
    // cycles from a random start), Cholesky inside Rayleigh-Ritz after
    LAPH_TRY(dense_steps(v, q, it < b->qr_cycles));
    LAPH_HIP(hipEventRecord(ln->ev[2], ln->st));
    LAPH_TRY(residuals_and_window(v, lc, &active));
    float a = 0, c = 0;
    LAPH_HIP(hipEventElapsedTime(&a, ln->ev[0], ln->ev[1]));
    LAPH_HIP(hipEventElapsedTime(&c, ln->ev[1], ln->ev[2]));
    ln->ms_filter += a;
    ln->ms_dense += c;
    if (b->trace_time && ln->t0 == 0) {
      const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc0).count();
      std::fprintf(stderr, "[laph] cycle %d time: filter %.1f ms  dense %.1f ms  wall %.1f ms  active width %d  "
                   "unconverged timeslices %d of %d\n", it, a, c, wall, q.a, active, nt);
    }
    if (b->trace && ln->t0 == 0) {  // per-cycle trace of the first timeslice
      const double worst = worst_residual(b, ln);
      const double* mu = &b->h_mu[0];
      const double* rs = &b->h_res[0];
      std::fprintf(stderr,
                   "[laph] cycle %d: degree %d passes %d  window %.4f  mu[nev-1] %.5f  mu[K-1] %.4f  res[0] %.2e  "
                   "res[nev-1] %.2e  worst %.2e  locked %d  lock %d  min pivot^2 %.2e (timeslice %d)  rr fallbacks %d  "
                   "shifted %d  unconverged timeslices %d of %d\n", it, m, passes, lc[0], mu[b->nev - 1], mu[K - 1],
                   std::sqrt(rs[0]), std::sqrt(rs[b->nev - 1]), worst, b->nlock[0], ln->lock, ln->min_pivot2,
                   ln->min_pivot_ts, ln->rr_fallbacks, ln->shifted, active, nt);
      // per 64-column block of timeslice 0: largest residual of its unconverged wanted columns, and
      // acosh(theta) (theta in the next cycle's window) of its slowest wanted column
      const double c = 2.0 / (b->p.lambda_l - lc[0]);
      std::fprintf(stderr, "[laph]   block degrees:");
      for (int blk = 0; blk < K / 64; ++blk) std::fprintf(stderr, " %d", ln->dblk[blk]);
      std::fprintf(stderr, "\n[laph]   blocks:");
      for (int blk = 0; blk * 64 < b->nev; ++blk) {
        double rmax = 0.0;
        const int jend = std::min((blk + 1) * 64, b->nev);
        for (int j = std::max(blk * 64, b->nlock[0]); j < jend; ++j) rmax = std::max(rmax, std::sqrt(rs[j]));
        const double th = 1.0 + c * (lc[0] - mu[jend - 1]);
        std::fprintf(stderr, "  [%d] res %.2e psi %.4f", blk, rmax, th > 1.0 ? std::acosh(th) : 0.0);
      }
      std::fprintf(stderr, "\n");
    }
  }
  ln->cycles = it;
  // eigenvectors of every lane end in B[0]
  if (ln->iX != 0)
    LAPH_HIP(hipMemcpyAsync(v.blk(0), v.X(), (size_t)nt * b->per_ts * sizeof(double2), hipMemcpyDeviceToDevice,
                            ln->st));
  LAPH_HIP(hipStreamSynchronize(ln->st));
  ln->iX = 0;
  return LAPH_OK;
}

// Device time of `reps` replays of lane 0's correction pass at the current
// block width (graph recaptured for that width).
// reject_above > 0: a layout whose first (warm-up) pass already takes longer
// than that many ms is reported at reps x that time, without further passes.
static int time_correction_pass(laph_batch* b, int reps, float* ms, float reject_above = 0.f) {
  laph_lane* ln = &b->lanes[0];
  const View v{b, ln};
  if (ln->graph) LAPH_HIP(hipGraphExecDestroy(ln->graph));
  ln->graph = nullptr;
  LAPH_TRY(capture(v, &ln->graph, [&] { return enqueue_filter(v); }));
  LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
  LAPH_HIP(hipGraphLaunch(ln->graph, ln->st));  // warm-up
  LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
  LAPH_HIP(hipEventSynchronize(ln->ev[1]));
  float first = 0.f;
  LAPH_HIP(hipEventElapsedTime(&first, ln->ev[0], ln->ev[1]));
  if (reject_above > 0.f && first > reject_above) {
    *ms = first * reps;
    return LAPH_OK;
  }
  LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
  for (int r = 0; r < reps; ++r) LAPH_HIP(hipGraphLaunch(ln->graph, ln->st));
  LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
  LAPH_HIP(hipEventSynchronize(ln->ev[1]));
  LAPH_HIP(hipEventElapsedTime(ms, ln->ev[0], ln->ev[1]));
  return LAPH_OK;
}

// Layout cache (LAPH_TUNE_FILE): the layouts the first solve chose, one text
// line per shape and device,
//   "<arch> cu<CUs> L<L> K<K> nt<lane 0 timeslices> s<bf16> f<prod><qform> z<zrun>
//    d<DEG> n<nev> t<tol> : bw <width> strip <strip> split <row blocks>
//    chunks <site chunks> mfma <0|1> warm <warm-up degree>"
// read before timing (the first valid line of the key wins) and appended
// after a miss, so that later processes on the same shape skip the timing
// (several seconds at 32^3 x 64) and pick the same layout.  Used only while no
// layout control (LAPH_BW, LAPH_STRIP, LAPH_GRAM_SPLIT, LAPH_GRAM_MFMA,
// LAPH_WARM_DEG) is set.
static std::string tune_key(const laph_batch* b) {
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, b->p.device) != hipSuccess) return std::string();
  char k[384];
  std::snprintf(k, sizeof k, "%s cu%d L%d K%d nt%d s%d f%d%d z%d d%d n%d t%g", pr.gcnArchName,
                pr.multiProcessorCount, b->p.L, b->K, b->lanes[0].nt, (int)b->bf16, (int)b->prod, (int)b->qform,
                b->zrun, b->deg, b->nev, b->o.tol);
  return std::string(k);
}

// Warm-up degrees the first solve times (per-cycle degree mode with a warm-up
// cycle, LAPH_WARM_DEG unset): 24, 32 and 40 above the configuration's degree,
// each at most the highest degree whose amplification at the start window stays
// under warm_amp_cap (that degree itself when it is below 40), highest first,
// then the configuration's degree.  A higher warm-up degree can save a whole
// cycle (24^3: 5 -> 4, 48^3: 6 -> 5) or only cost filter steps (32^3: 4 cycles
// at every degree), so it is measured rather than set.  The bound is wider than
// a cycle's amp_cap: every candidate is timed as a whole solve, judged by the
// same residual test, and one whose solve fails is dropped (at 24^3 degree 40,
// amplification 1.3e11, saves a cycle: 91.3 -> 102.1 timeslices/s).  Empty:
// nothing to choose.
static std::vector<int> warm_candidates(const laph_batch* b) {
  std::vector<int> c;
  if (b->fdeg_max <= 0 || b->warm_cycles <= 0 || b->warm_deg_fixed) return c;
  const double w = b->window0, lamL = b->p.lambda_l;
  const double p0 = std::acosh(1.0 + 2.0 * w / (lamL - w));
  int dcap = p0 > 0.0 ? (int)std::floor(std::acosh(b->warm_amp_cap) / p0) : b->fdeg_max;
  dcap = std::min(dcap, b->fdeg_max) / 4 * 4;
  for (int d : {std::min(40, dcap), 32, 24})
    if (d > b->deg && d <= dcap && std::find(c.begin(), c.end(), d) == c.end()) c.push_back(d);
  if (c.empty()) return c;
  c.push_back(b->deg);
  return c;
}

static bool tune_lookup(laph_batch* b, const char* path, const std::string& key) {
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool hit = false;
  const std::vector<int> cand = block_widths(b);
  while (!hit && std::fgets(line, sizeof line, f)) {
    const char* sep = std::strstr(line, " : ");
    if (!sep || (size_t)(sep - line) != key.size() || std::strncmp(line, key.c_str(), key.size()) != 0) continue;
    int bw = 0, strip = 0, split = 0, ks = 0, mf = 0, wd = 0;
    if (std::sscanf(sep + 3, "bw %d strip %d split %d chunks %d mfma %d warm %d", &bw, &strip, &split, &ks, &mf,
                    &wd) != 6)
      continue;
    const std::vector<int> wc = warm_candidates(b);
    const bool ok = std::find(cand.begin(), cand.end(), bw) != cand.end() && b->K % bw == 0 && strip >= 1 &&
                    b->p.L % strip == 0 && split >= 1 && split <= 8 && ks >= 1 && ks <= b->gram_ksplit_max &&
                    b->n % ks == 0 && (mf == 0 || (mf == 1 && split == 1 && gram_mfma_fits(b->n, b->K, ks))) &&
                    (wc.empty() || std::find(wc.begin(), wc.end(), wd) != wc.end());
    if (!ok) continue;
    b->bw = bw;
    b->strip = strip;
    b->gram_split = split;
    b->gram_ksplit = ks;
    b->gram_mfma = mf;
    if (!wc.empty()) {
      b->warm_deg = wd;
      b->warm_tuned = true;
    }
    hit = true;
  }
  std::fclose(f);
  return hit;
}

static void tune_store(const laph_batch* b, const char* path, const std::string& key) {
  FILE* f = std::fopen(path, "a");
  if (!f) return;  // an unwritable cache only costs the next process the timing
  std::fprintf(f, "%s : bw %d strip %d split %d chunks %d mfma %d warm %d\n", key.c_str(), b->bw, b->strip,
               b->gram_split, b->gram_ksplit, b->gram_mfma, b->warm_deg > 0 ? b->warm_deg : b->deg);
  std::fclose(f);
}

// Layout of the filter: block width (the candidates of the storage format,
// widest first) and sweep strip width (L, or 4 or 8 when they divide L), each
// pair timed on lane 0's timeslices with every column active (the kernels
// only stream the buffers; their contents do not change the time); the
// fastest is kept.  A pair whose warm-up pass already takes 1.5 x the best
// pass so far is not timed further (the 32-vector blocks at 32^3: 14.9
// against 6.0 ms per step).  LAPH_BW / LAPH_STRIP fix either.  Then the Gram
// layout.  LAPH_TUNE_FILE caches both.
static int autotune_block_width(laph_batch* b) {
  laph_lane* ln = &b->lanes[0];
  const char* tf = std::getenv("LAPH_TUNE_FILE");
  const bool cache = tf && *tf && b->bw == 0 && !b->strip_fixed && !b->gram_split_fixed && !b->warm_deg_fixed;
  const std::string key = cache ? tune_key(b) : std::string();
  if (!key.empty() && tune_lookup(b, tf, key)) {
    b->tuned = b->tuned_from_cache = true;
    if (b->trace)
      std::fprintf(stderr, "[laph] layout from %s: block width %d, strip %d, Gram %s, row blocks %d, site chunks %d\n",
                   tf, b->bw, b->strip, b->gram_mfma ? "MFMA kernel" : "rocBLAS", b->gram_split, b->gram_ksplit);
    return LAPH_OK;
  }
  ln->m = 0;  // layouts are timed at the batch's degree
  LAPH_HIP(hipMemsetAsync(b->first_blk + ln->t0, 0, ln->nt * sizeof(int), ln->st));
  std::vector<int> widths, strips;
  for (int w : block_widths(b))
    if (b->K % w == 0 && (b->bw == 0 || b->bw == w)) widths.push_back(w);
  std::sort(widths.rbegin(), widths.rend());
  if (b->strip_fixed) {
    strips.push_back(b->strip);
  } else {
    strips.push_back(b->p.L);
    if (b->p.L % 4 == 0 && b->p.L > 4) strips.push_back(4);
    if (b->p.L % 8 == 0 && b->p.L > 8) strips.push_back(8);
  }
  float best = 0;
  int wbest = widths[0], sbest = strips[0];
  bool first = true;
  for (int w : widths)
    for (int sw : strips) {
      b->bw = w;
      b->strip = sw;
      float ms = 0;
      LAPH_TRY(time_correction_pass(b, 2, &ms, first ? 0.f : 0.75f * best));
      if (first || ms < best) {
        best = ms;
        wbest = w;
        sbest = sw;
        first = false;
      }
    }
  b->bw = wbest;
  b->strip = sbest;
  b->tuned = true;
  if (!b->gram_split_fixed) {
    // Gram row blocks: the upper-triangle split saves a quarter of the flops
    // but launches fewer tiles per GEMM, which leaves CUs idle when the lane
    // holds few timeslices (24 at 48^3: 55% of the FP64 matrix peak against
    // 90% with 64 timeslices at 32^3); timed on the lane's block at full width
    const View v{b, ln};
    const Act q{0, b->K};
    float best_ms = 0;
    int best_split = 1, best_ks = 1, best_mfma = 0;
    bool first_g = true;
    for (int mf : {1, 0})
      for (int split : {1, 2})
      for (int ks : {1, 2, 4, 8}) {
        if (b->n % ks != 0 || ks > b->gram_ksplit_max) continue;
        if (mf && (split > 1 || !gram_mfma_fits(b->n, b->K, ks))) continue;  // one layout of the MFMA kernel
        b->gram_mfma = mf;
        b->gram_split = split;
        b->gram_ksplit = ks;
        // warm-up, timed: a layout already 1.5 x slower than the best is not timed further
        LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
        LAPH_TRY(gram(v, q, v.X(), v.X(), b->G2 + v.kk(), 1.0));
        LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
        LAPH_HIP(hipEventSynchronize(ln->ev[1]));
        float ms = 0;
        LAPH_HIP(hipEventElapsedTime(&ms, ln->ev[0], ln->ev[1]));
        if (!first_g && ms > 0.75f * best_ms) continue;
        LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
        for (int r = 0; r < 2; ++r) LAPH_TRY(gram(v, q, v.X(), v.X(), b->G2 + v.kk(), 1.0));
        LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
        LAPH_HIP(hipEventSynchronize(ln->ev[1]));
        LAPH_HIP(hipEventElapsedTime(&ms, ln->ev[0], ln->ev[1]));
        if (first_g || ms < best_ms) best_ms = ms, best_split = split, best_ks = ks, best_mfma = mf;
        first_g = false;
      }
    b->gram_split = best_split;
    b->gram_ksplit = best_ks;
    b->gram_mfma = best_mfma;
  }
  LAPH_HIP(hipGraphExecDestroy(ln->graph));  // captured for the last layout timed
  ln->graph = nullptr;
  b->tune_miss_key = key;  // appended once the first solve has chosen the warm-up degree
  if (b->trace)
    std::fprintf(stderr, "[laph] filter block width %d, strip %d, Gram %s, row blocks %d, site chunks %d\n", wbest,
                 sbest, b->gram_mfma ? "MFMA kernel" : "rocBLAS", b->gram_split, b->gram_ksplit);
  return LAPH_OK;
}

// Phase-fixed copy of one timeslice's eigenvectors (file layout, n x nev)
// and its phases: the output staging of laph_batch_write and
// laph_batch_source_shape, allocated on first use.
int output_buffers(laph_batch* b) {
  if (!b->wV) LAPH_HIP(hipMalloc(&b->wV, (size_t)b->n * b->nev * sizeof(double2)));
  return LAPH_OK;
}

// Every lane's solve on its own host thread.
static int solve_lanes(laph_batch* b, bool warm) {
  const int dev = b->p.device;
  std::vector<std::thread> th;
  for (auto& ln : b->lanes)
    th.emplace_back([b, &ln, warm, dev] {
      if (hipSetDevice(dev) != hipSuccess) {
        ln.rc = LAPH_ERR_HIP;
        ln.msg = "lane thread: hipSetDevice failed";
        return;
      }
      const auto l0 = std::chrono::steady_clock::now();
      ln.rc = guarded([&] { return lane_solve(b, &ln, warm); });
      ln.s_solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - l0).count();
      if (ln.rc != LAPH_OK) ln.msg = laph_last_error();
    });
  for (auto& t : th) t.join();
  for (auto& ln : b->lanes)
    if (ln.rc != LAPH_OK) return set_error(ln.rc, "%s", ln.msg.c_str());
  return LAPH_OK;
}

static int solve_impl(laph_batch* b, laph_batch_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  if (!b->tuned) LAPH_TRY(autotune_block_width(b));
  const bool warm = b->o.warm_start && b->have_start;
  b->have_start = false;
  b->solved = false;
  LAPH_HIP(hipStreamSynchronize(b->st));
  const std::vector<int> wc = b->warm_tuned || warm ? std::vector<int>() : warm_candidates(b);
  if (wc.empty()) {
    LAPH_TRY(solve_lanes(b, warm));
  } else {
    // first solve from the start block: one whole solve per warm-up degree
    // (each a valid solve), timed without the host time of first-use graph
    // captures; the fastest is kept, and solved again when it was not the last
    // one run, so that this solve's pairs are those every later solve repeats.
    // One untimed solve at DEG comes first: the first solve of a batch also
    // pays the first-use costs of the dense library kernels (code loading,
    // workspaces), which would otherwise fall on the candidate timed first
    // (at 24^3 they made degree 40, 12% faster once warm, look slower than 32)
    b->warm_deg = wc.back();
    LAPH_TRY(solve_lanes(b, false));
    double best = 0.0;
    int dbest = wc.back();
    bool any = false;
    for (size_t i = 0; i < wc.size(); ++i) {
      b->warm_deg = wc[i];
      const int rc = b->test_fail_warm == wc[i] && i + 1 < wc.size()  // LAPH_TEST_FAIL_WARM (tests)
                         ? set_error(LAPH_ERR_NOCONV, "warm-up degree %d failed (LAPH_TEST_FAIL_WARM)", wc[i])
                         : solve_lanes(b, false);
      if (rc == LAPH_ERR_NOCONV && i + 1 < wc.size()) {  // a candidate above DEG that fails is dropped
        if (b->trace) std::fprintf(stderr, "[laph] warm-up degree %d dropped: %s\n", wc[i], laph_last_error());
        continue;
      }
      LAPH_TRY(rc);
      double sec = 0.0;
      for (auto& ln : b->lanes) sec = std::max(sec, ln.s_solve - ln.s_capture);
      if (b->trace) std::fprintf(stderr, "[laph] warm-up degree %d: %.3f s\n", wc[i], sec);
      if (!any || sec < best) best = sec, dbest = wc[i];
      any = true;
    }
    b->warm_deg = dbest;
    if (dbest != wc.back()) LAPH_TRY(solve_lanes(b, false));
  }
  if (!warm) b->warm_tuned = true;
  if (!b->tune_miss_key.empty() && b->warm_tuned) {
    if (const char* tf = std::getenv("LAPH_TUNE_FILE")) tune_store(b, tf, b->tune_miss_key);
    b->tune_miss_key.clear();
  }
  // fix_phase (eigensystem.cpp:4-30) of every timeslice's nev eigenvectors,
  // in place in B[0]: the solve ends with phase-fixed eigenpairs resident
  LAPH_HIP(launch_fix_phases(b->B[0], b->phase, b->n, b->K, b->nev, b->nts, b->st));
  LAPH_HIP(hipMemcpyAsync(b->h_phase.data(), b->phase, b->h_phase.size() * sizeof(double), hipMemcpyDeviceToHost,
                          b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  b->solved = true;
  int conv = 0, cycles = 0;
  double maxres = 0;
  laph_batch_stats s{};
  for (int t = 0; t < b->nts; ++t) {
    if (b->nlock[t] >= b->nev) ++conv;
    for (int j = 0; j < b->nev; ++j) maxres = std::max(maxres, std::sqrt(b->h_res[(size_t)t * b->K + j]));
  }
  for (auto& ln : b->lanes) {
    cycles = std::max(cycles, ln.cycles);
    s.stencil_vecs += ln.sv32;
    s.stencil64_vecs += ln.sv64;
    s.filter_launches += ln.flaunch;
    s.recoveries += ln.recoveries;
    s.rr_cholqr_fallbacks += ln.rr_fallbacks;
    s.shifted_cholqr += ln.shifted;
    s.subset_cycles += ln.subset_cycles;
    s.filter_bytes += ln.fbytes;
    s.seconds_filter += ln.ms_filter * 1e-3;
    s.seconds_dense += ln.ms_dense * 1e-3;
  }
  s.iterations = cycles;
  s.filter_block = b->bw;
  s.gram_blocks = b->gram_split;
  s.gram_chunks = b->gram_ksplit;
  s.gram_mfma = b->gram_mfma;
  s.layout_from_cache = b->tuned_from_cache;
  s.filter_strip = b->strip;
  s.filter_zrun = b->bf16 == 0 && (b->bw == 128 || b->bw == 64) ? laph::stencil_zrun(b) : 0;
  s.start_window = b->window0;
  s.warmup_degree = b->warm_deg > 0 && b->fdeg_max >= b->warm_deg ? b->warm_deg : b->deg;
  s.converged = conv;
  s.max_residual = maxres;
  s.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (stats) *stats = s;
  if (conv < b->nts)
    return set_error(LAPH_ERR_NOCONV, "laph_batch_solve: %d of %d timeslices converged after %d cycles", conv, b->nts,
                     cycles);
  return LAPH_OK;
}

}  // namespace laph

extern "C" {

int laph_batch_create(laph_batch** out, const laph_params* p, const laph_batch_opts* o) {
  LAPH_CHECK(out && p && o, LAPH_ERR_INVALID, "laph_batch_create: null argument");
  *out = nullptr;
  LAPH_CHECK(p->L >= 3 && p->L <= 128, LAPH_ERR_INVALID, "laph_batch_create: L=%d out of range", p->L);
  LAPH_CHECK(p->deg >= 2, LAPH_ERR_INVALID, "laph_batch_create: DEG=%d < 2", p->deg);
  LAPH_CHECK(p->lambda_l > 0, LAPH_ERR_INVALID, "laph_batch_create: lambda_l must be positive");
  LAPH_CHECK(o->nts >= 1 && o->nev >= 1, LAPH_ERR_INVALID, "laph_batch_create: nts=%d nev=%d", o->nts, o->nev);
  LAPH_CHECK(o->tol > 0, LAPH_ERR_INVALID, "laph_batch_create: tol must be positive");
  LAPH_CHECK(o->filter_passes >= 0 && o->filter_passes <= LAPH_MAX_PASSES, LAPH_ERR_INVALID,
             "laph_batch_create: filter_passes=%d outside 0..%d", o->filter_passes, LAPH_MAX_PASSES);
  LAPH_CHECK(o->warm_start >= 0 && o->warm_start <= 2, LAPH_ERR_INVALID, "laph_batch_create: warm_start=%d outside 0..2",
             o->warm_start);
  LAPH_CHECK(o->filter_storage == 0 || o->filter_storage == 1, LAPH_ERR_INVALID,
             "laph_batch_create: filter_storage=%d (0 FP32, 1 BF16)", o->filter_storage);
  const int n = 3 * p->L * p->L * p->L;
  int K = o->nsub > 0 ? o->nsub : o->nev + std::max(16, o->nev / 4);
  K = (K + 63) / 64 * 64;
  LAPH_CHECK(K > o->nev && K <= n, LAPH_ERR_INVALID, "laph_batch_create: search space %d invalid for n=%d nev=%d", K,
             n, o->nev);
  int strip = 0;  // LAPH_STRIP: fixed sweep strip width
  if (const char* e = std::getenv("LAPH_STRIP")) {
    strip = std::atoi(e);
    LAPH_CHECK(strip >= 1 && p->L % strip == 0, LAPH_ERR_INVALID, "LAPH_STRIP=%s must divide L=%d", e, p->L);
  }
  // LAPH_ZRUN: z-sites per wavefront of the y-row work-group filter step (0: one site per wavefront)
  int zrun = p->L % 4 != 0 ? 0 : (p->L % 8 == 0 ? 8 : 4);
  if (const char* e = std::getenv("LAPH_ZRUN")) {
    zrun = std::atoi(e);
    LAPH_CHECK(zrun == 0 || ((zrun == 1 || zrun == 2 || zrun == 4 || zrun == 8) && p->L % zrun == 0 && p->L % 4 == 0),
               LAPH_ERR_INVALID, "LAPH_ZRUN=%s: 0, or 1, 2, 4 or 8 dividing L=%d (a multiple of 4)", e, p->L);
  }
  // automatic passes: one pass per cycle of the degree the window allows (up
  // to 128); an explicit pass count: the batch's degree, up to that many passes.
  // The amplification bound is what limits a cycle at 24^3 and 32^3 (about
  // 72); at 48^3 it allows about 100, and a ceiling of 72 cost a fifth cycle
  // there (7.00 against 7.36 timeslices/s)
  int fdeg_max = o->filter_passes > 0 ? 0 : 128;
  if (const char* e = std::getenv("LAPH_FILTER_DEGREE")) {
    fdeg_max = std::atoi(e);
    LAPH_CHECK(fdeg_max == 0 || (fdeg_max >= 4 && fdeg_max <= 256), LAPH_ERR_INVALID,
               "LAPH_FILTER_DEGREE=%s: 0, or a highest per-cycle degree in 4..256", e);
  }
  int ndev = 0;
  LAPH_HIP(hipGetDeviceCount(&ndev));
  LAPH_CHECK(p->device >= 0 && p->device < ndev, LAPH_ERR_INVALID, "laph_batch_create: no device %d", p->device);
  return laph::guarded([&]() -> int {
    laph_batch* b = new laph_batch();
    b->p = *p;
    b->o = *o;
    b->V3 = p->L * p->L * p->L;
    b->n = n;
    b->K = K;
    b->nts = o->nts;
    b->nev = o->nev;
    b->deg = p->deg;
    b->per_ts = (size_t)n * K;
    if (const char* e = std::getenv("LAPH_TRACE")) {
      b->trace_time = true;
      b->trace = std::strcmp(e, "time") != 0;
    }
    if (const char* e = std::getenv("LAPH_WARM")) b->warm_cycles = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("LAPH_WARM_DEG")) {
      b->warm_deg = std::max(0, std::atoi(e)) / 4 * 4;
      b->warm_deg_fixed = true;
    }
    b->force_shift = std::getenv("LAPH_SHIFTED_CHOLQR") != nullptr;
    if (const char* e = std::getenv("LAPH_QR_CYCLES")) b->qr_cycles = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("LAPH_TEST_POISON")) b->poison_ts = std::atoi(e);
    if (const char* e = std::getenv("LAPH_TEST_FAIL_WARM")) b->test_fail_warm = std::atoi(e);
    if (const char* e = std::getenv("LAPH_FILTER_FORM")) {
      b->prod = std::atoi(e) != 0;
      b->qform = std::atoi(e) == 2;
    }
    b->fdeg_max = fdeg_max;
    if (b->fdeg_max > 0) b->prod = b->qform = true;  // one correction pass per cycle: product over q's roots
    // cycle_degree returns a multiple of 4 in [4, fdeg_max]
    b->qrot.resize(std::max(p->deg, b->fdeg_max) + 1);
    for (int d = 2; d < (int)b->qrot.size(); ++d)
      if (d == p->deg || (b->fdeg_max > 0 && d % 4 == 0)) b->qrot[d] = laph::correction_root_rotations(d);
    if (const char* e = std::getenv("LAPH_START_RR")) b->start_rr = std::atoi(e) != 0;
    if (const char* e = std::getenv("LAPH_TRSM")) b->use_trsm = std::atoi(e) != 0;
    if (const char* e = std::getenv("LAPH_AMP_CAP")) b->amp_cap = std::max(10.0, std::atof(e));
    if (const char* e = std::getenv("LAPH_WARM_AMP_CAP")) b->warm_amp_cap = std::max(10.0, std::atof(e));
    if (const char* e = std::getenv("LAPH_DEGREE_SAFETY")) b->deg_safety = std::max(0.0, std::atof(e));
    b->zrun = zrun;
    // first-cycle window: 3 x the free Laplacian's K-th level (measured best
    // between 2.5 and 3.5 x at 24^3, 32^3 and 48^3), at most the usual cap
    b->window0 = std::min(3.0 * laph::free_level(p->L, K), 0.75 * p->lambda_l);
    if (const char* e = std::getenv("LAPH_WINDOW0")) {
      const double w = std::atof(e);
      if (w > 0.0) b->window0 = std::min(w, 0.75 * p->lambda_l);
    }
    if (const char* e = std::getenv("LAPH_WARMUP_WINDOW_CAP")) b->warmup_cap = std::atoi(e) != 0;
    if (const char* e = std::getenv("LAPH_BLOCK_DEGREE")) b->block_degrees = std::atoi(e) != 0;
    if (const char* e = std::getenv("LAPH_DENSE_SUBSET")) b->dense_subset = std::atoi(e) != 0;
    b->roots = laph::leja_roots(p->deg);
    b->strip = strip > 0 ? strip : p->L;
    b->strip_fixed = strip > 0;
    b->bf16 = o->filter_storage == 1;
    b->esz = b->bf16 ? 4 : 8;
    if (const char* e = std::getenv("LAPH_EIG")) b->eig_mode = std::atoi(e);
    if (const char* e = std::getenv("LAPH_GRAM_SPLIT")) {  // "rows" or "rows,chunks"
      b->gram_split = std::max(1, std::min(8, std::atoi(e)));
      const char* c = std::strchr(e, ',');
      b->gram_ksplit = c ? std::max(1, std::min(b->gram_ksplit_max, std::atoi(c + 1))) : 1;
      if (b->n % b->gram_ksplit != 0) b->gram_ksplit = 1;
      b->gram_split_fixed = true;
    }
    if (const char* e = std::getenv("LAPH_GRAM_MFMA")) {
      b->gram_mfma = std::atoi(e) != 0;
      b->gram_split_fixed = true;
    }
    if (const char* e = std::getenv("LAPH_NET_TARGET")) b->net_target_all = std::atoi(e) != 0;
    b->nlock.assign(b->nts, 0);
    b->h_mu.assign((size_t)b->nts * K, 0.0);
    b->h_res.assign((size_t)b->nts * K, 0.0);
    b->h_phase.assign((size_t)b->nts * o->nev, 0.0);
    auto fail = [&](int rc) {
      laph::batch_free(b);
      return rc;
    };
    if (hipSetDevice(p->device) != hipSuccess || hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_batch_create: stream creation failed"));
    int rc = laph::batch_alloc(b);
    if (rc != LAPH_OK) return fail(rc);
    *out = b;
    return LAPH_OK;
  });
}

int laph_batch_destroy(laph_batch* b) {
  laph::batch_free(b);
  return LAPH_OK;
}

int laph_batch_set_links(laph_batch* b, int ts, const double* links) {
  LAPH_CHECK(b && links && ts >= 0 && ts < b->nts, LAPH_ERR_INVALID, "laph_batch_set_links: bad arguments");
  LAPH_HIP(hipSetDevice(b->p.device));
  const size_t lk = (size_t)b->V3 * 27;
  LAPH_HIP(hipMemcpyAsync(b->U64 + lk * ts, links, lk * sizeof(double2), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(laph::launch_convert_links32(b->U64 + lk * ts, b->U32 + lk * ts, lk, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  b->solved = false;
  return LAPH_OK;
}

int laph_batch_set_timeslice(laph_batch* b, int ts, const double* cfg, double a1, double a2, int iter) {
  LAPH_CHECK(b && cfg && ts >= 0 && ts < b->nts && iter >= 0, LAPH_ERR_INVALID,
             "laph_batch_set_timeslice: bad arguments");
  LAPH_HIP(hipSetDevice(b->p.device));
  const size_t lk = (size_t)b->V3 * 27;
  std::vector<double> host(lk * 2);
  for (int s = 0; s < b->V3; ++s)  // timeslice.cpp:46-73
    for (int dir = 0; dir < 3; ++dir)
      std::memcpy(&host[((size_t)s * 3 + dir) * 18], cfg + (size_t)s * 72 + (dir + 1) * 18, 18 * sizeof(double));
  double2* U = b->U64 + lk * ts;
  LAPH_HIP(hipMemcpyAsync(U, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  if (iter > 0) {
    if (!b->dec) LAPH_HIP(hipMalloc(&b->dec, (size_t)b->V3 * 54 * sizeof(double2)));
    if (!b->Unew) LAPH_HIP(hipMalloc(&b->Unew, lk * sizeof(double2)));
    for (int it = 0; it < iter; ++it) {
      LAPH_HIP(laph::launch_hyp_level1(U, b->dec, b->p.L, a2, b->st));
      LAPH_HIP(laph::launch_hyp_level2(U, b->dec, b->Unew, b->p.L, a1, b->st));
      LAPH_HIP(hipMemcpyAsync(U, b->Unew, lk * sizeof(double2), hipMemcpyDeviceToDevice, b->st));
    }
  }
  LAPH_HIP(laph::launch_convert_links32(U, b->U32 + lk * ts, lk, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  b->solved = false;
  return LAPH_OK;
}

int laph_batch_solve(laph_batch* b, laph_batch_stats* stats) {
  LAPH_CHECK(b, LAPH_ERR_INVALID, "laph_batch_solve: null batch");
  LAPH_HIP(hipSetDevice(b->p.device));
  return laph::guarded([&] { return laph::solve_impl(b, stats); });
}

int laph_batch_time_filter(laph_batch* b, int reps, double* ms_per_launch, double* bytes_per_launch) {
  LAPH_CHECK(b && reps >= 1 && ms_per_launch && bytes_per_launch, LAPH_ERR_INVALID,
             "laph_batch_time_filter: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_time_filter: solve the batch first (coefficient tables)");
  LAPH_HIP(hipSetDevice(b->p.device));
  return laph::guarded([&]() -> int {
    laph_lane* ln = &b->lanes[0];
    ln->m = 0;  // the batch's degree
    const laph::View v{b, ln};
    const int K = b->K, nt = ln->nt;
    // every column of every timeslice of lane 0 active (also in every step of the correction pass)
    std::vector<int> zero(nt, 0);
    LAPH_HIP(hipMemcpyAsync(b->first_blk + ln->t0, zero.data(), nt * sizeof(int), hipMemcpyHostToDevice, ln->st));
    for (int s = 1; b->fdeg_max > 0 && s < b->deg; ++s)
      LAPH_HIP(hipMemcpyAsync(b->first_blk_steps + (size_t)(s - 1) * b->nts + ln->t0, zero.data(), nt * sizeof(int),
                              hipMemcpyHostToDevice, ln->st));
    void* w = nullptr;
    LAPH_TRY(laph::run_filter(v, 1, &w));  // warm-up (and capture if needed)
    LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
    for (int r = 0; r < reps; ++r) LAPH_TRY(laph::run_filter(v, 1, &w));
    LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
    LAPH_HIP(hipEventSynchronize(ln->ev[1]));
    float ms = 0;
    LAPH_HIP(hipEventElapsedTime(&ms, ln->ev[0], ln->ev[1]));
    const int steps = b->deg - 1;
    double bytes = 0;
    for (int s = 1; s < b->deg; ++s) {
      const int touches = s == 1 || b->qform ? 2 : (s == 2 || b->prod ? 3 : 4);
      bytes += (double)nt * b->V3 * ((double)K * touches * 3.0 * b->esz + (K / b->bw) * 216.0);
    }
    *ms_per_launch = ms / ((double)reps * steps);
    *bytes_per_launch = bytes / steps;
    return LAPH_OK;
  });
}

int laph_batch_time_gram(laph_batch* b, int reps, double* ms_per_gram, double* flops_per_gram) {
  LAPH_CHECK(b && reps >= 1 && ms_per_gram && flops_per_gram, LAPH_ERR_INVALID, "laph_batch_time_gram: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_time_gram: solve the batch first");
  LAPH_HIP(hipSetDevice(b->p.device));
  return laph::guarded([&]() -> int {
    laph_lane* ln = &b->lanes[0];
    const laph::View v{b, ln};
    const laph::Act q{0, b->K};
    double2* G2 = b->G2 + v.kk();
    LAPH_TRY(laph::gram(v, q, v.X(), v.X(), G2, 1.0));  // warm-up
    LAPH_HIP(hipEventRecord(ln->ev[0], ln->st));
    for (int r = 0; r < reps; ++r) LAPH_TRY(laph::gram(v, q, v.X(), v.X(), G2, 1.0));
    LAPH_HIP(hipEventRecord(ln->ev[1], ln->st));
    LAPH_HIP(hipEventSynchronize(ln->ev[1]));
    float ms = 0;
    LAPH_HIP(hipEventElapsedTime(&ms, ln->ev[0], ln->ev[1]));
    *ms_per_gram = ms / reps;
    *flops_per_gram = laph::gram_flops(b, b->K, ln->nt);
    return LAPH_OK;
  });
}

int laph_batch_get(laph_batch* b, int ts, double* evals, double* V_fix, double* phase, double* residuals) {
  LAPH_CHECK(b && ts >= 0 && ts < b->nts && evals, LAPH_ERR_INVALID, "laph_batch_get: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_get: batch not solved");
  LAPH_HIP(hipSetDevice(b->p.device));
  const int K = b->K, nev = b->nev;
  for (int j = 0; j < nev; ++j) evals[j] = -b->h_mu[(size_t)ts * K + j];
  if (residuals)
    for (int j = 0; j < nev; ++j) residuals[j] = std::sqrt(b->h_res[(size_t)ts * K + j]);
  if (phase)
    for (int j = 0; j < nev; ++j) phase[j] = b->h_phase[(size_t)ts * nev + j];
  if (V_fix) {  // the solve left them phase-fixed in B[0]: only the layout changes, in the
                // batch's output buffer (allocated once, shared with laph_batch_write)
    LAPH_TRY(laph::output_buffers(b));
    LAPH_HIP(laph::launch_extract_cols(b->B[0] + (size_t)ts * b->per_ts, b->wV, b->n, K, nev, b->st));
    LAPH_HIP(hipMemcpyAsync(V_fix, b->wV, (size_t)b->n * nev * sizeof(double2), hipMemcpyDeviceToHost, b->st));
    LAPH_HIP(hipStreamSynchronize(b->st));
  }
  return LAPH_OK;
}

int laph_batch_get_device(laph_batch* b, int ts, double* d_V, double* phase) {
  LAPH_CHECK(b && ts >= 0 && ts < b->nts && d_V, LAPH_ERR_INVALID, "laph_batch_get_device: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_get_device: batch not solved");
  LAPH_HIP(hipSetDevice(b->p.device));
  LAPH_HIP(laph::launch_extract_cols(b->B[0] + (size_t)ts * b->per_ts, reinterpret_cast<double2*>(d_V), b->n, b->K,
                                     b->nev, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  if (phase)
    for (int j = 0; j < b->nev; ++j) phase[j] = b->h_phase[(size_t)ts * b->nev + j];
  return LAPH_OK;
}

int laph_batch_write(laph_batch* b, int ts, const char* dir, int config, int t, double* trace_sum) {
  LAPH_CHECK(b && ts >= 0 && ts < b->nts && dir, LAPH_ERR_INVALID, "laph_batch_write: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_write: batch not solved");
  LAPH_HIP(hipSetDevice(b->p.device));
  return laph::guarded([&]() -> int {
    const int n = b->n, nev = b->nev, K = b->K;
    const size_t CH = (size_t)64 << 20;  // staging chunk
    LAPH_TRY(laph::output_buffers(b));
    if (!b->wG) {
      LAPH_HIP(hipMalloc(&b->wG, (size_t)nev * nev * sizeof(double2)));
      LAPH_CHECK(rocblas_create_handle(&b->wrb) == rocblas_status_success &&
                     rocblas_set_stream(b->wrb, b->st) == rocblas_status_success,
                 LAPH_ERR_HIP, "laph_batch_write: rocBLAS handle");
      for (int i = 0; i < 2; ++i) {
        LAPH_HIP(hipHostMalloc(&b->pin[i], CH, hipHostMallocDefault));
        LAPH_HIP(hipEventCreateWithFlags(&b->wev[i], hipEventDisableTiming));
      }
    }
    // phase-fixed eigenvectors (fixed in place by the solve, eigensystem.cpp:4-30)
    // in the file layout, and their Gram V^H V
    LAPH_HIP(laph::launch_extract_cols(b->B[0] + (size_t)ts * b->per_ts, b->wV, n, K, nev, b->st));
    const laph::cplx one(1.0, 0.0), zero(0.0, 0.0);
    LAPH_RB(rocblas_zgemm(b->wrb, rocblas_operation_conjugate_transpose, rocblas_operation_none, nev, nev, n, &one,
                          laph::zc(b->wV), n, laph::zc(b->wV), n, &zero, laph::zc(b->wG), nev));
    std::vector<double2> G((size_t)nev * nev);
    std::vector<double> ph(nev), ev(nev);
    LAPH_HIP(hipMemcpyAsync(G.data(), b->wG, G.size() * sizeof(double2), hipMemcpyDeviceToHost, b->st));
    std::copy(b->h_phase.begin() + (size_t)ts * nev, b->h_phase.begin() + (size_t)(ts + 1) * nev, ph.begin());
    LAPH_HIP(hipStreamSynchronize(b->st));
    double tr[2] = {0, 0}, sm[2] = {0, 0};
    for (int i = 0; i < nev; ++i) {
      tr[0] += G[(size_t)i * nev + i].x;
      tr[1] += G[(size_t)i * nev + i].y;
      for (int j = 0; j < nev; ++j) {
        sm[0] += G[(size_t)i * nev + j].x;
        sm[1] += G[(size_t)i * nev + j].y;
      }
    }
    if (trace_sum) {
      trace_sum[0] = tr[0];
      trace_sum[1] = tr[1];
      trace_sum[2] = sm[0];
      trace_sum[3] = sm[1];
    }
    const double eps = 10e-10;  // read_write.cpp:24-31 (one-sided, as there)
    LAPH_CHECK(!(std::fabs(tr[0]) - nev > eps || std::fabs(tr[1]) > eps || std::fabs(sm[0]) - nev > eps ||
                 std::fabs(sm[1]) > eps),
               LAPH_ERR_TRACE, "timeslice %d: eigenvectors damaged, tr(V^H V) = (%.12g, %.3g), sum = (%.12g, %.3g)", t,
               tr[0], tr[1], sm[0], sm[1]);
    // eigenvectors: chunk i+1 copies while chunk i is written
    char file[4096];
    std::snprintf(file, sizeof file, "%s/eigenvectors.%04d.%03d", dir, config, t);
    FILE* f = std::fopen(file, "wb");
    LAPH_CHECK(f, LAPH_ERR_IO, "could not create %s", file);
    const size_t total = (size_t)n * nev * sizeof(double2);
    const char* src = reinterpret_cast<const char*>(b->wV);
    const size_t nch = (total + CH - 1) / CH;
    hipError_t e = hipMemcpyAsync(b->pin[0], src, std::min(CH, total), hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess) e = hipEventRecord(b->wev[0], b->st);
    bool ok = e == hipSuccess;
    for (size_t i = 0; ok && i < nch; ++i) {
      const size_t len = std::min(CH, total - i * CH);
      e = hipEventSynchronize(b->wev[i % 2]);
      if (e == hipSuccess && i + 1 < nch) {
        e = hipMemcpyAsync(b->pin[(i + 1) % 2], src + (i + 1) * CH, std::min(CH, total - (i + 1) * CH),
                           hipMemcpyDeviceToHost, b->st);
        if (e == hipSuccess) e = hipEventRecord(b->wev[(i + 1) % 2], b->st);
      }
      ok = e == hipSuccess && std::fwrite(b->pin[i % 2], 1, len, f) == len;
    }
    const bool closed = std::fclose(f) == 0;
    (void)hipStreamSynchronize(b->st);  // no copy left in flight into the staging chunks
    LAPH_HIP(e);
    LAPH_CHECK(ok && closed, LAPH_ERR_IO, "write incomplete: %s", file);
    for (int j = 0; j < nev; ++j) ev[j] = -b->h_mu[(size_t)ts * K + j];
    LAPH_TRY(laph_write_eigenvalues_bin(dir, "eigenvalues", config, t, nev, ev.data()));
    return laph_write_eigenvalues_bin(dir, "phases", config, t, nev, ph.data());
  });
}

}  // extern "C"
