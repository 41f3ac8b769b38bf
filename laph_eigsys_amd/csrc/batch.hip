// Mixed-precision batched eigensolver: every timeslice of a configuration on
// one device, in lock-step, each with its own window and lock count.
//
// Method: Chebyshev-filtered subspace iteration (Zhou & Saad, J. Comput.
// Phys. 219 (2006)) on D with a K-column block per timeslice (K >= nev+guard,
// multiple of 64), in residual-correction form so the FP32 filter acts on a
// shrinking quantity.  For a Ritz pair (x, theta) of B = c D + a with
// r = (B - theta) x, the degree-m Chebyshev filter satisfies exactly
//     T_m(B) x = T_m(theta) x + e_m,
//     e_0 = 0, e_1 = r, e_{k+1} = 2 B e_k - e_{k-1} + 2 T_k(theta) r,
// so only e (of size |r|) is computed in FP32; x stays FP64 and the FP32
// rounding error scales with the residual.  The recurrence is carried in the
// scaled form f_k = e_k / s_k, s_k = max(|T_k(theta)|, 1), which keeps every
// column O(|r|) and turns each step into
//     f_{k+1} = a1 (B f_k) + a2 r + b f_{k-1}     (per-column a1, a2, b)
// and the filtered vector into y = (T_m(theta)/s_m) x + f_m.
//
// Per cycle (all timeslices batched):
//   1. FP32 filter, m-1 fused stencil steps (hipGraph, one launch per step
//      covering every active vector block of every timeslice)
//   2. Y = gamma X + F (FP64)
//   3. CholQR2 in FP64 (Gram zgemm, zpotrf, ztrsm; all strided-batched)
//   4. AY = D Y (FP64 blocked stencil)
//   5. Rayleigh-Ritz in FP64: H = Y^H AY, zheevd, X = Y Q, AX = AY Q
//   6. residuals R = AX - X Lambda in FP64; lock the leading columns with
//      |R_j| <= tol; window lower bound = largest Ritz value of -D.
// The window map is B = c_t D + a_t with c_t = 2/(lambda_l - lambda_c,t),
// a_t = 1 + c_t lambda_c,t (shell_matop.cpp:49-50) and lambda_c,t adapted per
// timeslice; eigenvalues come straight from the FP64 Rayleigh-Ritz, so
// recover_spectrum's inversion is not needed on this path.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "kernels.h"
#include "kernels_block.h"

#define LAPH_RB(call)                                                                               \
  do {                                                                                              \
    rocblas_status s_ = (call);                                                                     \
    if (s_ != rocblas_status_success)                                                               \
      return laph::set_error(LAPH_ERR_HIP, "%s failed: %s", #call, rocblas_status_to_string(s_));   \
  } while (0)

struct laph_batch {
  laph_params p{};
  laph_batch_opts o{};
  int V3 = 0, n = 0, K = 0, nts = 0, nev = 0, deg = 0;
  size_t per_ts = 0;  // n*K
  hipStream_t st = nullptr;
  rocblas_handle rb = nullptr;
  // links
  double2* U64 = nullptr;  // [nts][V3*27]
  float2* U32 = nullptr;
  double2* dec = nullptr;   // HYP scratch for one timeslice
  double2* Unew = nullptr;
  // blocks
  double2* B[3] = {nullptr, nullptr, nullptr};  // rotating roles X, AX, T
  int iX = 0, iAX = 1, iT = 2;
  float2* R32 = nullptr;
  float2* P[3] = {nullptr, nullptr, nullptr};
  // small per-timeslice data
  double2* G = nullptr;      // [nts][K*K]
  double* mu = nullptr;      // [nts][K]
  double* E = nullptr;       // [nts][K]
  rocblas_int* info = nullptr;
  double* nrm2 = nullptr;    // [nts][K]
  double* lam = nullptr;     // [nts][K]
  double* cts = nullptr;     // [nts]
  double* gam = nullptr;     // [nts][K]
  float4* coef = nullptr;    // [deg-1][nts][K]
  float2* cab = nullptr;     // [nts]
  int* first_blk = nullptr;  // [nts]
  // state
  std::vector<int> nlock;
  std::vector<double> h_mu, h_res;
  bool solved = false;
  hipGraphExec_t filter_graph = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace laph {

static int batch_alloc(laph_batch* b) {
  const size_t lk = (size_t)b->V3 * 27;
  const size_t nb = (size_t)b->nts * b->per_ts;
  LAPH_HIP(hipMalloc(&b->U64, lk * b->nts * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->U32, lk * b->nts * sizeof(float2)));
  for (int i = 0; i < 3; ++i) LAPH_HIP(hipMalloc(&b->B[i], nb * sizeof(double2)));
  LAPH_HIP(hipMalloc(&b->R32, nb * sizeof(float2)));
  for (int i = 0; i < 3; ++i) LAPH_HIP(hipMalloc(&b->P[i], nb * sizeof(float2)));
  const size_t kk = (size_t)b->nts * b->K * b->K;
  LAPH_HIP(hipMalloc(&b->G, kk * sizeof(double2)));
  const size_t nk = (size_t)b->nts * b->K;
  LAPH_HIP(hipMalloc(&b->mu, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->E, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->info, b->nts * sizeof(rocblas_int)));
  LAPH_HIP(hipMalloc(&b->nrm2, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->lam, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->cts, b->nts * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->gam, nk * sizeof(double)));
  LAPH_HIP(hipMalloc(&b->coef, (size_t)std::max(1, b->deg - 1) * nk * sizeof(float4)));
  LAPH_HIP(hipMalloc(&b->cab, b->nts * sizeof(float2)));
  LAPH_HIP(hipMalloc(&b->first_blk, b->nts * sizeof(int)));
  for (int i = 0; i < 4; ++i) LAPH_HIP(hipEventCreate(&b->ev[i]));
  return LAPH_OK;
}

static void batch_free(laph_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->p.device);
  if (b->st) (void)hipStreamSynchronize(b->st);
  if (b->filter_graph) (void)hipGraphExecDestroy(b->filter_graph);
  void* bufs[] = {b->U64, b->U32, b->dec, b->Unew, b->B[0], b->B[1], b->B[2], b->R32, b->P[0], b->P[1],
                  b->P[2], b->G, b->mu, b->E, b->info, b->nrm2, b->lam, b->cts, b->gam, b->coef, b->cab,
                  b->first_blk};
  for (void* q : bufs)
    if (q) (void)hipFree(q);
  for (auto e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->rb) rocblas_destroy_handle(b->rb);
  if (b->st) (void)hipStreamDestroy(b->st);
  delete b;
}

using cplx = rocblas_double_complex;
static inline cplx* zc(double2* p) { return reinterpret_cast<cplx*>(p); }

// CholQR on every timeslice: Yc (K x n column-major = row-major n x K)
// <- U^{-H} Yc with Yc Yc^H = U^H U.
static int cholqr(laph_batch* b, double2* Y) {
  const int K = b->K, n = b->n, nts = b->nts;
  const cplx one(1.0, 0.0), zero(0.0, 0.0);
  LAPH_RB(rocblas_zgemm_strided_batched(b->rb, rocblas_operation_none, rocblas_operation_conjugate_transpose, K, K,
                                        n, &one, zc(Y), K, (rocblas_stride)b->per_ts, zc(Y), K,
                                        (rocblas_stride)b->per_ts, &zero, zc(b->G), K, (rocblas_stride)K * K, nts));
  LAPH_RB(rocsolver_zpotrf_strided_batched(b->rb, rocblas_fill_upper, K, zc(b->G), K, (rocblas_stride)K * K,
                                           b->info, nts));
  std::vector<rocblas_int> info(nts);
  LAPH_HIP(hipMemcpyAsync(info.data(), b->info, nts * sizeof(rocblas_int), hipMemcpyDeviceToHost, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  for (int t = 0; t < nts; ++t)
    LAPH_CHECK(info[t] == 0, LAPH_ERR_NOCONV, "CholQR: Gram matrix of timeslice %d not positive definite (info %d)",
               t, (int)info[t]);
  LAPH_RB(rocblas_ztrsm_strided_batched(b->rb, rocblas_side_left, rocblas_fill_upper,
                                        rocblas_operation_conjugate_transpose, rocblas_diagonal_non_unit, K, n, &one,
                                        zc(b->G), K, (rocblas_stride)K * K, zc(Y), K, (rocblas_stride)b->per_ts, nts));
  return LAPH_OK;
}

// Rayleigh-Ritz: with Y orthonormal in B[iX] and AY in B[iAX]:
// M = -(AYc Yc^H) = -conj(H); M = W diag(mu) W^H (mu ascending = eigenvalues
// of -D, nearest 0 first); X <- Y conj(W) i.e. Xc = W^H Yc; same for AX.
static int rayleigh_ritz(laph_batch* b) {
  const int K = b->K, n = b->n, nts = b->nts;
  const cplx one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
  double2* Y = b->B[b->iX];
  double2* AY = b->B[b->iAX];
  double2* T = b->B[b->iT];
  LAPH_RB(rocblas_zgemm_strided_batched(b->rb, rocblas_operation_none, rocblas_operation_conjugate_transpose, K, K,
                                        n, &mone, zc(AY), K, (rocblas_stride)b->per_ts, zc(Y), K,
                                        (rocblas_stride)b->per_ts, &zero, zc(b->G), K, (rocblas_stride)K * K, nts));
  LAPH_RB(rocsolver_zheevd_strided_batched(b->rb, rocblas_evect_original, rocblas_fill_upper, K, zc(b->G), K,
                                           (rocblas_stride)K * K, b->mu, K, b->E, K, b->info, nts));
  // T <- W^H Y  (new X), then Y's buffer <- W^H AY (new AX)
  LAPH_RB(rocblas_zgemm_strided_batched(b->rb, rocblas_operation_conjugate_transpose, rocblas_operation_none, K, n,
                                        K, &one, zc(b->G), K, (rocblas_stride)K * K, zc(Y), K,
                                        (rocblas_stride)b->per_ts, &zero, zc(T), K, (rocblas_stride)b->per_ts, nts));
  LAPH_RB(rocblas_zgemm_strided_batched(b->rb, rocblas_operation_conjugate_transpose, rocblas_operation_none, K, n,
                                        K, &one, zc(b->G), K, (rocblas_stride)K * K, zc(AY), K,
                                        (rocblas_stride)b->per_ts, &zero, zc(Y), K, (rocblas_stride)b->per_ts, nts));
  // roles: X <- T, AX <- old Y buffer, T <- old AY buffer
  const int oX = b->iX, oAX = b->iAX, oT = b->iT;
  b->iX = oT;
  b->iAX = oX;
  b->iT = oAX;
  std::vector<rocblas_int> info(nts);
  LAPH_HIP(hipMemcpyAsync(info.data(), b->info, nts * sizeof(rocblas_int), hipMemcpyDeviceToHost, b->st));
  LAPH_HIP(hipMemcpyAsync(b->h_mu.data(), b->mu, b->h_mu.size() * sizeof(double), hipMemcpyDeviceToHost, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  for (int t = 0; t < nts; ++t)
    LAPH_CHECK(info[t] == 0, LAPH_ERR_NOCONV, "Rayleigh-Ritz: zheevd failed on timeslice %d (info %d)", t,
               (int)info[t]);
  return LAPH_OK;
}

// Window per timeslice from the Ritz values, residuals, lock counts.
// Returns the number of timeslices still unconverged.
static int residuals_and_window(laph_batch* b, std::vector<double>& lc, int* n_active) {
  const int K = b->K, nts = b->nts;
  const double lamL = b->p.lambda_l;
  std::vector<double> hl((size_t)nts * K), hc(nts);
  for (int t = 0; t < nts; ++t) {
    // lower end of the damped interval: largest Ritz value of -D, kept
    // well below lambda_l so the amplification stays inside FP32 range
    lc[t] = std::min(b->h_mu[(size_t)t * K + K - 1], 0.75 * lamL);
    hc[t] = 2.0 / (lamL - lc[t]);
    for (int j = 0; j < K; ++j) hl[(size_t)t * K + j] = -b->h_mu[(size_t)t * K + j];
  }
  LAPH_HIP(hipMemcpyAsync(b->lam, hl.data(), hl.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(hipMemcpyAsync(b->cts, hc.data(), hc.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(launch_residual(b->B[b->iX], b->B[b->iAX], b->lam, b->cts, b->R32, b->nrm2, b->n, K, nts, b->st));
  LAPH_HIP(hipMemcpyAsync(b->h_res.data(), b->nrm2, b->h_res.size() * sizeof(double), hipMemcpyDeviceToHost, b->st));
  LAPH_HIP(hipStreamSynchronize(b->st));
  int active = 0;
  for (int t = 0; t < nts; ++t) {
    int nl = 0;
    while (nl < K && std::sqrt(b->h_res[(size_t)t * K + nl]) <= b->o.tol) ++nl;
    b->nlock[t] = nl;
    if (nl < b->nev) ++active;
  }
  *n_active = active;
  return LAPH_OK;
}

// Per-column coefficients of the m-1 filter steps and the combine factor.
static int upload_coefficients(laph_batch* b, const std::vector<double>& lc) {
  const int K = b->K, nts = b->nts, m = b->deg;
  const double lamL = b->p.lambda_l;
  std::vector<float4> co((size_t)(m - 1) * nts * K);
  std::vector<double> ga((size_t)nts * K);
  std::vector<float2> cab(nts);
  std::vector<int> fb(nts);
  std::vector<double> tau(m + 1), s(m + 1);
  for (int t = 0; t < nts; ++t) {
    const double c = 2.0 / (lamL - lc[t]);
    const double a = 1.0 + c * lc[t];
    cab[t] = make_float2((float)c, (float)(a - 6.0 * c));
    const bool done = b->nlock[t] >= b->nev;
    fb[t] = done ? K / 64 : b->nlock[t] / 64;
    for (int j = 0; j < K; ++j) {
      const double th = 1.0 + c * (lc[t] - b->h_mu[(size_t)t * K + j]);  // Ritz value in B
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
        co[((size_t)(st - 1) * nts + t) * K + j] = q;
      }
      ga[(size_t)t * K + j] = tau[m] / s[m];
    }
  }
  LAPH_HIP(hipMemcpyAsync(b->coef, co.data(), co.size() * sizeof(float4), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(hipMemcpyAsync(b->gam, ga.data(), ga.size() * sizeof(double), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(hipMemcpyAsync(b->cab, cab.data(), cab.size() * sizeof(float2), hipMemcpyHostToDevice, b->st));
  LAPH_HIP(hipMemcpyAsync(b->first_blk, fb.data(), fb.size() * sizeof(int), hipMemcpyHostToDevice, b->st));
  return LAPH_OK;
}

// The m-1 FP32 filter steps; buffers rotate R32 -> P0 -> P1 -> P2 -> P0 ...
// Returns the buffer holding f_m.
static int enqueue_filter(laph_batch* b, float2** fm) {
  const int m = b->deg, K = b->K, nts = b->nts, L = b->p.L;
  const size_t step = (size_t)nts * K;
  for (int st = 1; st < m; ++st) {
    const float2* in = st == 1 ? b->R32 : b->P[(st - 2) % 3];
    const float2* old = st >= 3 ? b->P[(st - 3) % 3] : nullptr;
    float2* out = b->P[(st - 1) % 3];
    LAPH_HIP(launch_filter_step32(b->U32, in, b->R32, old, out, b->coef + (st - 1) * step, b->cab, b->first_blk,
                                  L, K, nts, b->st));
  }
  *fm = b->P[(m - 2) % 3];
  return LAPH_OK;
}

static int run_filter(laph_batch* b, float2** fm) {
  if (!b->filter_graph) {
    hipGraph_t g;
    LAPH_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_filter(b, fm);
    hipError_t e = hipStreamEndCapture(b->st, &g);
    LAPH_TRY(rc);
    LAPH_HIP(e);
    LAPH_HIP(hipGraphInstantiate(&b->filter_graph, g, nullptr, nullptr, 0));
    LAPH_HIP(hipGraphDestroy(g));
  }
  *fm = b->P[(b->deg - 2) % 3];
  LAPH_HIP(hipGraphLaunch(b->filter_graph, b->st));
  return LAPH_OK;
}

__global__ void k_combine_gamma(double2* __restrict__ X, const float2* __restrict__ F, const double* __restrict__ gam,
                                size_t per_ts, int K, const int* __restrict__ first_blk) {
  const int ts = blockIdx.y;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts) return;
  const int col = (int)(i % K);
  if (col < first_blk[ts] * 64) return;
  const size_t idx = (size_t)ts * per_ts + i;
  const double g = gam[(size_t)ts * K + col];
  const float2 f = F[idx];
  const double2 x = X[idx];
  X[idx] = make_double2(g * x.x + (double)f.x, g * x.y + (double)f.y);
}

static int solve_impl(laph_batch* b, laph_batch_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  const int K = b->K, nts = b->nts, L = b->p.L;
  hipStream_t st = b->st;
  b->iX = 0;
  b->iAX = 1;
  b->iT = 2;
  // random start, orthonormalised twice, first Rayleigh-Ritz
  LAPH_HIP(launch_random_block(b->B[b->iX], (size_t)nts * b->per_ts, b->o.seed, st));
  LAPH_TRY(cholqr(b, b->B[b->iX]));
  LAPH_TRY(cholqr(b, b->B[b->iX]));
  LAPH_HIP(launch_lap_block64(b->U64, b->B[b->iX], b->B[b->iAX], L, K, nts, nullptr, st));
  LAPH_TRY(rayleigh_ritz(b));
  std::vector<double> lc(nts);
  int active = 0;
  LAPH_TRY(residuals_and_window(b, lc, &active));
  long long sv32 = 0, sv64 = (long long)nts * K, flaunch = 0;
  double fbytes = 0;
  float ms_filter = 0, ms_dense = 0;
  const int max_iter = b->o.max_iter > 0 ? b->o.max_iter : 60;
  int it = 0;
  for (it = 0; it < max_iter && active > 0; ++it) {
    LAPH_TRY(upload_coefficients(b, lc));
    for (int t = 0; t < nts; ++t) {
      if (b->nlock[t] >= b->nev) continue;
      const int cols = K - (b->nlock[t] / 64) * 64;
      sv32 += (long long)(b->deg - 1) * cols;
      // algorithmic bytes: step 1 reads r and writes f_2 (2 vector touches),
      // step 2 adds r (3), later steps add f_{k-1} (4); a vector touch is
      // 3 colours x 8 B per site; links 27 x 8 B per site per 64-vector block
      for (int s = 1; s < b->deg; ++s) {
        const int touches = s == 1 ? 2 : (s == 2 ? 3 : 4);
        fbytes += (double)b->V3 * ((double)cols * touches * 24.0 + (cols / 64) * 216.0);
      }
    }
    flaunch += b->deg - 1;
    float2* fm = nullptr;
    LAPH_HIP(hipEventRecord(b->ev[0], st));
    LAPH_TRY(run_filter(b, &fm));
    LAPH_HIP(hipEventRecord(b->ev[1], st));
    dim3 g((unsigned)((b->per_ts + 255) / 256), nts);
    k_combine_gamma<<<g, 256, 0, st>>>(b->B[b->iX], fm, b->gam, b->per_ts, K, b->first_blk);
    LAPH_HIP(hipGetLastError());
    LAPH_TRY(cholqr(b, b->B[b->iX]));
    LAPH_TRY(cholqr(b, b->B[b->iX]));
    LAPH_HIP(launch_lap_block64(b->U64, b->B[b->iX], b->B[b->iAX], L, K, nts, nullptr, st));
    sv64 += (long long)nts * K;
    LAPH_TRY(rayleigh_ritz(b));
    LAPH_HIP(hipEventRecord(b->ev[2], st));
    LAPH_TRY(residuals_and_window(b, lc, &active));
    float a = 0, c = 0;
    LAPH_HIP(hipEventElapsedTime(&a, b->ev[0], b->ev[1]));
    LAPH_HIP(hipEventElapsedTime(&c, b->ev[1], b->ev[2]));
    ms_filter += a;
    ms_dense += c;
  }
  b->solved = true;
  int conv = 0;
  double maxres = 0;
  for (int t = 0; t < nts; ++t) {
    if (b->nlock[t] >= b->nev) ++conv;
    for (int j = 0; j < b->nev; ++j) maxres = std::max(maxres, std::sqrt(b->h_res[(size_t)t * K + j]));
  }
  if (stats) {
    stats->iterations = it;
    stats->stencil_vecs = sv32;
    stats->stencil64_vecs = sv64;
    stats->converged = conv;
    stats->max_residual = maxres;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    stats->seconds_filter = ms_filter * 1e-3;
    stats->seconds_dense = ms_dense * 1e-3;
    stats->filter_launches = flaunch;
    stats->filter_bytes = fbytes;
  }
  if (conv < nts)
    return set_error(LAPH_ERR_NOCONV, "laph_batch_solve: %d of %d timeslices converged after %d cycles", conv, nts,
                     it);
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
  const int n = 3 * p->L * p->L * p->L;
  int K = o->nsub > 0 ? o->nsub : o->nev + std::max(16, o->nev / 4);
  K = (K + 63) / 64 * 64;
  LAPH_CHECK(K > o->nev && K <= n, LAPH_ERR_INVALID, "laph_batch_create: search space %d invalid for n=%d nev=%d", K,
             n, o->nev);
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
    b->nlock.assign(b->nts, 0);
    b->h_mu.assign((size_t)b->nts * K, 0.0);
    b->h_res.assign((size_t)b->nts * K, 0.0);
    auto fail = [&](int rc) {
      laph::batch_free(b);
      return rc;
    };
    if (hipSetDevice(p->device) != hipSuccess || hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_batch_create: stream creation failed"));
    if (rocblas_create_handle(&b->rb) != rocblas_status_success || rocblas_set_stream(b->rb, b->st) != rocblas_status_success)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_batch_create: rocBLAS handle"));
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

int laph_batch_get(laph_batch* b, int ts, double* evals, double* V_fix, double* phase, double* residuals) {
  LAPH_CHECK(b && ts >= 0 && ts < b->nts && evals, LAPH_ERR_INVALID, "laph_batch_get: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_batch_get: batch not solved");
  LAPH_HIP(hipSetDevice(b->p.device));
  const int K = b->K, nev = b->nev;
  for (int j = 0; j < nev; ++j) evals[j] = -b->h_mu[(size_t)ts * K + j];
  if (residuals)
    for (int j = 0; j < nev; ++j) residuals[j] = std::sqrt(b->h_res[(size_t)ts * K + j]);
  if (V_fix || phase) {
    double2* dV = nullptr;
    double* dph = nullptr;
    LAPH_HIP(hipMalloc(&dV, (size_t)b->n * nev * sizeof(double2)));
    hipError_t e = hipMalloc(&dph, nev * sizeof(double));
    if (e == hipSuccess)
      e = laph::launch_extract_phase(b->B[b->iX] + (size_t)ts * b->per_ts, dV, dph, b->n, K, nev, b->st);
    if (e == hipSuccess && V_fix)
      e = hipMemcpyAsync(V_fix, dV, (size_t)b->n * nev * sizeof(double2), hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess && phase) e = hipMemcpyAsync(phase, dph, nev * sizeof(double), hipMemcpyDeviceToHost, b->st);
    if (e == hipSuccess) e = hipStreamSynchronize(b->st);
    (void)hipFree(dV);
    if (dph) (void)hipFree(dph);
    LAPH_HIP(e);
  }
  return LAPH_OK;
}

}  // extern "C"
