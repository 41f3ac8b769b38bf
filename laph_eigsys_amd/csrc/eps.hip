// Plain FP64 drop-in of the EPS solve (ev_ts.cpp:231-303): Hermitian
// Krylov-Schur (thick-restart Lanczos; Stewart, SIMAX 23 (2001); Wu & Simon,
// SIMAX 22 (2000)) for the nev largest-magnitude eigenvalues of T_DEG(B),
// with every basis vector resident on the device.
//
// Per Lanczos step: the filtered matvec (DEG fused stencil steps, written
// straight into the next basis column), classical Gram-Schmidt with one
// conditional refinement pass (DGKS criterion |w_1| < |w_0|/sqrt(2), the
// "if needed" refinement SLEPc uses by default), and the normalisation.
// Per restart: the m x m real symmetric projected matrix is diagonalised on
// the device (rocSOLVER dsyevd), Ritz values are sorted by decreasing
// magnitude, the converged prefix is counted with the relative criterion
// |beta s_{m,i}| < tol |theta_i|, and the basis is compressed to the kept
// Ritz vectors plus the residual direction (a real GEMM on the interleaved
// complex basis).
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <vector>

#include "ctx.h"

namespace laph {

#define LAPH_RB(call)                                                                    \
  do {                                                                                   \
    rocblas_status s_ = (call);                                                          \
    if (s_ != rocblas_status_success)                                                    \
      return laph::set_error(LAPH_ERR_HIP, "%s failed: %s", #call, rocblas_status_to_string(s_)); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T* as() {
    return static_cast<T*>(p);
  }
};

struct RbHandle {
  rocblas_handle h = nullptr;
  ~RbHandle() {
    if (h) rocblas_destroy_handle(h);
  }
};

static uint64_t splitmix_next(uint64_t* s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double splitmix_u01(uint64_t* s) {
  return ((splitmix_next(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Seeded Gaussian start vector, normalised (one sequential splitmix64 stream).
void start_vector(uint64_t seed, size_t n, std::vector<double2>& v) {
  v.resize(n);
  uint64_t st = seed;
  double nrm = 0;
  for (size_t i = 0; i < n; ++i) {
    double r = std::sqrt(-2.0 * std::log(splitmix_u01(&st)));
    double ph = 2.0 * M_PI * splitmix_u01(&st);
    v[i] = make_double2(r * std::cos(ph), r * std::sin(ph));
    nrm += v[i].x * v[i].x + v[i].y * v[i].y;
  }
  nrm = std::sqrt(nrm);
  for (auto& e : v) e = make_double2(e.x / nrm, e.y / nrm);
}

static int eps_solve_impl(laph_ctx* ctx, const laph_eps_opts* o, double* theta_out, double* Vout,
                          laph_eps_stats* stats) {
  auto t0 = std::chrono::steady_clock::now();
  const size_t n = ctx->n;
  const int nev = o->nev;
  LAPH_CHECK(nev >= 1 && (size_t)nev < n, LAPH_ERR_INVALID, "laph_eps_solve: nev=%d out of range", nev);
  int m = o->ncv > 0 ? o->ncv : std::max(2 * nev, nev + 15);
  if ((size_t)m > n) m = (int)n;
  LAPH_CHECK(m > nev, LAPH_ERR_INVALID, "laph_eps_solve: ncv=%d must exceed nev=%d", m, nev);
  const double tol = o->tol > 0 ? o->tol : 1e-8;
  const int max_it = o->max_it > 0 ? o->max_it : std::max<int>(100, (int)(2 * n / m));
  hipStream_t st = ctx->stream;

  RbHandle rb;
  LAPH_RB(rocblas_create_handle(&rb.h));
  LAPH_RB(rocblas_set_stream(rb.h, st));
  LAPH_RB(rocblas_set_pointer_mode(rb.h, rocblas_pointer_mode_host));

  DevBuf dV, dT, dH, dh2, dS, dW, dE, dInfo, dSsel;
  LAPH_HIP(hipMalloc(&dV.p, n * (m + 1) * sizeof(double2)));
  LAPH_HIP(hipMalloc(&dT.p, n * (size_t)m * sizeof(double2)));
  LAPH_HIP(hipMalloc(&dH.p, (size_t)m * (m + 1) * sizeof(double2)));  // GS coefficients per column
  LAPH_HIP(hipMalloc(&dh2.p, (size_t)(m + 1) * sizeof(double2)));
  LAPH_HIP(hipMalloc(&dS.p, (size_t)m * m * sizeof(double)));
  LAPH_HIP(hipMalloc(&dW.p, (size_t)m * sizeof(double)));
  LAPH_HIP(hipMalloc(&dE.p, (size_t)m * sizeof(double)));
  LAPH_HIP(hipMalloc(&dInfo.p, sizeof(rocblas_int)));
  LAPH_HIP(hipMalloc(&dSsel.p, (size_t)m * m * sizeof(double)));
  double2* V = dV.as<double2>();
  auto col = [&](int j) { return V + (size_t)j * n; };
  auto rbz = [](double2* p) { return reinterpret_cast<rocblas_double_complex*>(p); };

  {
    std::vector<double2> v0;
    start_vector(o->seed, n, v0);
    LAPH_HIP(hipMemcpyAsync(col(0), v0.data(), n * sizeof(double2), hipMemcpyHostToDevice, st));
  }

  std::vector<double> H((size_t)m * m, 0.0), S((size_t)m * m), th(m), betas(m, 0.0);
  std::vector<double2> hcol((size_t)m * (m + 1));
  std::vector<int> ord(m);
  const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0), mone(-1.0, 0.0);
  const double eta = 1.0 / std::sqrt(2.0);
  int k = 0, nconv = 0, it = 0;
  long long nmat = 0;
  double beta = 0;
  for (it = 0; it < max_it; ++it) {
    for (int j = k; j < m; ++j) {
      double2* w = col(j + 1);
      LAPH_TRY(cheb_apply(ctx, col(j), w));
      ++nmat;
      double nrm0 = 0, nrm1 = 0;
      LAPH_RB(rocblas_dznrm2(rb.h, (rocblas_int)n, rbz(w), 1, &nrm0));
      double2* hj = dH.as<double2>() + (size_t)j * (m + 1);
      LAPH_RB(rocblas_zgemv(rb.h, rocblas_operation_conjugate_transpose, (rocblas_int)n, j + 1, &one,
                            rbz(V), (rocblas_int)n, rbz(w), 1, &zero, rbz(hj), 1));
      LAPH_RB(rocblas_zgemv(rb.h, rocblas_operation_none, (rocblas_int)n, j + 1, &mone, rbz(V),
                            (rocblas_int)n, rbz(hj), 1, &one, rbz(w), 1));
      LAPH_RB(rocblas_dznrm2(rb.h, (rocblas_int)n, rbz(w), 1, &nrm1));
      if (nrm1 < eta * nrm0) {
        double2* h2 = dh2.as<double2>();
        LAPH_RB(rocblas_zgemv(rb.h, rocblas_operation_conjugate_transpose, (rocblas_int)n, j + 1, &one,
                              rbz(V), (rocblas_int)n, rbz(w), 1, &zero, rbz(h2), 1));
        LAPH_RB(rocblas_zgemv(rb.h, rocblas_operation_none, (rocblas_int)n, j + 1, &mone, rbz(V),
                              (rocblas_int)n, rbz(h2), 1, &one, rbz(w), 1));
        LAPH_RB(rocblas_zaxpy(rb.h, j + 1, &one, rbz(h2), 1, rbz(hj), 1));
        LAPH_RB(rocblas_dznrm2(rb.h, (rocblas_int)n, rbz(w), 1, &nrm1));
      }
      LAPH_CHECK(nrm1 > 0.0 && std::isfinite(nrm1), LAPH_ERR_NOCONV,
                 "laph_eps_solve: Lanczos breakdown at step %d (|w| = %g)", j, nrm1);
      const double inv = 1.0 / nrm1;
      LAPH_RB(rocblas_zdscal(rb.h, (rocblas_int)n, &inv, rbz(w), 1));
      betas[j] = nrm1;
    }
    LAPH_HIP(hipMemcpyAsync(hcol.data(), dH.p, hcol.size() * sizeof(double2), hipMemcpyDeviceToHost, st));
    LAPH_HIP(hipStreamSynchronize(st));
    for (int j = k; j < m; ++j) {
      for (int i = 0; i <= j; ++i) {
        double v = hcol[(size_t)j * (m + 1) + i].x;
        H[(size_t)i * m + j] = v;
        H[(size_t)j * m + i] = v;
      }
      if (j + 1 < m) {
        H[(size_t)(j + 1) * m + j] = betas[j];
        H[(size_t)j * m + j + 1] = betas[j];
      }
    }
    beta = betas[m - 1];
    // projected eigenproblem on the device (row-major H is its own transpose)
    LAPH_HIP(hipMemcpyAsync(dS.p, H.data(), H.size() * sizeof(double), hipMemcpyHostToDevice, st));
    LAPH_RB(rocsolver_dsyevd(rb.h, rocblas_evect_original, rocblas_fill_upper, m, dS.as<double>(), m,
                             dW.as<double>(), dE.as<double>(), dInfo.as<rocblas_int>()));
    rocblas_int info = 0;
    LAPH_HIP(hipMemcpyAsync(S.data(), dS.p, S.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    LAPH_HIP(hipMemcpyAsync(th.data(), dW.p, th.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    LAPH_HIP(hipMemcpyAsync(&info, dInfo.p, sizeof info, hipMemcpyDeviceToHost, st));
    LAPH_HIP(hipStreamSynchronize(st));
    LAPH_CHECK(info == 0, LAPH_ERR_NOCONV, "laph_eps_solve: dsyevd info=%d", (int)info);
    // S is column-major: eigenvector c is S[c*m + r]
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return std::fabs(th[a]) > std::fabs(th[b]); });
    nconv = 0;
    for (int i = 0; i < m; ++i) {
      int c = ord[i];
      if (std::fabs(beta * S[(size_t)c * m + m - 1]) < tol * std::fabs(th[c])) ++nconv;
      else break;
    }
    const bool done = nconv >= nev || it == max_it - 1;
    int keep = done ? nev : nconv + (m - nconv) / 2;
    if (!done && keep < nev) keep = std::min(nev, m - 1);
    keep = std::min(keep, m - 1);
    // Ssel (m x keep, column-major) = S[:, ord[0..keep)]
    std::vector<double> ssel((size_t)m * keep);
    for (int i = 0; i < keep; ++i)
      std::copy(S.begin() + (size_t)ord[i] * m, S.begin() + (size_t)(ord[i] + 1) * m, ssel.begin() + (size_t)i * m);
    LAPH_HIP(hipMemcpyAsync(dSsel.p, ssel.data(), ssel.size() * sizeof(double), hipMemcpyHostToDevice, st));
    const double done1 = 1.0, dzero = 0.0;
    LAPH_RB(rocblas_dgemm(rb.h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)(2 * n), keep, m,
                          &done1, reinterpret_cast<const double*>(V), (rocblas_int)(2 * n),
                          dSsel.as<double>(), m, &dzero, dT.as<double>(), (rocblas_int)(2 * n)));
    LAPH_HIP(hipMemcpyAsync(V, dT.p, n * keep * sizeof(double2), hipMemcpyDeviceToDevice, st));
    if (done) {
      for (int i = 0; i < nev; ++i) theta_out[i] = th[ord[i]];
      break;
    }
    LAPH_HIP(hipMemcpyAsync(col(keep), col(m), n * sizeof(double2), hipMemcpyDeviceToDevice, st));
    std::fill(H.begin(), H.end(), 0.0);
    for (int i = 0; i < keep; ++i) H[(size_t)i * m + i] = th[ord[i]];
    k = keep;
  }
  if (Vout)
    LAPH_HIP(hipMemcpyAsync(Vout, V, n * nev * sizeof(double2), hipMemcpyDeviceToHost, st));
  LAPH_HIP(hipStreamSynchronize(st));
  if (stats) {
    stats->restarts = it;
    stats->nconv = nconv;
    stats->matvecs = nmat;
    stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (nconv < nev)
    return set_error(LAPH_ERR_NOCONV, "laph_eps_solve: %d of %d pairs converged after %d restarts", nconv,
                     nev, it);
  return LAPH_OK;
}

}  // namespace laph

extern "C" int laph_eps_solve(laph_ctx* ctx, const laph_eps_opts* opts, double* theta, double* V,
                              laph_eps_stats* stats) {
  LAPH_CHECK(ctx && opts && theta, LAPH_ERR_INVALID, "laph_eps_solve: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  return laph::guarded([&] { return laph::eps_solve_impl(ctx, opts, theta, V, stats); });
}
