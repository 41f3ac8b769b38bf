// Times every Tensile solution rocBLAS lists for the batched solver's two
// large FP64 complex GEMM shapes at 32^3 x 64 timeslices with K = 256
// columns: the Gram C = Y Y^H (m = n = K, k = 98304 rows) and the rotation
// T = Z^H Y (m = K, n = 98304, k = K), strided-batched over the 64
// timeslices; solution 0 is rocBLAS's own choice.  Prints the default and the
// five fastest, with TFLOP/s (8 m n k flop per timeslice).  Needs about 52 GB
// of device memory.
#define ROCBLAS_BETA_FEATURES_API
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <cstdio>
#include <vector>
#include <algorithm>
#define CK(x)                                                    \
  do {                                                           \
    auto e_ = (x);                                               \
    if (e_ != 0) {                                               \
      std::printf("%s failed: %d\n", #x, (int)e_);               \
      return 1;                                                  \
    }                                                            \
  } while (0)
int main() {
  const int K = 256, n = 98304, nts = 64;
  const size_t per_ts = (size_t)n * K, kk = (size_t)K * K;
  rocblas_double_complex *Y, *T, *G;
  CK(hipMalloc(&Y, per_ts * nts * sizeof(rocblas_double_complex)));
  CK(hipMalloc(&T, per_ts * nts * sizeof(rocblas_double_complex)));
  CK(hipMalloc(&G, kk * nts * sizeof(rocblas_double_complex)));
  CK(hipMemset(Y, 0, per_ts * nts * sizeof(rocblas_double_complex)));
  CK(hipMemset(G, 0, kk * nts * sizeof(rocblas_double_complex)));
  rocblas_handle h;
  CK(rocblas_create_handle(&h));
  rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0)); CK(hipEventCreate(&e1));
  for (int shape = 0; shape < 2; ++shape) {
    // 0: Gram C = Y Y^H (m = n = K, k = sites); 1: rotation T = Z^H Y (m = K, n = sites, k = K)
    const rocblas_operation ta = shape == 0 ? rocblas_operation_none : rocblas_operation_conjugate_transpose;
    const rocblas_operation tb = shape == 0 ? rocblas_operation_conjugate_transpose : rocblas_operation_none;
    const int m = K, nn = shape == 0 ? K : n, k = shape == 0 ? n : K;
    const void* A = shape == 0 ? (void*)Y : (void*)G;
    const rocblas_stride sa = shape == 0 ? per_ts : kk;
    const void* B = Y;
    void* C = shape == 0 ? (void*)G : (void*)T;
    const rocblas_stride sc = shape == 0 ? kk : per_ts;
    auto run = [&](int idx) {
      return rocblas_gemm_strided_batched_ex(h, ta, tb, m, nn, k, &one, A, rocblas_datatype_f64_c, K, sa, B,
                                             rocblas_datatype_f64_c, K, per_ts, &zero, C, rocblas_datatype_f64_c, K,
                                             sc, C, rocblas_datatype_f64_c, K, sc, nts, rocblas_datatype_f64_c,
                                             rocblas_gemm_algo_solution_index, idx, rocblas_gemm_flags_none);
    };
    rocblas_int sz = 0;
    CK(rocblas_gemm_strided_batched_ex_get_solutions(h, ta, tb, m, nn, k, &one, A, rocblas_datatype_f64_c, K, sa, B,
                                                     rocblas_datatype_f64_c, K, per_ts, &zero, C,
                                                     rocblas_datatype_f64_c, K, sc, C, rocblas_datatype_f64_c, K, sc,
                                                     nts, rocblas_datatype_f64_c, rocblas_gemm_algo_solution_index,
                                                     rocblas_gemm_flags_none, nullptr, &sz));
    std::vector<rocblas_int> sol(sz);
    CK(rocblas_gemm_strided_batched_ex_get_solutions(h, ta, tb, m, nn, k, &one, A, rocblas_datatype_f64_c, K, sa, B,
                                                     rocblas_datatype_f64_c, K, per_ts, &zero, C,
                                                     rocblas_datatype_f64_c, K, sc, C, rocblas_datatype_f64_c, K, sc,
                                                     nts, rocblas_datatype_f64_c, rocblas_gemm_algo_solution_index,
                                                     rocblas_gemm_flags_none, sol.data(), &sz));
    sol.insert(sol.begin(), 0);  // 0 = the default choice
    std::vector<std::pair<float, int>> res;
    for (int idx : sol) {
      if (run(idx) != rocblas_status_success) continue;  // warm-up
      CK(hipEventRecord(e0));
      for (int r = 0; r < 3; ++r) run(idx);
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      float ms = 0;
      CK(hipEventElapsedTime(&ms, e0, e1));
      res.push_back({ms / 3, idx});
    }
    std::sort(res.begin(), res.end());
    const double flop = 8.0 * m * nn * (double)k * nts;
    float def = 0;
    for (auto& r : res) if (r.second == 0) def = r.first;
    std::printf("shape %s: %d solutions, default %.2f ms (%.1f TFLOP/s)\n", shape == 0 ? "gram" : "rotation", sz, def,
                flop / (def * 1e-3) / 1e12);
    for (size_t i = 0; i < std::min<size_t>(5, res.size()); ++i)
      std::printf("  solution %d: %.2f ms (%.1f TFLOP/s)\n", res[i].second, res[i].first,
                  flop / (res[i].first * 1e-3) / 1e12);
  }
  return 0;
}
