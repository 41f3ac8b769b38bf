// Launchers of the blocked multi-vector kernels (layout X[ts][(3s+c)*K + col],
// links AoS [ts][s][dir][9]); all asynchronous on `st`.
#pragma once

#include <hip/hip_runtime.h>

namespace laph {

// out = a1 (B in) + a3 in + a2 r + b old per column, B = c_t (H - 6) + a_t;
// coef[ts*K + col] = (a1, a2, b, w) with w = a1 (a_t - 6 c_t) + a3, U32 the
// links prescaled by c_t (launch_scale_links32), so out = a1 (c_t H in) +
// w in + a2 r + b old; column blocks below first_blk[ts] are skipped.  old
// and r may be null (plain recurrence without the source term).  A
// product-form step, a factor a1 (B - z) with a complex shift z, has no old
// vector: its shift term is (a3 + i b) in with a3 + i b = -a1 z.
// Filter vectors stored as FP32 complex (bf16 = 0; bw = 32: two sites per
// wavefront, 64: one vector per lane, 128: two per lane) or as BF16 complex
// (bf16 = 1; bw = 128 or 256: two or four per lane); buffers block-major with
// bw-vector blocks.
// zrun = 1, 2, 4 or 8 (FP32, bw = 128, zrun dividing L, W a multiple of 4):
// the 4 wavefronts of a work-group take 4 neighbouring y-rows and walk zrun
// z-sites each, sharing rows through LDS and registers; 0 = one site per
// wavefront.
hipError_t launch_filter_step(const float2* U32, const void* in, const void* r, const void* old, void* out,
                              const float4* coef, const float2* cab, const int* first_blk, int L, int K, int nts,
                              int bw, int bf16, int W, int zrun, hipStream_t st);
// out = D v on every (or, with first_blk, every active) column.
// W: sweep strip width (divides L; L = plain x-major order)
// zrun > 0 (dividing L, W a multiple of 4): work-groups of 4 y-rows, zrun
// z-sites per wavefront, as the filter step; 0: one site per wavefront.
hipError_t launch_lap_block64(const double2* U64, const double2* v, double2* out, int L, int K, int nts,
                              const int* first_blk, int W, int zrun, hipStream_t st);
// Residuals of the Ritz pairs from the FP64 stencil: r = c_t (D X + X diag(mu))
// in the filter storage (block-major, bw-vector blocks, FP32 or BF16), mu the
// Ritz values of -D ([nts][K]), c_t = 2 / (lamL - min(mu[t][K-1], wcap)) (wcap
// the cap of the window's lower end, 0.75 lamL outside a warm-up cycle that
// starts from a Rayleigh-Ritz step), for the 64-column blocks >= first_blk[ts].
hipError_t launch_lap_residual64(const double2* U64, const double2* X, const double* mu, double lamL, double wcap,
                                 void* r, int L, int K, int nts, const int* first_blk, int W, int bw, int bf16,
                                 int zrun, hipStream_t st);
// Column sums over a block's rows go through per-chunk partials of kRedRows
// rows, added in chunk order (deterministic: no atomics); `red` holds
// nts * red_chunks(n) * K doubles.
constexpr int kRedRows = 512;
inline int red_chunks(int n) { return (n + kRedRows - 1) / kRedRows; }
// nrm2 = column sums of |D X + X diag(mu)|^2 from that r, for the columns
// >= first_col (a multiple of 64).
hipError_t launch_residual_norms(const void* r, const double* mu, double lamL, double wcap, double* nrm2,
                                 double* red, int n, int K, int nts, int first_col, int bw, int bf16,
                                 hipStream_t st);
// fix_phase (eigensystem.cpp:4-30) in place on the first nev columns of every
// timeslice of a [ts][n][K] block: phase[ts][col] = arg X[ts][0][col], then
// each column times e^{-i phase}.
hipError_t launch_fix_phases(double2* X, double* phase, int n, int K, int nev, int nts, hipStream_t st);
// V (column-major n x nev, the eigenvector-file layout) = the first nev
// columns of one timeslice's [n][K] block.
hipError_t launch_extract_cols(const double2* X, double2* V, int n, int K, int nev, hipStream_t st);
// psi(r) = sum_x |V_{x+r} V_x^H|_F for r in {0, 2, ..}^3 below L, V in the
// eigenvector-file layout (n x nev column-major); psi has ((L+1)/2)^3 entries
// in the order r1, r2, r3 (r3 fastest).
hipError_t launch_source_shape(const double2* V, int L, int nev, double* psi, hipStream_t st);
// Scales every column >= first_col (a multiple of 64) of a [ts][n][K] block to
// unit norm; tmp holds nts*K doubles (the squared norms), red the partials.
hipError_t launch_normalize_columns(double2* Y, double* tmp, double* red, int n, int K, int nts, int first_col,
                                    hipStream_t st);
hipError_t launch_convert_links32(const double2* U, float2* U32, size_t n, hipStream_t st);
// C[ts*S + c] = alpha sum over site chunk c of A[s][l+i] conj(B[s][l+j]) on
// the 64x64 tiles ti <= tj of the leading a x a block (column-major, lda K),
// FP64 matrix cores (gram_mfma.hip).  Needs a % 64 == 0 and (n / S) % 16 == 0
// (gram_mfma_fits).
bool gram_mfma_fits(int n, int a, int S);
hipError_t launch_gram_mfma(const double2* A, const double2* B, double2* C, int n, int K, int l, int a, int nts,
                            int S, size_t per_ts, double alpha, hipStream_t st);
// U32c = c_t U32 per timeslice (c_t = cab[ts].x): the filter steps' links.
hipError_t launch_scale_links32(const float2* U32, const float2* cab, float2* U32c, size_t per_ts, int nts,
                                hipStream_t st);
hipError_t launch_random_block(double2* X, size_t total, unsigned long long seed, hipStream_t st);
// Gaussian columns [first_col, K) of the timeslices with flag[ts] != 0.
hipError_t launch_reseed_cols(double2* X, int n, int K, int nts, int first_col, const int* flag,
                              unsigned long long seed, hipStream_t st);

}  // namespace laph
