// Launchers of the blocked multi-vector kernels (layout X[ts][(3s+c)*K + col],
// links AoS [ts][s][dir][9]); all asynchronous on `st`.
#pragma once

#include <hip/hip_runtime.h>

namespace laph {

// out = a1 (B in) + a2 r + b old per column, B = c_t (H - 6) + a_t;
// coef[ts*K + col] = (a1, a2, b, -), cab[ts] = (c_t, a_t - 6 c_t); column
// blocks below first_blk[ts] are skipped.  old and r may be null (plain
// recurrence without the source term).
// bw = 64: one vector per lane; bw = 128: two vectors per lane (float4),
// buffers block-major with bw-vector blocks.
hipError_t launch_filter_step32(const float2* U32, const float2* in, const float2* r, const float2* old,
                                float2* out, const float4* coef, const float2* cab, const int* first_blk,
                                int L, int K, int nts, int bw, hipStream_t st);
// out = D v on every (or, with first_blk, every active) column.
hipError_t launch_lap_block64(const double2* U64, const double2* v, double2* out, int L, int K, int nts,
                              const int* first_blk, hipStream_t st);
// R = AX - X diag(lam), r32 = c_t R, nrm2 = column sums of |R|^2, for the
// columns >= first_col (a multiple of 64).
hipError_t launch_residual(const double2* X, const double2* AX, const double* lam, const double* cts,
                           float2* r32, double* nrm2, int n, int K, int nts, int first_col, int bw, hipStream_t st);
// Column-major, phase-fixed copy of the first nev columns of one timeslice.
hipError_t launch_extract_phase(const double2* X, double2* V, double* phase, int n, int K, int nev,
                                hipStream_t st);
// Scales every column >= first_col (a multiple of 64) of a [ts][n][K] block to
// unit norm; tmp holds nts*K doubles.
hipError_t launch_normalize_columns(double2* Y, double* tmp, int n, int K, int nts, int first_col, hipStream_t st);
hipError_t launch_convert_links32(const double2* U, float2* U32, size_t n, hipStream_t st);
hipError_t launch_random_block(double2* X, size_t total, unsigned long long seed, hipStream_t st);

}  // namespace laph
