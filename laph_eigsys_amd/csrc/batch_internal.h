// State of a laph_batch (the mixed-precision batched solver), shared by
// batch.hip and the warm-start network (net.hip).
#pragma once

#include <rocblas/rocblas.h>

#include <map>
#include <string>
#include <vector>

#include "common.h"

#define LAPH_RB(call)                                                                               \
  do {                                                                                              \
    rocblas_status s_ = (call);                                                                     \
    if (s_ != rocblas_status_success)                                                               \
      return laph::set_error(LAPH_ERR_HIP, "%s failed: %s", #call, rocblas_status_to_string(s_));   \
  } while (0)

// One lane = a contiguous range of timeslices [t0, t0+nt) solved by its own
// host thread on its own HIP stream and rocBLAS handle.  Two lanes run side
// by side so the HBM-bound FP32 filter of one overlaps the MFMA-bound dense
// FP64 steps of the other.
// Applications of T_DEG(B) per Rayleigh-Ritz cycle at most.
#define LAPH_MAX_PASSES 6

struct laph_lane {
  int t0 = 0, nt = 0;
  hipStream_t st = nullptr;
  rocblas_handle rb = nullptr;
  int iX = 0, iAX = 1, iT = 2;  // roles of the three FP64 block buffers
  int lock = 0;                 // hard-locked columns, common to the lane
  hipGraphExec_t graph = nullptr;  // correction-form filter pass of the batch's degree
  std::map<int, hipGraphExec_t> gdeg;  // correction passes by degree (per-cycle degree mode)
  int m = 0;                    // filter degree of the current cycle (0: the batch's degree)
  bool blk_cycle = false;       // this cycle's correction pass uses per-block degrees
  std::vector<int> dblk;        // [lane timeslice][K/64] filter degree of each 64-column block this cycle
  std::vector<int> fbs;         // [correction step][lane timeslice] host copy of the step table: kept
                                // here so that it outlives the asynchronous uploads of the cycle
  double wcap = 0.0;            // cap of the window's lower end: 0.75 lambda_l, or the start window
                                // in a warm-up cycle that follows a Rayleigh-Ritz start
  bool keep_roles = false;      // dense steps on a subset of the lane's timeslices (t0, nt narrowed to
                                // a run): the rotation is copied back into X, the roles stay
  hipGraphExec_t gplain[LAPH_MAX_PASSES - 1] = {};  // plain passes 2 .. LAPH_MAX_PASSES
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // per-solve statistics
  int cycles = 0, rc = 0;
  std::string msg;  // error message raised on the lane's thread
  long long sv32 = 0, sv64 = 0, flaunch = 0;
  int recoveries = 0;  // timeslice blocks re-seeded after a failed orthonormalisation
  int rr_fallbacks = 0;  // Rayleigh-Ritz Gram Cholesky failures (block then orthonormalised by CholQR2)
  int shifted = 0;       // shifted Cholesky QR passes (plain CholQR Gram not numerically definite)
  int subset_cycles = 0;  // cycles whose dense steps ran on the unconverged timeslices only
  double min_pivot2 = 0;  // LAPH_TRACE: smallest squared Cholesky pivot of the unit-diagonal Grams
                          // factored in the current cycle (CholQR and Rayleigh-Ritz), over the lane
  int min_pivot_ts = -1;  // the timeslice it belongs to
  double fbytes = 0;
  float ms_filter = 0, ms_dense = 0;
  double s_solve = 0, s_capture = 0;  // wall time of the lane's solve, host time capturing filter graphs
};

struct laph_batch {
  laph_params p{};
  laph_batch_opts o{};
  int V3 = 0, n = 0, K = 0, nts = 0, nev = 0, deg = 0;
  size_t per_ts = 0;  // n*K
  hipStream_t st = nullptr;  // set-up / extraction stream
  // links
  double2* U64 = nullptr;  // [nts][V3*27]
  float2* U32 = nullptr;
  float2* U32c = nullptr;   // [nts][V3*27] the filter steps' links: c_t U32 for the current window
  double2* dec = nullptr;   // HYP scratch for one timeslice
  double2* Unew = nullptr;
  // blocks [nts][n][K]; after a solve the eigenvectors are in B[0]
  double2* B[3] = {nullptr, nullptr, nullptr};
  // filter buffers: esz bytes per complex entry (8: FP32, 4: BF16 storage)
  void* R32 = nullptr;
  void* P[3] = {nullptr, nullptr, nullptr};
  int esz = 8;
  // small per-timeslice data
  double2* G = nullptr;      // [nts][K*K] Rayleigh-Ritz / CholQR matrix
  double2* G2 = nullptr;     // [nts][K*K] Gram of the active block
  double2* P2 = nullptr;     // [nts][K*K] projections on the locked block; Rayleigh-Ritz scratch
  double2* G3 = nullptr;     // [nts][K*K] inverse of the Gram's Cholesky factor
  double* mu = nullptr;      // [nts][K]
  double* E = nullptr;       // [nts][K]
  rocblas_int* info = nullptr;
  double* nrm2 = nullptr;    // [nts][K]
  double* cn = nullptr;      // [nts][K] column norms for CholQR scaling
  double* red = nullptr;     // [nts][red_chunks(n)][K] partial column sums (deterministic reductions)
  double* phase = nullptr;   // [nts][nev] phases of the phase-fixed eigenvectors (fix_phase)
  double* gam = nullptr;     // [nts][K]
  float4* coef = nullptr;    // [deg-1][nts][K]
  float4* coef2 = nullptr;   // [deg - 1 + LAPH_MAX_PASSES - 1][nts][K] plain passes
  float2* cab = nullptr;     // [nts]
  int* first_blk = nullptr;  // [nts] first filtered 64-column block
  int* first_blk_steps = nullptr;  // [correction step][nts] first filtered 64-column block of each step of
                                   // the per-cycle degree mode's correction pass (block degrees)
  int* lock_blk = nullptr;   // [nts] first active 64-column block of the dense steps
  // host mirrors, written by the lanes on disjoint ranges
  std::vector<int> nlock;
  std::vector<double> h_mu, h_res, h_phase;
  std::vector<laph_lane> lanes;
  bool solved = false;
  int eig_mode = 0;          // Rayleigh-Ritz eigensolver: 0 zheevd, 1 zheevdj, 2 zheevj (LAPH_EIG)
  rocblas_int* jsweeps = nullptr;
  bool trace = false;       // LAPH_TRACE set: per-cycle trace of timeslice 0 on stderr
  bool trace_time = false;  // LAPH_TRACE=time: only the per-cycle times (no extra synchronisation)
  bool have_start = false;  // B[0] holds a proposed start block (laph_net_propose)
  // vectors per block of the filter buffers; 0 until the first solve times
  // the candidates (LAPH_BW fixes it)
  int bw = 0;
  int bf16 = 0;  // filter vectors stored as BF16 (laph_batch_opts.filter_storage)
  int warm_cycles = 1;  // cycles of the batch's degree while the window settles (LAPH_WARM)
  double window0 = 0;   // window lower end of the first cycle from the start block (LAPH_WINDOW0)
  bool warmup_cap = true;  // warm-up cycle after a Rayleigh-Ritz start: window at most window0
                          // (LAPH_WARMUP_WINDOW_CAP=0: from the start block's Ritz values)
  int warm_deg = 0;     // degree of the warm-up cycles in the per-cycle degree mode (0: the batch's degree)
  bool warm_deg_fixed = false;  // LAPH_WARM_DEG set: the warm-up degree is not timed
  bool warm_tuned = false;      // the warm-up degree was timed (or read from the layout cache)
  bool dense_subset = true;     // dense steps on the unconverged timeslices only when that is cheaper
                                // (LAPH_DENSE_SUBSET=0: always on the whole lane)
  bool block_degrees = true;    // per-block degrees in the per-cycle degree mode (LAPH_BLOCK_DEGREE=0: off)
  std::string tune_miss_key;    // layout cache key to append once the warm-up degree is chosen
  int strip = 0;  // sweep strip width of the stencil kernels (divides L; L = x-major planes)
  bool strip_fixed = false;  // LAPH_STRIP given
  int zrun = 0;  // z-sites per wavefront of the y-row work-group filter step (LAPH_ZRUN; 0: one site each)
  bool tuned = false;        // filter layout chosen (first solve)
  bool tuned_from_cache = false;  // ... read from the LAPH_TUNE_FILE layout cache instead of timed
  int poison_ts = -1;  // LAPH_TEST_POISON: NaN into one row of this timeslice's active block at cycle 2 (tests)
  int test_fail_warm = -1;  // LAPH_TEST_FAIL_WARM: this warm-up degree candidate reports a failure (tests)
  int qr_cycles = 0;  // cycles (and the start block) orthonormalised by explicit CholQR2 (LAPH_QR_CYCLES)
  // laph_batch_write: phase-fixed copy in the file layout, V^H V, rocBLAS handle on st,
  // two pinned staging chunks and their events (created on first use)
  double2* wV = nullptr;
  double2* wG = nullptr;
  rocblas_handle wrb = nullptr;
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t wev[2] = {nullptr, nullptr};
  bool prod = true;          // product-form filter over Leja-ordered roots of T_m; LAPH_FILTER_FORM=0:
                             // three-term recurrence
  std::vector<double> roots; // roots of T_deg in the order the product form applies them
  int fdeg_max = 128;        // per-cycle filter degree up to this, one application per cycle after the
                             // warm-up cycles (LAPH_FILTER_DEGREE; set by laph_batch_create); 0: the
                             // batch's degree and up to filter_passes applications per cycle (an
                             // explicit filter_passes selects it)
  // cos, sin of 2 pi k / m of the correction roots in application order, by degree m
  std::vector<std::vector<std::pair<double, double>>> qrot;
  bool qform = true;         // correction pass as a product over the roots of q_theta (LAPH_FILTER_FORM=2,
                             // the default; 1: product form reading r in every step)
  double amp_cap = 1e10;     // bound on T_m(theta_max)^P of a cycle (LAPH_AMP_CAP)
  double warm_amp_cap = 1e12;  // bound on the warm-up degrees the first solve times (LAPH_WARM_AMP_CAP)
  double deg_safety = 30.0;  // margin S of the residual-based cycle degree (LAPH_DEGREE_SAFETY; 0: off)
  bool use_trsm = false;     // triangular solves instead of the explicit inverse in Rayleigh-Ritz (LAPH_TRSM=1)
  bool start_rr = false;     // Rayleigh-Ritz on the Gaussian start block before the first cycle (LAPH_START_RR=1)
  bool force_shift = false;  // every CholQR pass starts with a shifted one (LAPH_SHIFTED_CHOLQR)
  bool net_target_all = false;  // laph_net_train_step targets all K converged Ritz vectors (LAPH_NET_TARGET=1)
  int gram_split = 2;        // row blocks of the Hermitian Gram / Rayleigh-quotient products, 1 (the full
                             // product) or 2, timed on the first solve (LAPH_GRAM_SPLIT fixes it)
  bool gram_split_fixed = false;
  int gram_ksplit = 1;       // site chunks of the Gram products (1, 2, 4 or 8), timed with gram_split
  int gram_mfma = 0;         // 1: the hand-written FP64 MFMA kernel (64 x 64 upper tiles) instead of rocBLAS
  int gram_ksplit_max = 8;
  double2* Gpart = nullptr;  // [nts][gram_ksplit_max][K*K] partial Gram products of the site chunks
};

namespace laph {
// K-th lowest eigenvalue of -D on the free field (all links 1).
double free_level(int L, int K);
// z-sites per wavefront of the y-row work-group stencils for the batch's sweep
// strip (they need a strip width that is a multiple of 4); 0: one site each.
inline int stencil_zrun(const laph_batch* b) { return b->strip % 4 == 0 ? b->zrun : 0; }
// Allocates laph_batch::wV on first use.
int output_buffers(laph_batch* b);
}  // namespace laph
