/*
 * laph.h -- C-ABI of the MI355X (gfx950) LapH per-timeslice eigensystem
 * library (liblapheig.so).
 *
 * Conventions
 *   - extern "C", plain pointers and sizes, int return: 0 = LAPH_OK,
 *     negative = error (laph_error_string / laph_last_error explain it).
 *     No exception crosses this boundary.
 *   - Complex numbers are interleaved (re, im) doubles (layout-compatible
 *     with PetscScalar = double _Complex and std::complex<double>).
 *   - Geometry: isotropic L^3 timeslice, V3 = L^3, site s = x*L*L + y*L + z,
 *     dir 0/1/2 = x/y/z, vector index 3*s + colour (n = 3*V3 entries).
 *   - Links (host side): U[((s*3 + dir)*9 + 3*a + b)] complex, row a, col b.
 *   - Caller owns host buffers; the library owns every device buffer it
 *     allocates.  One context per host thread; a context is not re-entrant.
 *
 * Each entry point names the reference interface it replaces
 * (file:line in the LapH_EigSys sources).
 */
#ifndef LAPH_H
#define LAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LAPH_OK 0
#define LAPH_ERR_INVALID (-1)  /* bad argument or size */
#define LAPH_ERR_HIP (-2)      /* HIP runtime / library failure */
#define LAPH_ERR_NOMEM (-3)    /* device or host allocation failed */
#define LAPH_ERR_IO (-4)       /* file open / read / write failure */
#define LAPH_ERR_FORMAT (-5)   /* LIME record missing or wrong size */
#define LAPH_ERR_NOCONV (-6)   /* eigensolver stopped before nev pairs converged */
#define LAPH_ERR_TRACE (-7)    /* tr(V^H V) guard failed */

const char* laph_error_string(int code);
/* Message of the last error raised on the calling thread ("" if none). */
const char* laph_last_error(void);
/* Library version string. */
const char* laph_version(void);

/* Parameters that ev_ts reads from parameters.txt (par_io.cpp:99-135):
 * LX (= LY = LZ), DEG, LAM_L, LAM_C; plus the HIP device ordinal. */
typedef struct laph_params {
  int L;
  int deg;
  double lambda_l;
  double lambda_c;
  int device;
} laph_params;

typedef struct laph_ctx laph_ctx;

/* Creates a context on params->device, uploads the spatial links of one
 * timeslice given in the io.cpp config layout (72*V3 doubles, as produced by
 * laph_read_lime), or unit links when ts_config is NULL.
 * Replaces Tslice::init + Tslice::map_timeslice_to_eigen
 * (timeslice.cpp:19-31, 34-79) and the MatCreateShell context
 * (ev_ts.cpp:197). */
int laph_ctx_create(laph_ctx** out, const laph_params* params, const double* ts_config);
int laph_ctx_destroy(laph_ctx* ctx);

/* Tslice::map_timeslice_to_eigen (timeslice.cpp:34-79). */
int laph_map_timeslice(laph_ctx* ctx, const double* ts_config);
/* Tslice::set_gauge / get_gauge for the whole timeslice (timeslice.cpp:81-88):
 * V3*3*9 complex in the host link layout above. */
int laph_set_links(laph_ctx* ctx, const double* links);
int laph_get_links(laph_ctx* ctx, double* links);

/* Tslice::smearing_hyp(alpha_1, alpha_2, iter) (timeslice.cpp:243-369),
 * on the device links of ctx. */
int laph_hyp_smear(laph_ctx* ctx, double alpha_1, double alpha_2, int iter);

/* MatMult_Laplacian2D (shell_matop.cpp:226-244): y = T_DEG(B) x with
 * B = c*D + a, c = 2/(LAM_L-LAM_C), a = 1 + c*LAM_C.  x, y: host arrays of
 * n = 3*V3 complex.  ctx is the laph_ctx* stored as the MatShell context.
 * DEG >= 1 (T_1(B) = B); DEG < 1 is rejected at context creation. */
int laph_shell_mult(void* ctx, const double* x, double* y);
/* MatGetDiagonal_Laplacian2D (shell_matop.cpp:248-255): diag[i] = -6 + 0i,
 * the diagonal of D (the value the reference registers), n complex. */
int laph_shell_get_diagonal(void* ctx, double* diag);
/* Device-pointer variant of laph_shell_mult (n complex on the context's
 * device; completes before returning). */
int laph_shell_mult_device(laph_ctx* ctx, const double* dx, double* dy);
/* Variant of laph_shell_mult_device that allocates nothing and needs no
 * context.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  y = T_DEG(B) x with the parameters of
 * *params (L, DEG, LAM_L, LAM_C, device); d_links: the smeared links on the
 * device in the host link layout above (V3*3*9 complex); dx, dy: device
 * vectors of n = 3*V3 complex, distinct.  The kernels are enqueued on
 * `stream` (a hipStream_t, NULL = the null stream) and the call returns
 * without waiting; dy is also used as one of the three recurrence buffers.
 * The other two live in a fixed device array of the library (2 * 3 * 48^3
 * complex doubles per device, 10.6 MB, held until the process ends): L > 48
 * is rejected with LAPH_ERR_INVALID.  Every call on a device shares that
 * array, so calls that overlap on one device (other streams, other host
 * threads) race; calls in one stream are ordered and safe.  Results are
 * bitwise equal to laph_shell_mult_device's. */
int laph_shell_mult_noalloc(const laph_params* params, const double* d_links, const double* dx, double* dy,
                            void* stream);
/* y = D x (tv2 with c = 1, a = 0; shell_matop.cpp:68-77), host arrays. */
int laph_laplacian(laph_ctx* ctx, const double* x, double* y);

/* ------------------------------------------------------------------------
 * EPS-shaped solve of the plain FP64 drop-in (ev_ts.cpp:231-303): Hermitian
 * Krylov-Schur for the nev eigenvalues of T_DEG(B) of largest magnitude,
 * with the search space, every Lanczos vector and the filtered matvec
 * resident on the device.  Outputs: theta (nev doubles, |theta| descending,
 * as EPSGetEigenpair returns them), V (host n x nev complex column-major,
 * may be NULL). */
typedef struct laph_eps_opts {
  int nev;                 /* EPSSetDimensions nev */
  int ncv;                 /* 0 = max(2*nev, nev+15) */
  double tol;              /* -eps_tol: |r|/|theta| < tol */
  int max_it;              /* 0 = max(100, 2n/ncv) */
  unsigned long long seed; /* start vector */
} laph_eps_opts;

typedef struct laph_eps_stats {
  int restarts;
  int nconv;
  long long matvecs; /* applications of T_DEG(B) */
  double seconds;
} laph_eps_stats;

int laph_eps_solve(laph_ctx* ctx, const laph_eps_opts* opts, double* theta, double* V,
                   laph_eps_stats* stats);

/* ------------------------------------------------------------------------
 * Post-processing, host arrays (recover_spec.cpp, eigensystem.cpp,
 * read_write.cpp). */
/* recover_spectrum (recover_spec.cpp:45-54). */
int laph_recover_spectrum(int nev, const double* theta, int deg, double lambda_l,
                          double lambda_c, double* evals);
/* fix_phase (eigensystem.cpp:4-30): V n x nev complex column-major. */
int laph_fix_phase(int n, int nev, const double* V, double* V_fix, double* phase);
/* reset_phase (eigensystem.cpp:32-57). */
int laph_reset_phase(int n, int nev, const double* V, const double* phase_want,
                     double* V_reset);
/* check_trace (read_write.cpp:20-35): returns LAPH_OK or LAPH_ERR_TRACE;
 * trace/sum (complex) may be NULL. */
int laph_check_trace(int n, int nev, const double* V, double* trace, double* sum);
/* write_eig_sys_bin (read_write.cpp:176-201): "<dir>/<prefix>.%04d.%03d",
 * guarded by check_trace. */
int laph_write_eig_sys_bin(const char* dir, const char* prefix, int config, int t, int n,
                           int nev, const double* V);
/* write_eigenvalues_bin (read_write.cpp:237-250). */
int laph_write_eigenvalues_bin(const char* dir, const char* prefix, int config, int t,
                               int nev, const double* values);
/* read_evectors_bin_ts (read_write.cpp:40-96): nev eigenvectors of n complex
 * doubles from "<dir>/<prefix>.%04d.%03d" into V (nev x n, row i =
 * eigenvector i, i.e. the column-major file order), then check_trace.
 * LAPH_ERR_IO: no such file; LAPH_ERR_FORMAT: fewer than nev vectors in it
 * (the reference exits in eof_check); LAPH_ERR_TRACE: the guard fails (the
 * reference exits with "Eigenvectors damaged"); V holds what was read. */
int laph_read_evectors_bin_ts(const char* dir, const char* prefix, int config, int t, int n, int nev,
                              double* V);
/* read_eigenvalues_bin (read_write.cpp:137-152): nev doubles.  A missing file
 * is LAPH_ERR_IO (the reference silently keeps its old vector), a short one
 * LAPH_ERR_FORMAT. */
int laph_read_eigenvalues_bin(const char* dir, const char* prefix, int config, int t, int nev, double* values);
/* read_eigenvalues_ascii (read_write.cpp:98-115): nev whitespace-separated
 * values from "<prefix>.%03d.%04d" (timeslice before config, as there). */
int laph_read_eigenvalues_ascii(const char* prefix, int config, int t, int nev, double* values);

/* ------------------------------------------------------------------------
 * Gauge-field ingest (io.cpp:22-146) and the seeded synthetic field. */
/* read_lime_gauge_field_doubleprec_timeslices: keeps t in [ti, tf],
 * output (tf-ti+1)*72*V3 doubles in the io.cpp config layout. */
int laph_read_lime(const char* path, int T, int L, int ti, int tf, double* config);
/* ILDG LIME writer for a config in the same layout (T*72*V3 doubles). */
int laph_write_lime(const char* path, int T, int L, const double* config);
/* U = exp(i eps H), H = sum_a xi_a lambda_a / 2 with xi_a ~ N(0,1) from a
 * counter-based hash of (seed, t, x, y, z, mu); T*72*V3 doubles. */
int laph_synth_config(unsigned long long seed, double eps, int L, int T, double* config);

/* ------------------------------------------------------------------------
 * Mixed-precision batched solver: all timeslices of a configuration on one
 * device, FP32 Chebyshev filter (correction form), FP64 CholQR2,
 * Rayleigh-Ritz and residual locking.  Replaces the per-rank timeslice loop
 * of ev_ts.cpp:176-336 (which spreads timeslices over MPI ranks,
 * ev_ts.cpp:96-126) for the EPS solve + fix_phase + recover_spectrum part. */
typedef struct laph_batch laph_batch;

/* Environment controls of a batch, read by laph_batch_create (for tuning and
 * tests; the defaults are the measured best):
 *   LAPH_LANES=1..4         host lanes (threads, streams) the timeslices are split over (1)
 *   LAPH_BW=32|64|128|256   vectors per filter block (else timed on the first solve)
 *   LAPH_STRIP=w            sweep strip width of the stencils, w divides L (else timed)
 *   LAPH_ZRUN=0|1|2|4|8     FP32 64-/128-vector filter step in work-groups of 4 y-rows (L a multiple of 4):
 *                           z-sites per wavefront, dividing L (8, else 4); 0: one site per wavefront
 *   LAPH_FILTER_FORM=0|1|2  T_m(B) applied as a product over its Leja-ordered roots, the correction
 *                           pass as a product over the roots of (T_m(y) - T_m(theta)) / (y - theta)
 *                           applied to r (2); the same with the correction pass reading r in every
 *                           step (1); the three-term recurrence, which also reads the previous vector (0)
 *   LAPH_WARM=n             cycles of T_DEG(B) before the per-cycle degree (or T_m(B)^P) cycles (1)
 *   LAPH_WARM_DEG=d         degree of the warm-up cycles in the per-cycle degree mode (a multiple of 4
 *                           up to LAPH_FILTER_DEGREE; else DEG or a higher degree timed on the first
 *                           solve: see laph_batch_stats.warmup_degree)
 *   LAPH_WARMUP_WINDOW_CAP=0  the warm-up cycle after a Rayleigh-Ritz start (a network proposal,
 *                           LAPH_START_RR) takes its window from the start block's Ritz values (else
 *                           its lower end is at most the start window)
 *   LAPH_WINDOW0=x          window lower end of the first cycle from the start block (3 x the K-th
 *                           level of the free Laplacian, at most 0.75 lambda_l)
 *   LAPH_QR_CYCLES=n        cycles (and start block) orthonormalised by explicit CholQR2 (0)
 *   LAPH_START_RR=1         Rayleigh-Ritz on the Gaussian start block before the first cycle (off: the
 *                           first cycle filters the block as it is)
 *   LAPH_AMP_CAP=x          bound on the amplification T_d(theta_max) (or T_DEG(theta_max)^P) of a cycle (1e10)
 *   LAPH_WARM_AMP_CAP=x     bound on the amplification at the start window of the warm-up degrees the
 *                           first solve times (1e12; each candidate is a whole solve, judged by the
 *                           residual test, and one that fails is dropped)
 *   LAPH_FILTER_DEGREE=d    highest per-cycle filter degree (128); 0: applications of T_DEG(B) per cycle
 *                           (up to 5, or filter_passes)
 *   LAPH_DEGREE_SAFETY=S    a cycle's degree is lowered to the least d with T_d(theta_j) >= S r_j / tol
 *                           for every unconverged wanted pair (theta_j, r_j) of the lane (30); 0: always the
 *                           highest degree the amplification bound allows
 *   LAPH_BLOCK_DEGREE=0     one degree per cycle for every 64-column block (else each block is filtered
 *                           to the degree its own unconverged wanted pairs need by the same rule, at most
 *                           the cycle's degree)
 *   LAPH_SHIFTED_CHOLQR=1   every CholQR pass starts with a shifted one
 *   LAPH_TRSM=1             Rayleigh-Ritz through triangular solves with the Gram's Cholesky factor
 *                           (else its explicit inverse and GEMMs)
 *   LAPH_EIG=0|1|2          Rayleigh-Ritz eigensolver: zheevd, zheevdj, zheevj (0)
 *   LAPH_GRAM_SPLIT=s[,c]   row blocks s of the Hermitian Gram and Rayleigh-quotient products (upper
 *                           triangle only; 1: the full products) and site chunks c (1, 2, 4, 8) of their
 *                           sums (else both timed on the first solve, with the kernel)
 *   LAPH_GRAM_MFMA=0|1      Gram products on rocBLAS zgemm (0) or on the hand-written FP64 MFMA
 *                           kernel (1); else timed on the first solve
 *   LAPH_TUNE_FILE=path     layout cache: the filter and Gram layouts and the warm-up degree of a shape
 *                           (device, L, K, timeslices, DEG, nev, tol) are read from this text file
 *                           instead of timed on the first solve, and appended to it after timing (not
 *                           used while LAPH_BW, LAPH_STRIP, LAPH_GRAM_SPLIT, LAPH_GRAM_MFMA or
 *                           LAPH_WARM_DEG is set)
 *   LAPH_DENSE_SUBSET=0     dense steps always on every timeslice of a lane, converged ones included
 *   LAPH_TRACE              (set) per-cycle trace of the first timeslice on stderr; "time": only
 *                           the per-cycle filter, dense and wall times (no extra synchronisation)
 *   LAPH_NET_TARGET=1       laph_net_train_step fits the span of all K Ritz vectors of the last solve
 *                           (else its nev eigenvectors)
 *   LAPH_TEST_POISON=t      (tests) NaN written into one row of timeslice t's unlocked columns in
 *                           cycle 2, to exercise the recovery counted in laph_batch_stats
 *   LAPH_TEST_FAIL_WARM=d   (tests) the first solve's warm-up degree candidate d reports a failure
 *                           without solving, to exercise the dropping of a failed candidate */

typedef struct laph_batch_opts {
  int nts;           /* timeslices held by the batch */
  int nev;           /* eigenpairs per timeslice */
  int nsub;          /* search-space width (0 = nev + max(16, nev/4)), rounded up to a multiple of 64 */
  double tol;        /* lock when |D v - lambda v| <= tol (v unit norm) */
  int max_iter;      /* filter / Rayleigh-Ritz cycles (0 = 60) */
  unsigned long long seed;
  int warm_start;    /* 1: start from the block proposed by laph_net_propose, when one was (Rayleigh-Ritz
                        on it, then the usual warm-up cycle); 2: the same without the warm-up cycle, the
                        proposal's Ritz values setting the first window */
  int filter_passes; /* 0: after one warm-up cycle of T_DEG(B), one Chebyshev filter per cycle of
                        the highest degree (a multiple of 4, at most 128) whose amplification
                        T_d(theta_max) stays below 1e10; 1 .. 6: up to that many applications of
                        T_DEG(B) per cycle, fewer while T_DEG(theta_max)^P would exceed 1e10 */
  int filter_storage; /* filter vectors in HBM: 0 FP32 complex, 1 BF16 complex (half the
                         bytes; arithmetic stays FP32, the FP64 Rayleigh-Ritz and residual
                         test are unchanged, so only the cycle count can differ) */
} laph_batch_opts;

typedef struct laph_batch_stats {
  int iterations;          /* filter + RR cycles of the slowest timeslice */
  long long stencil_vecs;  /* FP32 stencil applications (vector count) */
  long long stencil64_vecs;/* FP64 stencil applications (vector count) */
  int converged;           /* timeslices with all nev pairs locked */
  double max_residual;     /* largest |D v - lambda v| over all locked pairs */
  double seconds;
  double seconds_filter;   /* device time in the FP32 filter kernels */
  double seconds_dense;    /* device time in Gram / CholQR / RR / rotations */
  long long filter_launches; /* FP32 filter-step kernel launches */
  double filter_bytes;     /* algorithmic bytes of those launches (vectors read/written
                              once, links once per vector block) */
  int filter_block;        /* vectors per block of the filter (FP32 storage: 32, 64 or 128,
                              BF16: 128 or 256), chosen by timing on the first solve
                              unless LAPH_BW fixes it */
  int filter_strip;        /* sweep strip width of the stencil kernels (L = x-major
                              planes, 4 or 8), chosen with the block width unless
                              LAPH_STRIP fixes it */
  int recoveries;          /* times the active columns of failing timeslices were re-seeded
                              after even shifted Cholesky QR failed (a non-finite block) */
  int filter_zrun;         /* FP32 64- or 128-vector filter step in work-groups of 4 y-rows:
                              z-sites per wavefront (8 or 4; LAPH_ZRUN fixes it); 0 = one site
                              per wavefront (other layouts, L not a multiple of 4, LAPH_ZRUN=0) */
  double start_window;     /* window lower end of the first cycle from the start block: 3 x the
                              K-th level of the free Laplacian, at most 0.75 lambda_l
                              (LAPH_WINDOW0 fixes it) */
  int rr_cholqr_fallbacks; /* Rayleigh-Ritz cycles whose Gram Cholesky failed, after which the
                              block was orthonormalised by explicit CholQR2 */
  int shifted_cholqr;      /* shifted Cholesky QR passes (a CholQR Gram that was not numerically
                              positive definite) */
  int gram_blocks;         /* row blocks of the Hermitian Gram products (1: full zgemm, 2: upper
                              triangle in two row blocks), timed on the first solve */
  int gram_chunks;         /* site chunks of the Gram products (1, 2, 4 or 8: partial products as
                              extra batch entries, added in order), timed with gram_blocks */
  int gram_mfma;           /* 1: the Gram products ran on the hand-written FP64 MFMA kernel (the
                              64 x 64 tiles of the upper triangle), 0: on rocBLAS zgemm */
  int layout_from_cache;   /* 1: the filter and Gram layouts were read from the LAPH_TUNE_FILE
                              cache instead of timed on the first solve */
  int warmup_degree;       /* Chebyshev degree of the warm-up cycle: on the first solve from the
                              start block DEG and up to three higher degrees (24, 32, 40, within
                              LAPH_WARM_AMP_CAP at the start window) are each solved and
                              timed, and the fastest is kept (LAPH_WARM_DEG fixes it) */
  int subset_cycles;       /* cycles whose dense steps (projection, Rayleigh-Ritz, residuals) ran on
                              the runs of unconverged timeslices only, the converged ones keeping
                              their pairs (LAPH_DENSE_SUBSET=0 turns this off) */
} laph_batch_stats;

int laph_batch_create(laph_batch** out, const laph_params* params, const laph_batch_opts* opts);
int laph_batch_destroy(laph_batch* b);
/* Upload timeslice ts (io.cpp layout, 72*V3 doubles) and HYP-smear it on
 * the device with (alpha_1, alpha_2, iter). */
int laph_batch_set_timeslice(laph_batch* b, int ts, const double* ts_config, double alpha_1,
                             double alpha_2, int iter);
/* Upload already-smeared links of timeslice ts (host link layout). */
int laph_batch_set_links(laph_batch* b, int ts, const double* links);
/* Solve every timeslice of the batch; the solve ends with the nev eigenvectors
 * of every timeslice phase-fixed (fix_phase, eigensystem.cpp:4-30) in device
 * memory, and the phases recorded. */
int laph_batch_solve(laph_batch* b, laph_batch_stats* stats);
/* The three files of one solved timeslice straight from device memory
 * (ev_ts.cpp:304-318 for that timeslice): "<dir>/eigenvectors.%04d.%03d"
 * (phase-fixed, column-major n x nev complex128, guarded by check_trace with
 * V^H V computed on the device), "eigenvalues" and "phases".  The
 * eigenvectors stream to the file through pinned chunks, copy of the next
 * overlapping the write of the current.  trace_sum (4 doubles, may be null)
 * receives tr(V^H V) and sum(V^H V) as (re, im, re, im). */
int laph_batch_write(laph_batch* b, int ts, const char* dir, int config, int t, double* trace_sum);
/* Source shape of a solved timeslice (sourceshape_funcs.cpp:31-81):
 * psi[i] = sum_x |V_{x+r} V_x^H|_F over the ((L+1)/2)^3 vectors r of the
 * grid {0, 2, 4, ..}^3 below L, i in the order of the reference's loops (r1,
 * r2, r3; r3 fastest). */
int laph_batch_source_shape(laph_batch* b, int ts, double* psi);
/* The same for eigenvectors in host memory (nev x n, row i = eigenvector i:
 * the eigenvectors-file layout) of an L^3 timeslice, on device `device`. */
int laph_source_shape(int L, int nev, const double* V, int device, double* psi);

/* Device time of the FP32 filter-step kernel, alone on the device: replays
 * the captured filter of the first half of the timeslices (every column
 * active) reps times, timed with HIP events on its stream.  Outputs the
 * average launch duration and the algorithmic bytes per launch (vectors read
 * or written once per step, links once per 64-vector block).  Requires a
 * solved batch (its coefficient tables).  No counterpart in the reference. */
int laph_batch_time_filter(laph_batch* b, int reps, double* ms_per_launch, double* bytes_per_launch);
/* Device-pointer variant of laph_batch_get for the eigenvectors: the nev
 * phase-fixed eigenvectors of timeslice ts in the eigenvector-file layout (n x
 * nev complex column-major) into d_V, device memory of n*nev complex doubles
 * on the batch's device (caller-owned), and the phases into the host array
 * phase (may be NULL).  For keeping the eigensystems of every timeslice of a
 * configuration resident while the solver's work buffers serve groups of
 * timeslices in turn.  Completes before returning. */
int laph_batch_get_device(laph_batch* b, int ts, double* d_V, double* phase);
/* Device time of the Gram product Y^H Y of the first lane's timeslices at
 * full width (the solve's FP64 matrix-core kernel, strided-batched zgemm of
 * the upper-triangle row blocks), alone on the device: reps replays timed
 * with HIP events on the lane's stream.  Outputs the average duration and
 * the floating-point operations it executes.  Requires a solved batch (the
 * block it multiplies).  No counterpart in the reference. */
int laph_batch_time_gram(laph_batch* b, int reps, double* ms_per_gram, double* flops_per_gram);
/* Results of timeslice ts: eigenvalues of D (descending, i.e. nearest 0
 * first, like recover_spectrum's output), phase-fixed eigenvectors (n x nev
 * complex column-major, may be NULL), phases (may be NULL), per-pair
 * residuals |D v - lambda v| (may be NULL).  The eigenvectors pass through
 * one n x nev device buffer that the batch allocates on first use and keeps
 * (shared with laph_batch_write); a page-locked V_fix (hipHostMalloc) takes
 * the direct copy (55 GB/s against about 16 GB/s into pageable memory). */
int laph_batch_get(laph_batch* b, int ts, double* evals, double* V_fix, double* phase,
                   double* residuals);

/* ------------------------------------------------------------------------
 * Learned warm start: a gauge-covariant parallel-transport network (16
 * complex channels per group of 16 output vectors, BF16 activations, FP32
 * arithmetic) that maps Gaussian colour vectors to the starting block of the
 * batched solver.  No counterpart in the reference (its EPS starts from
 * SLEPc's random vector); the random start stays the fallback: a batch
 * solved after laph_net_propose is checked by the same residual test. */
typedef struct laph_net laph_net;

int laph_net_create(laph_net** out, int layers, unsigned long long seed, int device);
int laph_net_destroy(laph_net* net);
/* Number of float parameters: layers*7*16*16 complex weights, then
 * layers*16 gate biases. */
int laph_net_num_params(laph_net* net);
int laph_net_get_params(laph_net* net, float* params);
int laph_net_set_params(laph_net* net, const float* params);
/* Writes the network's output for fresh Gaussian inputs into the start block
 * of every timeslice of b (K columns each); laph_batch_solve uses it when
 * the batch was created with warm_start = 1. */
int laph_net_propose(laph_net* net, laph_batch* b, unsigned long long seed);
/* One Adam step (learning rate lr) on the converged eigenvectors of the last
 * laph_batch_solve of b; loss = -(mean fraction of each output vector inside
 * the span of the nev eigenvectors). */
int laph_net_train_step(laph_net* net, laph_batch* b, unsigned long long seed, double lr, double* loss);
/* Loss, output and parameter gradient for one timeslice given on the host:
 * links (host link layout), input (n x K complex FP32, [3*s+colour][K]),
 * targets V (n x nev complex FP64, [3*s+colour][nev]); output (n x K complex
 * FP64, may be NULL), grad (num_params floats, real gradients of Re W, Im W
 * and b, may be NULL).  fp32_act = 1 stores activations in FP32. */
int laph_net_loss_grad(laph_net* net, int L, int K, int fp32_act, const double* links, const float* input,
                       const double* V, int nev, double* loss, double* output, float* grad);

#ifdef __cplusplus
}
#endif
#endif
