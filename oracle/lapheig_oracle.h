/*
 * lapheig_oracle.h -- CPU FP64 restatement of the LapH_EigSys per-timeslice
 * eigensystem path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (laph_eigsys_amd, include/laph.h) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the LapH_EigSys source tree).  Layout conventions, shared with the product:
 *   site  s = x*L*L + y*L + z          (navigation.cpp:158-192)
 *   dir   0 = x (stride L*L), 1 = y (stride L), 2 = z (stride 1)
 *   links U[(s*3 + dir)*9 + 3*a + b]   complex, row a / column b
 *                                       (timeslice.cpp:46-73)
 *   vector v[3*s + colour]              (shell_matop.cpp:33-37, 90-93)
 * Complex numbers are interleaved (re, im) doubles.
 */
#ifndef LAPHEIG_ORACLE_H
#define LAPHEIG_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* navigation.cpp:27-88 -- periodic neighbour of site s in direction dir. */
int orc_up(int L, int s, int dir);
int orc_dn(int L, int s, int dir);

/* io.cpp:22-146: read the ildg-binary-data record of a LIME file, keep
 * timeslices [ti, tf] in the io.cpp config layout (72 doubles per site,
 * slot mu'*18 + 6i + 2j + {re,im}, mu'=0 is U_t).  Returns 0 on success,
 * -1 open failure, -2 no binary record, -3 size mismatch, -4 short read. */
int orc_read_lime(const char* path, int T, int L, int ti, int tf, double* config);

/* timeslice.cpp:34-79: config slice (72*V3 doubles) -> U (V3*3*9 complex). */
void orc_map_timeslice(const double* ts_config, int L, double* U);

/* Hermitian 3x3 eigen-decomposition by cyclic complex Jacobi:
 * A = Q diag(w) Q^H.  Used for the matrix square root in the projection. */
void orc_herm3_eig(const double* A, double* w, double* Q);

/* timeslice.cpp:90-104: in -> in * ((in^H in)^{1/2})^{-1} * det(.)^{-1/3}. */
void orc_proj_su3(double* M);

/* timeslice.cpp:243-369: iter Jacobi-style HYP iterations in place. */
void orc_hyp(double* U, int L, double alpha1, double alpha2, int iter);

/* shell_matop.cpp:16-103: y = c*(sum_mu[U x(x+mu) + U^H(x-mu) x(x-mu)] - 6x) + a*x,
 * c = 2/(lamL-lamC), a = 1 + c*lamC.  Passing c=1, a=0 applies D itself. */
void orc_tv2_ca(const double* U, int L, double c, double a, const double* x, double* y);
void orc_tv2(const double* U, int L, double lamL, double lamC, const double* x, double* y);

/* shell_matop.cpp:136-168: y = T_deg(B) x.  deg >= 1 (T_1(B) = B); returns
 * -1 for deg < 1 (the reference leaves y unwritten there). */
int orc_tv_iter(const double* U, int L, double lamL, double lamC, int deg,
                const double* x, double* y);

/* Krylov-Schur (thick-restart Lanczos) for the nev largest-magnitude
 * eigenpairs of the Hermitian operator T_deg(B); the EPS configuration of
 * ev_ts.cpp:231-263 (EPS_HEP, EPS_LARGEST_MAGNITUDE, nev, ncv default
 * max(2 nev, nev+15), relative convergence |r|/|theta| < tol).
 * theta: nev values in order of decreasing |theta|; V: n x nev column-major
 * complex (n = 3*V3).  Returns the number of converged pairs (>= nev on
 * success), stats[0] = restarts, stats[1] = operator applications. */
int orc_eps_solve(const double* U, int L, double lamL, double lamC, int deg,
                  int nev, int ncv, double tol, int max_it, unsigned long long seed,
                  double* theta, double* V, long long* stats);

/* recover_spec.cpp:13-54: theta -> eigenvalue of D. */
double orc_invert_B(double b, double lamL, double lamC);
double orc_invert_Tn(double t, int deg);
void orc_recover_spectrum(int nev, const double* theta, int deg, double lamL,
                          double lamC, double* evals);

/* eigensystem.cpp:4-30 (fix_phase) and 32-57 (reset_phase). */
void orc_fix_phase(int n, int nev, const double* V, double* Vfix, double* phase);
void orc_reset_phase(int n, int nev, const double* V, const double* phase_want,
                     double* Vreset);

/* read_write.cpp:20-35: 1 when tr(V^H V) and sum(V^H V) pass, else 0. */
int orc_check_trace(int n, int nev, const double* V, double* trace2, double* sum2);

/* Real symmetric eigen-decomposition (Householder tridiagonalisation plus
 * implicit QL), ascending eigenvalues; A is m x m row-major and is
 * overwritten by the eigenvectors (column j = eigenvector j). */
int orc_dsyev(int m, double* A, double* w);

#ifdef __cplusplus
}
#endif
#endif
