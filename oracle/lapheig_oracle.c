/*
 * lapheig_oracle.c -- CPU FP64 restatement of the LapH_EigSys per-timeslice
 * eigensystem path (HYP smearing, shell stencil + Chebyshev, Krylov-Schur,
 * spectrum recovery, phase fixing, trace guard, LIME ingest).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product and the CPU
 * baseline.  The product never links or calls this file.
 *
 * Parity status per function (how each is pinned; see tests/test_oracle_*.py):
 *   orc_read_lime       coordinate-encoded known-answer field written by an
 *                       independent numpy LIME writer
 *   orc_map_timeslice   known-answer field (position-encoded entries)
 *   orc_proj_su3        polar uniqueness: P[W]=W, P[sW]=W, P[W H]=W (H > 0),
 *                       det = 1, unitarity; numpy restatement (eigh-based)
 *   orc_hyp             fixed points (unit field, diagonal abelian field),
 *                       alpha = 0 identity, gauge covariance, numpy restatement
 *   orc_tv2 / tv_iter   free-field plane waves (closed form), Hermiticity,
 *                       assembled dense matrix, T_n(B) on eigenvectors
 *   orc_eps_solve       dense numpy eigh of the assembled Laplacian (4^3, 8^3)
 *   orc_recover_*       exact round trip lambda -> theta -> lambda
 *   orc_fix_phase       Im V(0,n) = 0, Re V(0,n) > 0, norms preserved
 *   orc_check_trace     orthonormal / damaged bases
 *   orc_dsyev           numpy eigh on random symmetric matrices
 */
#include "lapheig_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double _Complex cplx;

/* ------------------------------------------------------------------------ */
/* navigation.cpp:27-88: site = x*L^2 + y*L + z, periodic neighbours.        */
int orc_up(int L, int s, int dir) {
  int x = s / (L * L), y = (s / L) % L, z = s % L;
  if (dir == 0) x = (x + 1) % L;
  else if (dir == 1) y = (y + 1) % L;
  else z = (z + 1) % L;
  return (x * L + y) * L + z;
}
int orc_dn(int L, int s, int dir) {
  int x = s / (L * L), y = (s / L) % L, z = s % L;
  if (dir == 0) x = (x + L - 1) % L;
  else if (dir == 1) y = (y + L - 1) % L;
  else z = (z + L - 1) % L;
  return (x * L + y) * L + z;
}

/* ------------------------------------------------------------------------ */
/* LIME container (record header: magic 0x456789ab, version, MB/ME flags,
 * 64-bit big-endian length, 128-byte type; payload padded to 8 bytes) and
 * the ILDG read of io.cpp:22-146.                                           */
static uint64_t be64(const unsigned char* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; ++i) v = (v << 8) | p[i];
  return v;
}
static uint32_t be32(const unsigned char* p) {
  return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static double swap_double(const unsigned char* p) { /* io_utils.cpp:89-111 */
  uint64_t v = be64(p);
  double d;
  memcpy(&d, &v, 8);
  return d;
}

int orc_read_lime(const char* path, int T, int L, int ti, int tf, double* config) {
  FILE* f = fopen(path, "rb");
  if (!f) return -1;
  unsigned char hdr[144];
  uint64_t len = 0;
  int found = 0;
  while (fread(hdr, 1, 144, f) == 144) {
    if (be32(hdr) != 0x456789abu) break;
    len = be64(hdr + 8);
    char type[129];
    memcpy(type, hdr + 16, 128);
    type[128] = 0;
    if (strcmp(type, "ildg-binary-data") == 0) { found = 1; break; }
    uint64_t padded = (len + 7) & ~(uint64_t)7;
    if (fseeko(f, (off_t)padded, SEEK_CUR) != 0) break;
  }
  if (!found) { fclose(f); return -2; }
  /* io.cpp:63: double precision byte count */
  if (len != (uint64_t)L * L * L * T * 72 * 8) { fclose(f); return -3; }
  unsigned char raw[72 * 8];
  const int LX = L, LY = L, LZ = L;
  /* io.cpp:85-88: file order t, z, y, x (x fastest) */
  for (int t = 0; t < T; ++t)
    for (int z = 0; z < LZ; ++z)
      for (int y = 0; y < LY; ++y)
        for (int x = 0; x < LX; ++x) {
          if (fread(raw, 1, sizeof raw, f) != sizeof raw) { fclose(f); return -4; }
          if (t < ti || t > tf) continue; /* io.cpp:109-110 */
          /* io.cpp:114-131: ILDG mu x,y,z,t -> slots 1,2,3,0 */
          size_t p = ((size_t)(t - ti) * LX * LY * LZ + (size_t)x * LY * LZ + (size_t)y * LZ + z) * 72;
          int k = 0;
          for (int mu = 1; mu <= 4; ++mu) {
            size_t index = (mu != 4) ? p + (size_t)mu * 18 : p;
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j) {
                config[index + 6 * i + 2 * j] = swap_double(raw + 16 * k);
                config[index + 6 * i + 2 * j + 1] = swap_double(raw + 16 * k + 8);
                ++k;
              }
          }
        }
  fclose(f);
  return 0;
}

/* timeslice.cpp:34-79: U[s][dir](a,b) = config[s*72 + (dir+1)*18 + 6a + 2b] */
void orc_map_timeslice(const double* ts, int L, double* U) {
  const int V3 = L * L * L;
  for (int s = 0; s < V3; ++s)
    for (int dir = 0; dir < 3; ++dir)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
          size_t src = (size_t)s * 72 + (size_t)(dir + 1) * 18 + 6 * a + 2 * b;
          size_t dst = ((size_t)(s * 3 + dir) * 9 + 3 * a + b) * 2;
          U[dst] = ts[src];
          U[dst + 1] = ts[src + 1];
        }
}

/* ------------------------------------------------------------------------ */
/* 3x3 complex helpers (row-major, element (a,b) at [3a+b]).                  */
static void m3_mul(const cplx* A, const cplx* B, cplx* C) {
  cplx t[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      t[3 * a + b] = A[3 * a] * B[b] + A[3 * a + 1] * B[3 + b] + A[3 * a + 2] * B[6 + b];
  memcpy(C, t, sizeof t);
}
static void m3_dag(const cplx* A, cplx* C) {
  cplx t[9];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) t[3 * a + b] = conj(A[3 * b + a]);
  memcpy(C, t, sizeof t);
}
/* A * B^H and A^H * B without temporaries for the dagger */
static void m3_mul_dag(const cplx* A, const cplx* B, cplx* C) {
  cplx d[9];
  m3_dag(B, d);
  m3_mul(A, d, C);
}
static void m3_dag_mul(const cplx* A, const cplx* B, cplx* C) {
  cplx d[9];
  m3_dag(A, d);
  m3_mul(d, B, C);
}
static cplx m3_det(const cplx* A) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
         A[2] * (A[3] * A[7] - A[4] * A[6]);
}
static void m3_inv(const cplx* A, cplx* C) { /* cofactor inverse */
  cplx d = m3_det(A);
  cplx t[9];
  t[0] = (A[4] * A[8] - A[5] * A[7]) / d;
  t[1] = (A[2] * A[7] - A[1] * A[8]) / d;
  t[2] = (A[1] * A[5] - A[2] * A[4]) / d;
  t[3] = (A[5] * A[6] - A[3] * A[8]) / d;
  t[4] = (A[0] * A[8] - A[2] * A[6]) / d;
  t[5] = (A[2] * A[3] - A[0] * A[5]) / d;
  t[6] = (A[3] * A[7] - A[4] * A[6]) / d;
  t[7] = (A[1] * A[6] - A[0] * A[7]) / d;
  t[8] = (A[0] * A[4] - A[1] * A[3]) / d;
  memcpy(C, t, sizeof t);
}

/* Cyclic complex Jacobi: each rotation J = diag-phase * real Givens zeroes
 * one off-diagonal element of the Hermitian matrix; A <- J^H A J, Q <- Q J. */
void orc_herm3_eig(const double* Ain, double* w, double* Qout) {
  cplx A[9], Q[9];
  memcpy(A, Ain, sizeof A);
  for (int i = 0; i < 9; ++i) Q[i] = (i % 4 == 0) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = cabs(A[1]) + cabs(A[2]) + cabs(A[5]);
    double diag = fabs(creal(A[0])) + fabs(creal(A[4])) + fabs(creal(A[8]));
    if (off <= 1e-300 || off <= 1e-18 * diag) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        cplx apq = A[3 * p + q];
        double g = cabs(apq);
        if (g == 0.0) continue;
        cplx ph = apq / g; /* e^{i phi} */
        double app = creal(A[4 * p]), aqq = creal(A[4 * q]);
        double tau = (aqq - app) / (2.0 * g);
        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        cplx J[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        J[4 * p] = c;
        J[3 * p + q] = s;
        J[3 * q + p] = -s * conj(ph);
        J[4 * q] = c * conj(ph);
        cplx tmp[9];
        m3_mul(A, J, tmp);
        m3_dag_mul(J, tmp, A);
        A[3 * p + q] = 0.0;
        A[3 * q + p] = 0.0;
        m3_mul(Q, J, Q);
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = creal(A[4 * i]);
  memcpy(Qout, Q, sizeof Q);
}

/* timeslice.cpp:90-104.  (M^H M)^{1/2} is the principal square root; for a
 * Hermitian positive definite argument it is Q diag(sqrt w) Q^H.  The
 * reference then inverts it (Eigen .inverse(), cofactors for 3x3) and
 * multiplies by the principal det^{-1/3}. */
void orc_proj_su3(double* Md) {
  cplx* M = (cplx*)Md;
  cplx H[9], Q[9], S[9], Si[9];
  double w[3];
  m3_dag_mul(M, M, H);
  orc_herm3_eig((const double*)H, w, (double*)Q);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      cplx acc = 0;
      for (int k = 0; k < 3; ++k) acc += Q[3 * a + k] * sqrt(w[k]) * conj(Q[3 * b + k]);
      S[3 * a + b] = acc;
    }
  m3_inv(S, Si);
  m3_mul(M, Si, M);
  cplx det = m3_det(M);
  cplx fac = 1.0 / cpow(det, 1.0 / 3.0);
  for (int i = 0; i < 9; ++i) M[i] *= fac;
}

/* timeslice.cpp:106-140 */
static void get_dirs(int mu, int* d) {
  if (mu == 0) { d[0] = 1; d[1] = 2; }
  else if (mu == 1) { d[0] = 0; d[1] = 2; }
  else { d[0] = 0; d[1] = 1; }
}
static int decor(int dir, int plane) {
  if (dir == 0) return plane == 1 ? 0 : 3;
  if (dir == 1) return plane == 0 ? 1 : 4;
  return plane == 0 ? 2 : 5;
}

/* timeslice.cpp:243-369 */
void orc_hyp(double* Ud, int L, double alpha1, double alpha2, int iter) {
  const int V3 = L * L * L;
  cplx* U = (cplx*)Ud;
  cplx* dec = (cplx*)malloc(sizeof(cplx) * 9 * 6 * (size_t)V3);
  cplx* nxt = (cplx*)malloc(sizeof(cplx) * 9 * 3 * (size_t)V3);
#define LNK(s, d) (U + ((size_t)(s) * 3 + (d)) * 9)
#define DEC(s, d) (dec + ((size_t)(s) * 6 + (d)) * 9)
  for (int run = 0; run < iter; ++run) {
    /* level 1: decorated links, timeslice.cpp:263-306 */
#pragma omp parallel for schedule(static)
    for (int vol = 0; vol < V3; ++vol) {
      for (int dir = 0; dir < 3; ++dir) {
        int perp[2];
        get_dirs(dir, perp);
        for (int n_el = 0; n_el < 2; ++n_el) {
          int pd = perp[n_el];
          cplx st[9], t1[9], t2[9];
          int mu = orc_up(L, vol, pd);
          int eta = orc_up(L, vol, dir);
          m3_mul_dag(LNK(mu, dir), LNK(eta, pd), t1);
          m3_mul(LNK(vol, pd), t1, st);
          mu = orc_dn(L, vol, pd);
          int nu = orc_up(L, mu, dir);
          m3_mul(LNK(mu, dir), LNK(nu, pd), t1);
          m3_dag_mul(LNK(mu, pd), t1, t2);
          cplx stac[9];
          for (int i = 0; i < 9; ++i)
            stac[i] = LNK(vol, dir)[i] * (1.0 - alpha2) + (st[i] + t2[i]) * (alpha2 / 2.0);
          orc_proj_su3((double*)stac);
          memcpy(DEC(vol, n_el == 0 ? dir : dir + 3), stac, sizeof stac);
        }
      }
    }
    /* level 2: outer staples, timeslice.cpp:311-351 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < V3; ++i) {
      for (int dir = 0; dir < 3; ++dir) {
        cplx outer[9] = {0};
        for (int nd = 0; nd < 3; ++nd) {
          if (nd == dir) continue;
          int plane = ((dir + 1) ^ (nd + 1)) - 1;
          int a = decor(nd, plane), b = decor(dir, plane);
          int mu = orc_up(L, i, nd);
          int nu = orc_up(L, mu, dir);
          int eta = orc_dn(L, nu, nd);
          cplx t1[9], t2[9];
          m3_mul_dag(DEC(mu, b), DEC(eta, a), t1);
          m3_mul(DEC(i, a), t1, t2);
          for (int k = 0; k < 9; ++k) outer[k] += t2[k];
          mu = orc_dn(L, i, nd);
          nu = orc_up(L, mu, dir);
          m3_mul(DEC(mu, b), DEC(nu, a), t1);
          m3_dag_mul(DEC(mu, a), t1, t2);
          for (int k = 0; k < 9; ++k) outer[k] += t2[k];
        }
        cplx* o = nxt + ((size_t)i * 3 + dir) * 9;
        for (int k = 0; k < 9; ++k) o[k] = LNK(i, dir)[k] * (1.0 - alpha1) + outer[k] * (alpha1 / 4.0);
      }
    }
    /* projection and write-back, timeslice.cpp:352-357 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < V3 * 3; ++i) {
      orc_proj_su3((double*)(nxt + (size_t)i * 9));
      memcpy(U + (size_t)i * 9, nxt + (size_t)i * 9, sizeof(cplx) * 9);
    }
  }
#undef LNK
#undef DEC
  free(dec);
  free(nxt);
}

/* ------------------------------------------------------------------------ */
/* shell_matop.cpp:16-103 */
void orc_tv2_ca(const double* Ud, int L, double c, double a, const double* xd, double* yd) {
  const int V3 = L * L * L;
  const cplx* U = (const cplx*)Ud;
  const cplx* x = (const cplx*)xd;
  cplx* y = (cplx*)yd;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < V3; ++k) {
    cplx acc[3] = {0, 0, 0};
    for (int mu = 0; mu < 3; ++mu) {
      int up = orc_up(L, k, mu), dn = orc_dn(L, k, mu);
      const cplx* Uf = U + ((size_t)k * 3 + mu) * 9;
      const cplx* Ub = U + ((size_t)dn * 3 + mu) * 9;
      const cplx* xu = x + 3 * (size_t)up;
      const cplx* xm = x + 3 * (size_t)dn;
      for (int r = 0; r < 3; ++r) {
        acc[r] += Uf[3 * r] * xu[0] + Uf[3 * r + 1] * xu[1] + Uf[3 * r + 2] * xu[2];
        acc[r] += conj(Ub[r]) * xm[0] + conj(Ub[3 + r]) * xm[1] + conj(Ub[6 + r]) * xm[2];
      }
    }
    for (int r = 0; r < 3; ++r) y[3 * (size_t)k + r] = c * (acc[r] - 6.0 * x[3 * (size_t)k + r]) + a * x[3 * (size_t)k + r];
  }
}

void orc_tv2(const double* U, int L, double lamL, double lamC, const double* x, double* y) {
  const double c = 2.0 / (lamL - lamC); /* shell_matop.cpp:49-50 */
  const double a = 1.0 + c * lamC;
  orc_tv2_ca(U, L, c, a, x, y);
}

/* shell_matop.cpp:136-168: T0 = x, T1 = Bx, T_n = 2 B T_{n-1} - T_{n-2} */
int orc_tv_iter(const double* U, int L, double lamL, double lamC, int deg, const double* xd,
                double* yd) {
  if (deg < 1) return -1;
  const size_t n = 3 * (size_t)L * L * L;
  cplx* Told = (cplx*)malloc(sizeof(cplx) * n);
  cplx* Tcur = (cplx*)malloc(sizeof(cplx) * n);
  cplx* tmp = (cplx*)malloc(sizeof(cplx) * n);
  cplx* y = (cplx*)yd;
  orc_tv2(U, L, lamL, lamC, xd, (double*)Tcur);
  memcpy(Told, xd, sizeof(cplx) * n);
  for (int k = 2; k <= deg; ++k) {
    orc_tv2(U, L, lamL, lamC, (const double*)Tcur, (double*)tmp);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
      cplx t = 2.0 * tmp[i] - Told[i];
      Told[i] = Tcur[i];
      Tcur[i] = t;
    }
  }
  memcpy(y, Tcur, sizeof(cplx) * n);
  free(Told);
  free(Tcur);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Real symmetric eigensolver: Householder tridiagonalisation with Q
 * accumulation, then implicit QL with Wilkinson-type shifts (the
 * Wilkinson-Reinsch tred2/tql2 scheme, Handbook for Automatic Computation
 * vol. II).  A is m x m row-major; on return column j of A is the
 * eigenvector of w[j], w ascending. */
int orc_dsyev(int m, double* A, double* w) {
  double* Q = (double*)calloc((size_t)m * m, sizeof(double));
  double* v = (double*)malloc(sizeof(double) * m);
  double* p = (double*)malloc(sizeof(double) * m);
  double* e = (double*)calloc((size_t)m, sizeof(double));
  for (int i = 0; i < m; ++i) Q[(size_t)i * m + i] = 1.0;
  for (int k = 0; k + 2 < m; ++k) {
    double nrm = 0;
    for (int i = k + 1; i < m; ++i) nrm += A[(size_t)i * m + k] * A[(size_t)i * m + k];
    nrm = sqrt(nrm);
    if (nrm == 0.0) continue;
    double x0 = A[(size_t)(k + 1) * m + k];
    double alpha = x0 > 0 ? -nrm : nrm;
    for (int i = k + 1; i < m; ++i) v[i] = A[(size_t)i * m + k];
    v[k + 1] -= alpha;
    double vn2 = 0;
    for (int i = k + 1; i < m; ++i) vn2 += v[i] * v[i];
    if (vn2 == 0.0) continue;
    double beta = 2.0 / vn2;
    /* p = beta * A22 v ; K = beta/2 * v.p ; q = p - K v ; A22 -= v q^T + q v^T */
    for (int i = k + 1; i < m; ++i) {
      double acc = 0;
      for (int j = k + 1; j < m; ++j) acc += A[(size_t)i * m + j] * v[j];
      p[i] = beta * acc;
    }
    double K = 0;
    for (int i = k + 1; i < m; ++i) K += v[i] * p[i];
    K *= 0.5 * beta;
    for (int i = k + 1; i < m; ++i) p[i] -= K * v[i];
    for (int i = k + 1; i < m; ++i)
      for (int j = k + 1; j < m; ++j) A[(size_t)i * m + j] -= v[i] * p[j] + p[i] * v[j];
    A[(size_t)(k + 1) * m + k] = alpha;
    A[(size_t)k * m + k + 1] = alpha;
    for (int i = k + 2; i < m; ++i) {
      A[(size_t)i * m + k] = 0;
      A[(size_t)k * m + i] = 0;
    }
    /* Q <- Q H */
    for (int r = 0; r < m; ++r) {
      double acc = 0;
      for (int j = k + 1; j < m; ++j) acc += Q[(size_t)r * m + j] * v[j];
      acc *= beta;
      for (int j = k + 1; j < m; ++j) Q[(size_t)r * m + j] -= acc * v[j];
    }
  }
  for (int i = 0; i < m; ++i) w[i] = A[(size_t)i * m + i];
  for (int i = 0; i + 1 < m; ++i) e[i] = A[(size_t)(i + 1) * m + i];
  e[m - 1] = 0;
  int status = 0;
  for (int l = 0; l < m; ++l) {
    int iter = 0, mm;
    do {
      for (mm = l; mm < m - 1; ++mm) {
        double dd = fabs(w[mm]) + fabs(w[mm + 1]);
        if (fabs(e[mm]) <= 2.2e-16 * dd) break;
      }
      if (mm != l) {
        if (iter++ == 100) { status = -1; break; }
        double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = w[mm] - w[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, pp = 0.0;
        int i;
        for (i = mm - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = hypot(f, g));
          if (r == 0.0) {
            w[i + 1] -= pp;
            e[mm] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = w[i + 1] - pp;
          r = (w[i] - g) * s + 2.0 * c * b;
          w[i + 1] = g + (pp = s * r);
          g = c * r - b;
          for (int k = 0; k < m; ++k) {
            f = Q[(size_t)k * m + i + 1];
            Q[(size_t)k * m + i + 1] = s * Q[(size_t)k * m + i] + c * f;
            Q[(size_t)k * m + i] = c * Q[(size_t)k * m + i] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        w[l] -= pp;
        e[l] = g;
        e[mm] = 0.0;
      }
    } while (mm != l);
  }
  /* sort ascending (selection sort keeps the column pairing) */
  for (int i = 0; i < m - 1; ++i) {
    int kmin = i;
    for (int j = i + 1; j < m; ++j)
      if (w[j] < w[kmin]) kmin = j;
    if (kmin != i) {
      double t = w[i];
      w[i] = w[kmin];
      w[kmin] = t;
      for (int r = 0; r < m; ++r) {
        t = Q[(size_t)r * m + i];
        Q[(size_t)r * m + i] = Q[(size_t)r * m + kmin];
        Q[(size_t)r * m + kmin] = t;
      }
    }
  }
  memcpy(A, Q, sizeof(double) * (size_t)m * m);
  free(Q);
  free(v);
  free(p);
  free(e);
  return status;
}

/* ------------------------------------------------------------------------ */
/* Seeded start vector: splitmix64 stream -> Box-Muller Gaussians.           */
static uint64_t splitmix64(uint64_t* s) {
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
static double u01(uint64_t* s) { return ((splitmix64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

/* Krylov-Schur for Hermitian operators (Stewart, SIAM J. Matrix Anal. Appl.
 * 23 (2001); thick-restart Lanczos of Wu & Simon, SIMAX 22 (2000)), with
 * full re-orthogonalisation (classical Gram-Schmidt, two passes), in the
 * configuration of ev_ts.cpp:231-263: Hermitian problem, nev eigenpairs of
 * largest magnitude, relative convergence criterion |r_i| / |theta_i| < tol
 * on the Krylov residual estimate |beta_m s_{m,i}|, restart keeping the
 * converged Ritz vectors plus half of the remaining ones. */
int orc_eps_solve(const double* U, int L, double lamL, double lamC, int deg, int nev, int ncv,
                  double tol, int max_it, unsigned long long seed, double* theta_out,
                  double* Vout, long long* stats) {
  const size_t n = 3 * (size_t)L * L * L;
  if (nev < 1 || (size_t)nev > n) return -1;
  if (ncv <= 0) ncv = (2 * nev > nev + 15) ? 2 * nev : nev + 15;
  if ((size_t)ncv > n) ncv = (int)n;
  if (ncv <= nev) return -1;
  if (max_it <= 0) max_it = (int)((2 * n / ncv) > 100 ? (2 * n / ncv) : 100);
  const int m = ncv;
  cplx* V = (cplx*)malloc(sizeof(cplx) * n * (m + 1));
  cplx* w = (cplx*)malloc(sizeof(cplx) * n);
  cplx* h = (cplx*)malloc(sizeof(cplx) * (m + 1));
  cplx* h2 = (cplx*)malloc(sizeof(cplx) * (m + 1));
  double* H = (double*)calloc((size_t)m * m, sizeof(double));
  double* S = (double*)malloc(sizeof(double) * m * m);
  double* th = (double*)malloc(sizeof(double) * m);
  int* ord = (int*)malloc(sizeof(int) * m);
  cplx* tmpV = (cplx*)malloc(sizeof(cplx) * n * m);
  long long nmat = 0;
  uint64_t st = seed;
  double nrm = 0;
  for (size_t i = 0; i < n; ++i) {
    double r = sqrt(-2.0 * log(u01(&st))), ph = 2.0 * M_PI * u01(&st);
    V[i] = r * cos(ph) + I * r * sin(ph);
    nrm += creal(V[i]) * creal(V[i]) + cimag(V[i]) * cimag(V[i]);
  }
  nrm = sqrt(nrm);
  for (size_t i = 0; i < n; ++i) V[i] /= nrm;
  int k = 0, nconv = 0, it;
  double beta = 0;
  for (it = 0; it < max_it; ++it) {
    /* Lanczos expansion of columns k .. m-1 */
    for (int j = k; j < m; ++j) {
      orc_tv_iter(U, L, lamL, lamC, deg, (const double*)(V + (size_t)j * n), (double*)w);
      ++nmat;
      for (int pass = 0; pass < 2; ++pass) {
        cplx* hh = pass == 0 ? h : h2;
#pragma omp parallel for schedule(static)
        for (int i = 0; i <= j; ++i) {
          const cplx* vi = V + (size_t)i * n;
          cplx acc = 0;
          for (size_t r = 0; r < n; ++r) acc += conj(vi[r]) * w[r];
          hh[i] = acc;
        }
#pragma omp parallel for schedule(static)
        for (size_t r = 0; r < n; ++r) {
          cplx acc = 0;
          for (int i = 0; i <= j; ++i) acc += V[(size_t)i * n + r] * hh[i];
          w[r] -= acc;
        }
      }
      for (int i = 0; i <= j; ++i) h[i] += h2[i];
      double b = 0;
      for (size_t r = 0; r < n; ++r) b += creal(w[r]) * creal(w[r]) + cimag(w[r]) * cimag(w[r]);
      b = sqrt(b);
      for (int i = 0; i <= j; ++i) {
        H[(size_t)i * m + j] = creal(h[i]);
        H[(size_t)j * m + i] = creal(h[i]);
      }
      beta = b;
      if (j + 1 < m) {
        H[(size_t)(j + 1) * m + j] = b;
        H[(size_t)j * m + j + 1] = b;
      }
      cplx* vn = V + (size_t)(j + 1) * n;
      for (size_t r = 0; r < n; ++r) vn[r] = w[r] / b;
    }
    /* projected problem */
    memcpy(S, H, sizeof(double) * m * m);
    orc_dsyev(m, S, th);
    for (int i = 0; i < m; ++i) ord[i] = i;
    for (int i = 0; i < m; ++i) /* stable sort by |theta| descending */
      for (int j = i + 1; j < m; ++j)
        if (fabs(th[ord[j]]) > fabs(th[ord[i]])) {
          int t = ord[i];
          ord[i] = ord[j];
          ord[j] = t;
        }
    nconv = 0;
    for (int i = 0; i < m; ++i) {
      int c = ord[i];
      double res = fabs(beta * S[(size_t)(m - 1) * m + c]);
      if (res < tol * fabs(th[c])) ++nconv;
      else break;
    }
    int done = nconv >= nev || it == max_it - 1;
    int keep = done ? nev : nconv + (m - nconv) / 2;
    if (keep < nev && !done) keep = nev < m - 1 ? nev : m - 1;
    if (keep > m - 1) keep = m - 1;
    /* V[:, 0..keep) <- V_m S[:, ord[0..keep)] */
#pragma omp parallel for schedule(static)
    for (size_t r = 0; r < n; ++r) {
      for (int i = 0; i < keep; ++i) {
        cplx acc = 0;
        int c = ord[i];
        for (int j = 0; j < m; ++j) acc += V[(size_t)j * n + r] * S[(size_t)j * m + c];
        tmpV[(size_t)i * n + r] = acc;
      }
    }
    memcpy(V, tmpV, sizeof(cplx) * n * keep);
    if (done) {
      for (int i = 0; i < nev; ++i) theta_out[i] = th[ord[i]];
      break;
    }
    memcpy(V + (size_t)keep * n, V + (size_t)m * n, sizeof(cplx) * n);
    memset(H, 0, sizeof(double) * m * m);
    for (int i = 0; i < keep; ++i) {
      int c = ord[i];
      H[(size_t)i * m + i] = th[c];
      double b = beta * S[(size_t)(m - 1) * m + c];
      H[(size_t)i * m + keep] = b;
      H[(size_t)keep * m + i] = b;
    }
    k = keep;
  }
  memcpy(Vout, V, sizeof(cplx) * n * nev);
  if (stats) {
    stats[0] = it;
    stats[1] = nmat;
  }
  free(V);
  free(w);
  free(h);
  free(h2);
  free(H);
  free(S);
  free(th);
  free(ord);
  free(tmpV);
  return nconv;
}

/* ------------------------------------------------------------------------ */
/* recover_spec.cpp:13-54 */
double orc_invert_B(double b, double lamL, double lamC) { return (lamL - lamC) * (b - 1) * 0.5 - lamC; }
double orc_invert_Tn(double t, int deg) { return cosh(1.0 / deg * acosh(t)); }
void orc_recover_spectrum(int nev, const double* theta, int deg, double lamL, double lamC,
                          double* evals) {
  for (int i = 0; i < nev; ++i) evals[i] = orc_invert_B(orc_invert_Tn(theta[i], deg), lamL, lamC);
}

/* eigensystem.cpp:4-30: phi_n = arg V(0,n); V_fix(:,n) = e^{-i phi_n} V(:,n) */
void orc_fix_phase(int n, int nev, const double* Vd, double* Vfd, double* phase) {
  const cplx* V = (const cplx*)Vd;
  cplx* Vf = (cplx*)Vfd;
  for (int c = 0; c < nev; ++c) {
    double ph = carg(V[(size_t)c * n]);
    phase[c] = ph;
    cplx fac = cexp(-I * ph);
    for (int r = 0; r < n; ++r) Vf[(size_t)c * n + r] = fac * V[(size_t)c * n + r];
  }
}

/* eigensystem.cpp:32-57 */
void orc_reset_phase(int n, int nev, const double* Vd, const double* want, double* Vrd) {
  const cplx* V = (const cplx*)Vd;
  cplx* Vr = (cplx*)Vrd;
  for (int c = 0; c < nev; ++c) {
    double ph = carg(V[(size_t)c * n]);
    cplx fac = cexp(-I * (ph - want[c]));
    for (int r = 0; r < n; ++r) Vr[(size_t)c * n + r] = fac * V[(size_t)c * n + r];
  }
}

/* read_write.cpp:20-35 (eps = 10e-10, one-sided on |tr| - nev) */
int orc_check_trace(int n, int nev, const double* Vd, double* tr2, double* sum2) {
  const cplx* V = (const cplx*)Vd;
  cplx tr = 0, sum = 0;
  for (int i = 0; i < nev; ++i)
    for (int j = 0; j < nev; ++j) {
      cplx acc = 0;
      for (int r = 0; r < n; ++r) acc += conj(V[(size_t)i * n + r]) * V[(size_t)j * n + r];
      sum += acc;
      if (i == j) tr += acc;
    }
  if (tr2) { tr2[0] = creal(tr); tr2[1] = cimag(tr); }
  if (sum2) { sum2[0] = creal(sum); sum2[1] = cimag(sum); }
  const double eps = 10e-10;
  if (fabs(creal(tr)) - nev > eps || fabs(cimag(tr)) > eps || fabs(creal(sum)) - nev > eps ||
      fabs(cimag(sum)) > eps)
    return 0;
  return 1;
}
