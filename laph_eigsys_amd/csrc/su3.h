// Device-side complex 3x3 algebra in FP64 (double2 = re, im), row-major
// element (a,b) at e[3a+b].  Used by the HYP kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace laph {

struct M3 {
  double2 e[9];
};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ double cabs2(double2 a) { return sqrt(a.x * a.x + a.y * a.y); }
// a / b
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

__device__ __forceinline__ M3 load_m3(const double2* p) {
  M3 m;
#pragma unroll
  for (int i = 0; i < 9; ++i) m.e[i] = p[i];
  return m;
}
__device__ __forceinline__ void store_m3(double2* p, const M3& m) {
#pragma unroll
  for (int i = 0; i < 9; ++i) p[i] = m.e[i];
}

// A * B
__device__ __forceinline__ M3 mul(const M3& A, const M3& B) {
  M3 C;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double2 acc = cmul(A.e[3 * a], B.e[b]);
      acc = cadd(acc, cmul(A.e[3 * a + 1], B.e[3 + b]));
      acc = cadd(acc, cmul(A.e[3 * a + 2], B.e[6 + b]));
      C.e[3 * a + b] = acc;
    }
  return C;
}
// A * B^H
__device__ __forceinline__ M3 mul_dag(const M3& A, const M3& B) {
  M3 C;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double2 acc = cmul(A.e[3 * a], cconj(B.e[3 * b]));
      acc = cadd(acc, cmul(A.e[3 * a + 1], cconj(B.e[3 * b + 1])));
      acc = cadd(acc, cmul(A.e[3 * a + 2], cconj(B.e[3 * b + 2])));
      C.e[3 * a + b] = acc;
    }
  return C;
}
// A^H * B
__device__ __forceinline__ M3 dag_mul(const M3& A, const M3& B) {
  M3 C;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double2 acc = cmul(cconj(A.e[a]), B.e[b]);
      acc = cadd(acc, cmul(cconj(A.e[3 + a]), B.e[3 + b]));
      acc = cadd(acc, cmul(cconj(A.e[6 + a]), B.e[6 + b]));
      C.e[3 * a + b] = acc;
    }
  return C;
}

__device__ __forceinline__ double2 det3(const M3& A) {
  const double2* m = A.e;
  double2 t0 = csub(cmul(m[4], m[8]), cmul(m[5], m[7]));
  double2 t1 = csub(cmul(m[3], m[8]), cmul(m[5], m[6]));
  double2 t2 = csub(cmul(m[3], m[7]), cmul(m[4], m[6]));
  return cadd(csub(cmul(m[0], t0), cmul(m[1], t1)), cmul(m[2], t2));
}

// Cofactor inverse.
__device__ __forceinline__ M3 inv3(const M3& A) {
  const double2* m = A.e;
  double2 d = det3(A);
  M3 C;
  C.e[0] = cdiv(csub(cmul(m[4], m[8]), cmul(m[5], m[7])), d);
  C.e[1] = cdiv(csub(cmul(m[2], m[7]), cmul(m[1], m[8])), d);
  C.e[2] = cdiv(csub(cmul(m[1], m[5]), cmul(m[2], m[4])), d);
  C.e[3] = cdiv(csub(cmul(m[5], m[6]), cmul(m[3], m[8])), d);
  C.e[4] = cdiv(csub(cmul(m[0], m[8]), cmul(m[2], m[6])), d);
  C.e[5] = cdiv(csub(cmul(m[2], m[3]), cmul(m[0], m[5])), d);
  C.e[6] = cdiv(csub(cmul(m[3], m[7]), cmul(m[4], m[6])), d);
  C.e[7] = cdiv(csub(cmul(m[1], m[6]), cmul(m[0], m[7])), d);
  C.e[8] = cdiv(csub(cmul(m[0], m[4]), cmul(m[1], m[3])), d);
  return C;
}

// Hermitian eigen-decomposition A = Q diag(w) Q^H by cyclic complex Jacobi
// rotations (each rotation: diagonal phase times a real Givens rotation,
// zeroing one off-diagonal element).  A is destroyed.
__device__ __forceinline__ void herm3_eig(M3& A, double w[3], M3& Q) {
#pragma unroll
  for (int i = 0; i < 9; ++i) Q.e[i] = make_double2((i % 4 == 0) ? 1.0 : 0.0, 0.0);
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = cabs2(A.e[1]) + cabs2(A.e[2]) + cabs2(A.e[5]);
    double diag = fabs(A.e[0].x) + fabs(A.e[4].x) + fabs(A.e[8].x);
    if (off <= 1e-300 || off <= 1e-18 * diag) break;
#pragma unroll
    for (int pq = 0; pq < 3; ++pq) {
      const int p = pq == 2 ? 1 : 0;
      const int q = pq == 0 ? 1 : 2;
      double g = cabs2(A.e[3 * p + q]);
      if (g == 0.0) continue;
      double2 ph = cscale(A.e[3 * p + q], 1.0 / g);
      double tau = (A.e[4 * q].x - A.e[4 * p].x) / (2.0 * g);
      double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
      double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
      double2 Jqp = cscale(cconj(ph), -s), Jqq = cscale(cconj(ph), c);
#pragma unroll
      for (int k = 0; k < 3; ++k) {  // columns p, q of A and Q
        double2 akp = A.e[3 * k + p], akq = A.e[3 * k + q];
        A.e[3 * k + p] = cadd(cscale(akp, c), cmul(akq, Jqp));
        A.e[3 * k + q] = cadd(cscale(akp, s), cmul(akq, Jqq));
        double2 qkp = Q.e[3 * k + p], qkq = Q.e[3 * k + q];
        Q.e[3 * k + p] = cadd(cscale(qkp, c), cmul(qkq, Jqp));
        Q.e[3 * k + q] = cadd(cscale(qkp, s), cmul(qkq, Jqq));
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {  // rows p, q of A
        double2 apk = A.e[3 * p + k], aqk = A.e[3 * q + k];
        A.e[3 * p + k] = cadd(cscale(apk, c), cmulc(Jqp, aqk));
        A.e[3 * q + k] = cadd(cscale(apk, s), cmulc(Jqq, aqk));
      }
      A.e[3 * p + q] = make_double2(0.0, 0.0);
      A.e[3 * q + p] = make_double2(0.0, 0.0);
    }
  }
  w[0] = A.e[0].x;
  w[1] = A.e[4].x;
  w[2] = A.e[8].x;
}

// SU(3) projection M -> M ((M^H M)^{1/2})^{-1} det(.)^{-1/3} with the
// principal cube root (timeslice.cpp:90-104).  The square root of the
// Hermitian positive definite M^H M is Q diag(sqrt w) Q^H.
__device__ __forceinline__ void proj_su3(M3& M) {
  M3 H = dag_mul(M, M);
  double w[3];
  M3 Q;
  herm3_eig(H, w, Q);
  double sw[3] = {sqrt(w[0]), sqrt(w[1]), sqrt(w[2])};
  M3 S;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      double2 acc = cscale(cmul(Q.e[3 * a], cconj(Q.e[3 * b])), sw[0]);
      acc = cadd(acc, cscale(cmul(Q.e[3 * a + 1], cconj(Q.e[3 * b + 1])), sw[1]));
      acc = cadd(acc, cscale(cmul(Q.e[3 * a + 2], cconj(Q.e[3 * b + 2])), sw[2]));
      S.e[3 * a + b] = acc;
    }
  M = mul(M, inv3(S));
  double2 d = det3(M);
  // principal d^{1/3} = |d|^{1/3} e^{i arg(d)/3}; factor = 1 / that
  double r = cbrt(sqrt(d.x * d.x + d.y * d.y));
  double ph = atan2(d.y, d.x) / 3.0;
  double2 root = make_double2(r * cos(ph), r * sin(ph));
  double2 fac = cdiv(make_double2(1.0, 0.0), root);
#pragma unroll
  for (int i = 0; i < 9; ++i) M.e[i] = cmul(M.e[i], fac);
}

}  // namespace laph
