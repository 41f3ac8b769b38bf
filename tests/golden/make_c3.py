"""Writes tests/golden/c3_32x32x32.npz: configuration C3 of the per-timeslice
path (32^3 x 64 seeded field U = exp(i 0.3 H), seed 1227, HYP 0.62/0.62/3,
N_ev = 220, DEG 16), timeslices 0 and 63 -- the headline workload's size.

Source: scipy's ARPACK (eigsh, largest magnitude) on T_16(B), B = c D + a,
the Chebyshev filter of the reference's MatShell (shell_matop.cpp:16-168)
applied with -D assembled entry by entry from the numpy HYP restatement's
links (oracle/lapheig_np.sparse_minus_D, the BuildLaplacian layout) -- the
reference's own method (ev_ts.cpp:231-303), in a Krylov code independent of
this project's solvers.  lambda_c = 1.35 x the 232nd free-field level, above
the 232 wanted eigenvalues (asserted).  The eigenvalues stored are the
Rayleigh quotients v^H D v of the converged vectors, with their residuals
|D v - lambda v|.  Timeslice 0 is cross-checked against the C oracle's
Krylov-Schur on T_16(B) with recover_spectrum (recover_spec.cpp:45-54).
(Shift-invert with a sparse LU, the C2 recipe, does not finish in useful time
at n = 98304.)  Stored: the 232 eigenvalues of D nearest 0 (descending), the
first 8 sites (24 components) of the 220 phase-fixed eigenvectors, the
phase-invariant overlaps |g^H v_i|^2 with a seeded Gaussian vector g, and
the residuals.

Run from the repository root: python tests/golden/make_c3.py   (~5 minutes on 8 cores)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import scipy.sparse as sp  # noqa: E402
import scipy.sparse.linalg as sla  # noqa: E402

from oracle import lapheig_np as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

L, T, SEED, EPS, NEV, NKEEP, DEG, LAM_L = 32, 64, 1227, 0.3, 220, 232, 16, 12.0
TSLICES = (0, 63)
G_SEED = 99


def free_level(L, k):
    s2 = 4 * np.sin(np.pi * np.arange(L) / L) ** 2
    return np.sort(np.repeat((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel(), 3))[k - 1]


def eig_lowest(U, k):
    """k lowest eigenpairs of -D by ARPACK on T_DEG(c D + a)."""
    A = sp.csr_matrix(N.sparse_minus_D(U))  # -D
    n = A.shape[0]
    lam_c = 1.35 * free_level(L, k)
    c = 2.0 / (LAM_L - lam_c)
    a = 1.0 + c * lam_c

    def B(v):
        return a * v - c * (A @ v)

    def T(v):
        t_old, t_cur = v, B(v)
        for _ in range(2, DEG + 1):
            t_old, t_cur = t_cur, 2.0 * B(t_cur) - t_old
        return t_cur

    op = sla.LinearOperator((n, n), matvec=T, dtype=np.complex128)
    v0 = np.random.default_rng(SEED).standard_normal(n) + 0j
    theta, X = sla.eigsh(op, k=k, which="LM", tol=1e-14, ncv=2 * k + 16, v0=v0)
    AX = A @ X
    w = np.real(np.einsum("ij,ij->j", X.conj(), AX)) / np.real(np.einsum("ij,ij->j", X.conj(), X))
    order = np.argsort(w)
    w, X, AX = w[order], X[:, order], AX[:, order]
    X = X / np.linalg.norm(X, axis=0)[None, :]
    res = np.linalg.norm(A @ X - X * w[None, :], axis=0)
    assert w[-1] < lam_c, (w[-1], lam_c)  # the filter amplifies exactly the wanted part
    return w, X, res


def main():
    t0 = time.time()
    cfg = N.synth_config(SEED, EPS, L, T)
    vol = 72 * L ** 3
    rng = np.random.default_rng(G_SEED)
    g = rng.normal(size=3 * L ** 3) + 1j * rng.normal(size=3 * L ** 3)
    out = {}
    for t in TSLICES:
        ts = cfg[t * vol:(t + 1) * vol]
        U = N.hyp(N.links_from_config(ts, L), 0.62, 0.62, 3)
        w, X, res = eig_lowest(U, NKEEP)
        assert res.max() < 1e-10, res.max()
        V = X[:, :NEV].T  # (nev, n)
        ph = np.angle(V[:, 0])
        Vf = V * np.exp(-1j * ph)[:, None]
        out[f"evals_{t}"] = -w  # eigenvalues of D, nearest 0 first
        out[f"head_{t}"] = Vf[:, :24].copy()
        out[f"overlap_{t}"] = np.abs(V.conj() @ g) ** 2
        out[f"residual_{t}"] = res[:NEV]
        print(f"t={t}: lambda_1..3 {-w[:3]}, lambda_220 {-w[NEV - 1]:.10f}, max res {res.max():.1e} "
              f"({time.time() - t0:.0f}s)", flush=True)
        if t == 0:  # the C oracle's Krylov-Schur on T_16(B) must agree
            Uo = O.hyp(O.map_timeslice(ts, L), L, 0.62, 0.62, 3)
            assert abs(Uo.reshape(U.shape) - U).max() < 1e-12
            lam_c = 1.05 * w[NEV + 4]
            nconv, theta, _, _ = O.eps_solve(Uo, L, LAM_L, lam_c, DEG, NEV, tol=1e-10, seed=SEED)
            ev_o = O.recover_spectrum(theta, DEG, LAM_L, lam_c)
            assert nconv >= NEV and np.allclose(ev_o, -w[:NEV], rtol=1e-9, atol=1e-12), \
                np.abs(ev_o + w[:NEV]).max()
            out["oracle_evals_0"] = ev_o[:NEV]
            print(f"C oracle agrees, max rel diff {np.max(np.abs(ev_o[:NEV] + w[:NEV]) / np.abs(w[:NEV])):.1e} "
                  f"({time.time() - t0:.0f}s)", flush=True)
    path = os.path.join(HERE, "c3_32x32x32.npz")
    np.savez_compressed(path, seed=SEED, eps=EPS, L=L, T=T, nev=NEV, deg=DEG, lambda_l=LAM_L,
                        alpha=np.array([0.62, 0.62]), hyp_iter=3, tslices=np.array(TSLICES), g_seed=G_SEED, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
