"""Writes tests/golden/c2_24x24x24.npz: configuration C2 of the per-timeslice
path (24^3 x 48 seeded field U = exp(i 0.4 H), seed 1227, HYP 0.62/0.62/3,
N_ev = 120), timeslices 0 and 47.

Source: scipy's ARPACK (eigsh, shift-invert about -0.05 with a sparse LU) on
-D assembled entry by entry from the numpy HYP restatement's links
(oracle/lapheig_np.sparse_minus_D, the BuildLaplacian layout).  Timeslice 0
is cross-checked against the C oracle's Krylov-Schur on T_12(B).  Stored:
the 128 eigenvalues of D nearest 0 (descending), the first 8 sites (24
components) of the 120 phase-fixed eigenvectors, and the phase-invariant
overlaps |g^H v_i|^2 with a seeded Gaussian vector g.

Run from the repository root: python tests/golden/make_c2.py   (a few minutes)
"""
import os
import sys
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import lapheig_np as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

L, T, SEED, EPS, NEV, NKEEP, DEG, LAM_L = 24, 48, 1227, 0.4, 120, 128, 12, 12.0
TSLICES = (0, 47)
G_SEED = 99


def eig_lowest(U, k):
    A = N.sparse_minus_D(U)
    sigma = -0.05
    lu = sla.splu((A - sigma * sp.identity(A.shape[0], dtype=A.dtype, format="csr")).tocsc())
    op = sla.LinearOperator(A.shape, matvec=lu.solve, dtype=A.dtype)
    nu, X = sla.eigsh(op, k=k, which="LM", tol=1e-14, ncv=2 * k + 16)
    w = sigma + 1.0 / nu  # eigenvalues of -D
    order = np.argsort(w)
    w, X = w[order], X[:, order]
    res = np.linalg.norm(A @ X - X * w[None, :], axis=0)
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
        print(f"t={t}: lambda_1..3 {-w[:3]}, lambda_120 {-w[NEV - 1]:.10f}, max res {res.max():.1e} "
              f"({time.time() - t0:.0f}s)", flush=True)
        if t == 0:  # the C oracle's Krylov-Schur on T_12(B) must agree
            Uo = O.hyp(O.map_timeslice(ts, L), L, 0.62, 0.62, 3)
            assert abs(Uo.reshape(U.shape) - U).max() < 1e-12
            lam_c = 1.05 * w[NEV + 4]
            nconv, theta, _, _ = O.eps_solve(Uo, L, LAM_L, lam_c, DEG, NEV, tol=1e-10, seed=SEED)
            ev_o = O.recover_spectrum(theta, DEG, LAM_L, lam_c)
            assert nconv >= NEV and np.allclose(ev_o, -w[:NEV], rtol=1e-9, atol=1e-12), \
                np.abs(ev_o + w[:NEV]).max()
            print(f"C oracle agrees ({time.time() - t0:.0f}s)", flush=True)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c2_24x24x24.npz")
    np.savez_compressed(path, seed=SEED, eps=EPS, L=L, T=T, nev=NEV, deg=DEG, lambda_l=LAM_L,
                        alpha=np.array([0.62, 0.62]), hyp_iter=3, tslices=np.array(TSLICES), g_seed=G_SEED, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
