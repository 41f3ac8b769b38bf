"""Writes tests/golden/c1_8x8x8.npz: configuration C1 of the per-timeslice
path (one 8^3 timeslice of the seeded field U = exp(i 0.4 H), seed 1227,
HYP 0.62/0.62/3, N_ev = 20, DEG = 8, lambda_l = 12, tolerance 1e-12).

Two independent sources: the C oracle (map + HYP + Krylov-Schur + recover +
fix_phase) and numpy's dense Hermitian eigensolver (LAPACK zheevd) on the
assembled 1536 x 1536 Laplacian of the numpy HYP restatement.  The fixture
stores the dense-eigh results; the script asserts that the oracle agrees.

Run from the repository root: python tests/golden/make_c1.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import lapheig_np as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

L, SEED, EPS, NEV, DEG, LAM_L, TOL = 8, 1227, 0.4, 20, 8, 12.0, 1e-12


def main():
    cfg = N.synth_config(SEED, EPS, L, 1)
    U0 = N.links_from_config(cfg, L)
    U = N.hyp(U0, 0.62, 0.62, 3)
    D = N.dense_D(U)
    w, Q = np.linalg.eigh(D)
    order = np.argsort(w)[::-1]  # eigenvalues of D, nearest 0 first
    w, Q = w[order], Q[:, order]
    lam_c = float(-w[NEV + 4] * 1.05)
    V = Q[:, :NEV].T.copy()  # (nev, n)
    Vf, ph = O.fix_phase(V)
    # the oracle path must agree
    Uo = O.hyp(O.map_timeslice(cfg, L), L, 0.62, 0.62, 3)
    assert abs(Uo.reshape(U.shape) - U).max() < 1e-13
    nconv, theta, Vo, _ = O.eps_solve(Uo, L, LAM_L, lam_c, DEG, NEV, tol=TOL, seed=SEED)
    ev_o = O.recover_spectrum(theta, DEG, LAM_L, lam_c)
    assert nconv >= NEV and np.allclose(ev_o, w[:NEV], rtol=1e-10, atol=1e-13)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_8x8x8.npz")
    np.savez_compressed(out, seed=SEED, eps=EPS, L=L, nev=NEV, deg=DEG, lambda_l=LAM_L, lambda_c=lam_c, tol=TOL,
                        alpha=np.array([0.62, 0.62]), hyp_iter=3,
                        links_hyp=U.reshape(-1, 3, 3, 3), evals=w[:NEV + 8], evecs_fixed=Vf, phases=ph)
    print(f"wrote {out}: lambda_c={lam_c:.6f} evals[:3]={w[:3]}")


if __name__ == "__main__":
    main()
