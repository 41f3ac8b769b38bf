"""Writes tests/golden/c3_32x32x32.npz: configuration C3 of the per-timeslice
path (32^3 x 64 seeded field U = exp(i 0.3 H), seed 1227, HYP 0.62/0.62/3,
N_ev = 220, DEG 16), timeslices 0 and 63 -- the headline workload's size.

Source: scipy's ARPACK (eigsh, shift-invert about -0.05 with a sparse LU) on
-D assembled entry by entry from the numpy HYP restatement's links
(oracle/lapheig_np.sparse_minus_D, the BuildLaplacian layout), the same
recipe as make_c2.py.  Timeslice 0 is cross-checked against the C oracle's
Krylov-Schur on T_16(B) (the reference's EPS configuration, ev_ts.cpp:231-303,
with recover_spectrum, recover_spec.cpp:45-54).  Stored: the 232 eigenvalues
of D nearest 0 (descending), the first 8 sites (24 components) of the 220
phase-fixed eigenvectors, and the phase-invariant overlaps |g^H v_i|^2 with a
seeded Gaussian vector g.

Run from the repository root: python tests/golden/make_c3.py   (tens of minutes)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_c2 import eig_lowest  # noqa: E402
from oracle import lapheig_np as N  # noqa: E402
from oracle import oracle as O  # noqa: E402

L, T, SEED, EPS, NEV, NKEEP, DEG, LAM_L = 32, 64, 1227, 0.3, 220, 232, 16, 12.0
TSLICES = (0, 63)
G_SEED = 99


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
