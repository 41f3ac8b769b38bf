"""Configuration C1 (8^3, eps 0.4, seed 1227, HYP 0.62/0.62/3, N_ev 20, DEG 8)
against the committed golden fixture (tests/golden/c1_8x8x8.npz, written by
tests/golden/make_c1.py from numpy's dense zheevd): the oracle on the CPU,
the HIP path (HYP, plain FP64 Krylov-Schur, mixed batched solver) on the GPU.

Bar: links 1e-13, eigenvalues 1e-10 relative, phase-fixed eigenvectors 1e-9
per component where the gap to both neighbours exceeds 1e-3, the others by
projector distance, tr(V^H V) = N_ev."""
import os

import numpy as np
import pytest

from oracle import lapheig_np as N
from oracle import oracle as O

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_8x8x8.npz"))
L, NEV, DEG = int(G["L"]), int(G["nev"]), int(G["deg"])
LAM_L, LAM_C, TOL = float(G["lambda_l"]), float(G["lambda_c"]), float(G["tol"])


def config():
    return N.synth_config(int(G["seed"]), float(G["eps"]), L, 1)


def check_vectors(V, gap=1e-3):
    """V (nev, n) eigenvectors (any phase) against the fixture."""
    ev = G["evals"]
    Vf, _ = O.fix_phase(V)
    ref = G["evecs_fixed"]
    i = 0
    while i < NEV:
        j = i + 1
        while j < NEV and abs(ev[j] - ev[j - 1]) < gap:
            j += 1
        iso = (j - i == 1) and abs(ev[i + 1] - ev[i]) >= gap
        if iso:
            assert np.abs(Vf[i] - ref[i]).max() < 1e-9, i
        else:
            P = V[i:j].T @ V[i:j].conj() - ref[i:j].T @ ref[i:j].conj()
            assert np.linalg.norm(P) < 1e-8, (i, j)
        i = j


def test_oracle_matches_golden():
    U = O.hyp(O.map_timeslice(config(), L), L, 0.62, 0.62, 3)
    np.testing.assert_allclose(U, G["links_hyp"], atol=1e-13, rtol=0)
    nconv, theta, V, _ = O.eps_solve(U, L, LAM_L, LAM_C, DEG, NEV, tol=TOL, seed=7)
    ev = O.recover_spectrum(theta, DEG, LAM_L, LAM_C)
    np.testing.assert_allclose(ev, G["evals"][:NEV], rtol=1e-10, atol=1e-13)
    check_vectors(V)
    assert O.check_trace(O.fix_phase(V)[0])[0]


@pytest.mark.gpu
def test_hip_path_matches_golden(gpu):
    from laph_eigsys_amd import lapheig as lp
    cfg = config()
    with lp.Tslice(L, DEG, LAM_L, LAM_C, device=gpu, ts_config=cfg) as t:
        t.smearing_hyp(0.62, 0.62, 3)
        np.testing.assert_allclose(t.get_links(), G["links_hyp"], atol=1e-13, rtol=0)
        theta, V, _ = t.eps_solve(NEV, tol=TOL, seed=11)
    ev = lp.recover_spectrum(theta, DEG, LAM_L, LAM_C)
    np.testing.assert_allclose(ev, G["evals"][:NEV], rtol=1e-10, atol=1e-13)
    check_vectors(V)
    Vf, ph = lp.fix_phase(V)
    assert lp.check_trace(Vf)[0]
    # mixed-precision batched solver on the same timeslice
    b = lp.Batch(L, DEG, LAM_L, 0.0, 1, NEV, nsub=64, tol=1e-12, seed=3, device=gpu)
    b.set_timeslice(0, cfg, 0.62, 0.62, 3)
    b.solve()
    evb, Vb, phb, res = b.get(0)
    np.testing.assert_allclose(evb, G["evals"][:NEV], rtol=1e-10, atol=1e-13)
    assert res.max() <= 1e-12
    check_vectors(Vb)
    b.close()
