"""Configuration C2 (24^3 x 48, eps 0.4, seed 1227, HYP 0.62/0.62/3,
N_ev 120, DEG 12) against the committed fixture tests/golden/c2_24x24x24.npz
(scipy ARPACK shift-invert on the sparse Laplacian of the numpy restatement,
cross-checked with the C oracle's Krylov-Schur; written by
tests/golden/make_c2.py), timeslices 0 and 47.

Bar: eigenvalues 1e-10 relative for both HIP paths (the plain FP64 drop-in
through the shell operator, the mixed batched solver); phase-fixed leading
components 1e-9 and overlaps |g^H v|^2 1e-8 relative where the eigenvalue
gap to both neighbours exceeds 1e-4 (|g|^2 = 2n, so the overlap bar is
1e-8 max(|g^H v|^2, 1))."""
import os

import numpy as np
import pytest

G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_24x24x24.npz"))
L, T, NEV, DEG, LAM_L = int(G["L"]), int(G["T"]), int(G["nev"]), int(G["deg"]), float(G["lambda_l"])
TSL = [int(t) for t in G["tslices"]]


def test_fixture_consistent():
    for t in TSL:
        ev = G[f"evals_{t}"]
        assert ev.shape == (128,) and np.all(np.diff(ev) <= 0) and ev[0] < 0
        head = G[f"head_{t}"]
        assert head.shape == (NEV, 24) and np.all(np.abs(head[:, 0].imag) < 1e-15) and np.all(head[:, 0].real >= 0)


def isolated(ev, gap=1e-4):
    d = np.abs(np.diff(ev[:NEV + 1]))
    return [i for i in range(NEV) if d[i] > gap and (i == 0 or d[i - 1] > gap)]


def check(ev, V, t):
    ref = G[f"evals_{t}"][:NEV]
    np.testing.assert_allclose(ev, ref, rtol=1e-10, atol=0)
    if V is None:
        return
    rng = np.random.default_rng(int(G["g_seed"]))
    g = rng.normal(size=3 * L ** 3) + 1j * rng.normal(size=3 * L ** 3)
    ov = np.abs(V.conj() @ g) ** 2
    ph = np.angle(V[:, 0])
    head = (V * np.exp(-1j * ph)[:, None])[:, :24]
    iso = isolated(G[f"evals_{t}"])
    assert len(iso) > NEV // 2
    for i in iso:
        assert np.abs(head[i] - G[f"head_{t}"][i]).max() < 1e-9, (t, i)
        assert abs(ov[i] - G[f"overlap_{t}"][i]) <= 1e-8 * max(G[f"overlap_{t}"][i], 1.0), (t, i)


@pytest.mark.gpu
def test_c2_plain_and_batched(gpu):
    from laph_eigsys_amd import lapheig as lp
    cfg = lp.synth_config(int(G["seed"]), float(G["eps"]), L, T)
    vol = 72 * L ** 3
    lam_c = 1.05 * max(-G[f"evals_{t}"][NEV + 4] for t in TSL)
    for t in TSL:
        with lp.Tslice(L, DEG, LAM_L, lam_c, device=gpu, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            theta, V, st = sl.eps_solve(NEV, tol=1e-10)
        check(lp.recover_spectrum(theta, DEG, LAM_L, lam_c), V, t)
    b = lp.Batch(L, DEG, LAM_L, 0.0, len(TSL), NEV, tol=1e-10, seed=5, device=gpu)
    for k, t in enumerate(TSL):
        b.set_timeslice(k, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
    st = b.solve()
    assert st.converged == len(TSL) and st.max_residual <= 1e-10
    for k, t in enumerate(TSL):
        ev, V, ph, res = b.get(k)
        check(ev, V, t)
        assert lp.check_trace(V)[0]
    b.close()


@pytest.mark.gpu
def test_c2_every_timeslice_batched_equals_plain(gpu):
    """All 48 timeslices of C2 in one batch (the configuration's whole
    lattice, as bench.py --workload c2 runs it) against the plain FP64
    drop-in timeslice by timeslice: eigenvalues 1e-10 relative, every pair's
    residual within the bound, and the fixture's timeslices 0 and 47 as
    above."""
    from laph_eigsys_amd import lapheig as lp
    cfg = lp.synth_config(int(G["seed"]), float(G["eps"]), L, T)
    vol = 72 * L ** 3
    b = lp.Batch(L, DEG, LAM_L, 0.0, T, NEV, tol=1e-10, seed=5, device=gpu)
    for t in range(T):
        b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
    st = b.solve()
    assert st.converged == T and st.max_residual <= 1e-10
    evb = [b.get(t, want_vectors=False) for t in range(T)]
    b.close()
    print(f"C2 batch of {T}: cycles {st.iterations}, warm-up degree {st.warmup_degree}, max res {st.max_residual:.2e}")
    lam_c = 1.35 * max(-ev[-1] for ev, _, _, _ in evb)
    worst = 0.0
    for t in range(T):
        ev, _, _, res = evb[t]
        assert res.max() <= 1e-10
        with lp.Tslice(L, DEG, LAM_L, lam_c, device=gpu, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            theta, _, _ = sl.eps_solve(NEV, tol=1e-10, want_vectors=False)
        evp = lp.recover_spectrum(theta, DEG, LAM_L, lam_c)
        np.testing.assert_allclose(ev, evp, rtol=1e-10, atol=0, err_msg=str(t))
        worst = max(worst, float(np.max(np.abs(ev - evp) / np.abs(evp))))
        if t in TSL:
            check(ev, None, t)
    print(f"largest relative eigenvalue difference batched vs plain over {T} timeslices: {worst:.2e}")
