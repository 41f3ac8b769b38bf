"""Configuration C3 -- the headline workload, 32^3 x 64, eps 0.3, seed 1227,
HYP 0.62/0.62/3, N_ev 220, DEG 16 -- against the committed fixture
tests/golden/c3_32x32x32.npz (scipy ARPACK shift-invert on the sparse
Laplacian of the numpy restatement, timeslice 0 cross-checked with the C
oracle's Krylov-Schur; written by tests/golden/make_c3.py), timeslices 0
and 63.

The batched solver runs exactly as bench.py times it: all 64 timeslices in
one batch, search space 256, the warm-up degree and the filter layout
picked on the first solve, then per-cycle degrees up to 128 (72 here, where
the amplification bound stops them).  It is solved twice; both solves must
match the fixture and each other bit for bit.  The plain FP64 drop-in (the
MatShell path) solves the same two timeslices with the lambda_c rule
bench.py uses, and every one of the 64 timeslices for the eigenvalues.

Bar: eigenvalues 1e-10 relative (ev_ts.cpp:263-303 with recover_spectrum,
recover_spec.cpp:45-54, for the plain path); phase-fixed leading components
1e-9 and overlaps |g^H v|^2 1e-8 relative where the eigenvalue gap to both
neighbours exceeds 1e-4; per-pair residuals <= 1e-10; tr(V^H V) = N_ev."""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "c3_32x32x32.npz")
G = np.load(PATH)
L, T, NEV, DEG, LAM_L = int(G["L"]), int(G["T"]), int(G["nev"]), int(G["deg"]), float(G["lambda_l"])
TSL = [int(t) for t in G["tslices"]]
NSUB = 256


def test_fixture_consistent():
    for t in TSL:
        ev = G[f"evals_{t}"]
        assert ev.shape == (232,) and np.all(np.diff(ev) <= 0) and ev[0] < 0
        head = G[f"head_{t}"]
        assert head.shape == (NEV, 24) and np.all(np.abs(head[:, 0].imag) < 1e-15) and np.all(head[:, 0].real >= 0)
        assert G[f"residual_{t}"].max() < 1e-10
    # the C oracle's Krylov-Schur (EPS configuration) agreed when the fixture was written
    np.testing.assert_allclose(G["oracle_evals_0"], G["evals_0"][:NEV], rtol=1e-9)


def free_lambda_c(L, nev, factor=1.35):
    """lambda_c for the plain drop-in: factor x the nev-th level of the free
    Laplacian (levels 4 sum_i sin^2(pi k_i / L), each three-fold), an a-priori
    rule that needs no solve of the field."""
    s2 = 4 * np.sin(np.pi * np.arange(L) / L) ** 2
    free = np.sort(np.repeat((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel(), 3))
    return factor * free[nev - 1]


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
    assert len(iso) > NEV // 3, len(iso)
    for i in iso:
        assert np.abs(head[i] - G[f"head_{t}"][i]).max() < 1e-9, (t, i)
        assert abs(ov[i] - G[f"overlap_{t}"][i]) <= 1e-8 * max(G[f"overlap_{t}"][i], 1.0), (t, i)


@pytest.mark.gpu
def test_c3_batched_at_bench_settings_and_plain(gpu):
    from laph_eigsys_amd import lapheig as lp
    cfg = lp.synth_config(int(G["seed"]), float(G["eps"]), L, T)
    vol = 72 * L ** 3
    b = lp.Batch(L, DEG, LAM_L, 0.0, T, NEV, nsub=NSUB, tol=1e-10, seed=int(G["seed"]), device=gpu)
    try:
        for t in range(T):
            b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
        got = []
        for rep in range(2):
            st = b.solve()
            print(f"C3 solve {rep}: {st.seconds:.2f}s cycles {st.iterations} block {st.filter_block} strip "
                  f"{st.filter_strip} zrun {st.filter_zrun} max res {st.max_residual:.2e} rr fallbacks "
                  f"{st.rr_cholqr_fallbacks} shifted {st.shifted_cholqr} recoveries {st.recoveries}", flush=True)
            assert st.converged == T and st.max_residual <= 1e-10 and st.recoveries == 0
            got.append({t: b.get(t) for t in TSL})
            if rep == 0:
                ev_all = [b.get(t, want_vectors=False)[0] for t in range(T)]
        for t in TSL:
            ev, V, ph, res = got[0][t]
            check(ev, V, t)
            assert res.max() <= 1e-10
            assert lp.check_trace(V)[0]
            assert np.all(np.abs(V[:, 0].imag) < 1e-15) and np.all(V[:, 0].real >= 0)  # phase-fixed in the solve
            np.testing.assert_allclose(np.angle(V[:, 0] * np.exp(1j * ph)), ph, rtol=0, atol=1e-12)
            for a, c in zip(got[0][t], got[1][t]):
                np.testing.assert_array_equal(a, c)
    finally:
        b.close()
    lam_c = free_lambda_c(L, NEV)
    for t in TSL:
        with lp.Tslice(L, DEG, LAM_L, lam_c, device=gpu, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            theta, V, pst = sl.eps_solve(NEV, tol=1e-10)
        print(f"C3 plain t={t}: restarts {pst.restarts} matvecs {pst.matvecs} {pst.seconds:.2f}s", flush=True)
        ev = lp.recover_spectrum(theta, DEG, LAM_L, lam_c)
        check(ev, V, t)
    # every timeslice of the headline batch against the plain drop-in
    worst = 0.0
    for t in range(T):
        with lp.Tslice(L, DEG, LAM_L, lam_c, device=gpu, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            theta, _, _ = sl.eps_solve(NEV, tol=1e-10, want_vectors=False)
        ev = lp.recover_spectrum(theta, DEG, LAM_L, lam_c)
        np.testing.assert_allclose(ev_all[t], ev, rtol=1e-10, atol=0, err_msg=str(t))
        worst = max(worst, float(np.max(np.abs(ev_all[t] - ev) / np.abs(ev))))
    print(f"C3: largest relative eigenvalue difference batched vs plain over {T} timeslices {worst:.2e}", flush=True)
