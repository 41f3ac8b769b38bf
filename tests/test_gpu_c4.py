"""Configuration C4 at its size: 48^3 field (eps 0.3, seed 1227, HYP
0.62/0.62/3), N_ev 200, DEG 16, residual 1e-10 -- timeslices 0 and 1 (the
synthetic field's timeslices do not depend on its extent, so these are
timeslices 0 and 1 of the 48^3 x 96 configuration).

The batched solver (the settings bench.py --workload c4 uses: search space
256, the filter layout autotuned on the first solve) against the plain FP64
drop-in (Krylov-Schur on T_16(B) through the MatShell operator, lambda_c =
1.35 x the 200th free-field level): two independent algorithms, each pinned
against dense and ARPACK spectra at smaller sizes.  Bar: eigenvalues 1e-10
relative, phase-fixed leading components 1e-9 where the eigenvalue gap to
both neighbours exceeds 1e-4, per-pair residuals <= 1e-10, tr(V^H V) = N_ev."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L, NEV, DEG, LAM_L, TOL = 48, 200, 16, 12.0, 1e-10


def free_lambda_c(L, nev, factor=1.35):
    s2 = 4 * np.sin(np.pi * np.arange(L) / L) ** 2
    free = np.sort(np.repeat((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel(), 3))
    return factor * free[nev - 1]


def test_c4_batched_against_plain(gpu):
    from laph_eigsys_amd import lapheig as lp
    T = 2
    cfg = lp.synth_config(1227, 0.3, L, T)
    vol = 72 * L ** 3
    b = lp.Batch(L, DEG, LAM_L, 0.0, T, NEV, tol=TOL, seed=1227, device=gpu)
    try:
        for t in range(T):
            b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
        st = b.solve()
        print(f"C4 batched: {st.seconds:.2f}s cycles {st.iterations} block {st.filter_block} strip {st.filter_strip} "
              f"max res {st.max_residual:.2e} rr fallbacks {st.rr_cholqr_fallbacks} shifted {st.shifted_cholqr}",
              flush=True)
        assert st.converged == T and st.max_residual <= TOL and st.recoveries == 0
        got = [b.get(t) for t in range(T)]
    finally:
        b.close()
    lam_c = free_lambda_c(L, NEV)
    for t in range(T):
        with lp.Tslice(L, DEG, LAM_L, lam_c, device=gpu, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            theta, Vp, pst = sl.eps_solve(NEV, tol=TOL)
        evp = lp.recover_spectrum(theta, DEG, LAM_L, lam_c)
        ev, V, ph, res = got[t]
        dev = np.max(np.abs(ev - evp) / np.abs(evp))
        print(f"C4 t={t}: plain {pst.seconds:.2f}s restarts {pst.restarts}; max rel eigenvalue difference {dev:.1e}",
              flush=True)
        np.testing.assert_allclose(ev, evp, rtol=1e-10, atol=0)
        assert res.max() <= TOL and lp.check_trace(V)[0]
        Vpf, _ = lp.fix_phase(Vp)
        d = np.abs(np.diff(evp))
        iso = [i for i in range(1, NEV - 1) if d[i] > 1e-4 and d[i - 1] > 1e-4]
        assert len(iso) > NEV // 4, len(iso)
        for i in iso:
            assert np.abs(V[i, :24] - Vpf[i, :24]).max() < 1e-9, (t, i)
