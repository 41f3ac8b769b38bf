"""GPU parity of the plain FP64 drop-in against the oracle, through the C-ABI:
HYP links, the MatShell callback (T_DEG(B) x), D x, and the EPS-shaped
Krylov-Schur solve with recover_spectrum + fix_phase.

Tolerances: links 1e-13 (abs), T_DEG(B) x 1e-13 relative to max|y|,
eigenvalues 1e-10 relative, phase-fixed eigenvectors 1e-9 per component
where the gap to the neighbouring eigenvalues exceeds 1e-4 (cluster
subspaces compared by projector distance otherwise)."""
import numpy as np
import pytest

from oracle import lapheig_np as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lp():
    from laph_eigsys_amd import lapheig
    lapheig.lib()
    return lapheig


def seeded_timeslice(L, seed=1227, eps=0.4, t=0, T=1):
    cfg = N.synth_config(seed, eps, L, T)
    return cfg[t * 72 * L ** 3:(t + 1) * 72 * L ** 3]


@pytest.mark.parametrize("L", [4, 8, 12])
def test_hyp_parity(lp, gpu, L):
    ts = seeded_timeslice(L)
    U0 = O.map_timeslice(ts, L)
    ref = O.hyp(U0, L, 0.62, 0.62, 3)
    with lp.Tslice(L, 8, 12.0, 0.6, device=gpu, ts_config=ts) as t:
        np.testing.assert_array_equal(t.get_links(), U0)  # map_timeslice_to_eigen
        t.smearing_hyp(0.62, 0.62, 3)
        got = t.get_links()
    print(f"L={L} max |U_gpu - U_oracle| = {abs(got - ref).max():.3e}")
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-13)


@pytest.mark.parametrize("L,deg", [(4, 1), (6, 2), (8, 8), (16, 12)])
def test_shell_mult_parity(lp, gpu, L, deg):
    ts = seeded_timeslice(L, seed=11)
    U = O.hyp(O.map_timeslice(ts, L), L, 0.62, 0.62, 3)
    rng = np.random.default_rng(L * 100 + deg)
    x = rng.standard_normal(3 * L ** 3) + 1j * rng.standard_normal(3 * L ** 3)
    with lp.Tslice(L, deg, 12.0, 0.6, device=gpu) as t:
        t.set_links(U)
        y = t.MatMult_Laplacian2D(x)
        d = t.laplacian(x)
        diag = t.MatGetDiagonal_Laplacian2D()
    ref = O.tv_iter(U, L, 12.0, 0.6, deg, x)
    err = abs(y - ref).max() / abs(ref).max()
    print(f"L={L} deg={deg} rel err {err:.3e}")
    assert err < 1e-13
    np.testing.assert_allclose(d, O.apply_D(U, L, x), rtol=0, atol=1e-13 * abs(x).max())
    assert np.all(diag == -6.0)


def test_bad_params_rejected(lp, gpu):
    with pytest.raises(lp.LaphError):
        lp.Tslice(8, 0, 12.0, 0.6, device=gpu)  # DEG < 1
    with pytest.raises(lp.LaphError):
        lp.Tslice(8, 8, 0.5, 0.6, device=gpu)   # lambda_l <= lambda_c
    with pytest.raises(lp.LaphError):
        lp.Tslice(8, 8, 12.0, 0.6, device=99)


def compare_eigensystems(ev, V, ev_ref, V_ref, gap_tol=1e-4, comp_tol=1e-9):
    """Eigenvalues: relative 1e-10.  Eigenvectors: per component after phase
    fixing where isolated, projector distance inside clusters."""
    np.testing.assert_allclose(ev, ev_ref, rtol=1e-10, atol=1e-13)
    Vf, _ = O.fix_phase(V)
    Vf_ref, _ = O.fix_phase(V_ref)
    nev = len(ev)
    i = 0
    worst = 0.0
    while i < nev:
        j = i + 1
        while j < nev and abs(ev_ref[j] - ev_ref[j - 1]) < gap_tol:
            j += 1
        if j - i == 1:
            d = abs(Vf[i] - Vf_ref[i]).max()
            worst = max(worst, d)
            assert d < comp_tol, (i, d)
        else:
            A, B = V[i:j], V_ref[i:j]
            P = A.T @ A.conj() - B.T @ B.conj()
            assert np.linalg.norm(P) < 1e-7, (i, j)
        i = j
    return worst


def test_eps_solve_parity_c1(lp, gpu):
    """Configuration C1: 8^3, eps=0.4, HYP 0.62/0.62/3, nev=20, DEG=8, tol 1e-12."""
    L, nev, deg = 8, 20, 8
    ts = seeded_timeslice(L)
    U = O.hyp(O.map_timeslice(ts, L), L, 0.62, 0.62, 3)
    D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
    w = np.sort(np.linalg.eigvalsh(D))[::-1]
    lamC = -w[nev + 4] * 1.05
    with lp.Tslice(L, deg, 12.0, lamC, device=gpu, ts_config=ts) as t:
        t.smearing_hyp(0.62, 0.62, 3)
        theta, V, st = t.eps_solve(nev, tol=1e-12, seed=1227)
    nconv, th_ref, V_ref, st_ref = O.eps_solve(U, L, 12.0, lamC, deg, nev, tol=1e-12, seed=1227)
    ev = lp.recover_spectrum(theta, deg, 12.0, lamC)
    ev_ref = O.recover_spectrum(th_ref, deg, 12.0, lamC)
    print(f"gpu restarts {st.restarts} matvecs {st.matvecs}; oracle {st_ref}")
    np.testing.assert_allclose(ev, w[:nev], rtol=1e-10, atol=1e-13)
    worst = compare_eigensystems(ev, V, ev_ref, V_ref)
    print(f"max phase-fixed component diff (isolated pairs) {worst:.3e}")
    res = [np.linalg.norm(D @ V[i] - ev[i] * V[i]) for i in range(nev)]
    print("max residual |D v - lambda v|", max(res))
    assert max(res) < 1e-9
    ok, tr, _ = lp.check_trace(V)
    assert ok and abs(tr - nev) < 1e-10
