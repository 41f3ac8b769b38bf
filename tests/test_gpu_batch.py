"""GPU parity of the mixed-precision batched solver (FP32 correction-form
Chebyshev filter, FP64 CholQR2 / Rayleigh-Ritz / locking) against the oracle:
dense diagonalisation of the assembled D per timeslice, the oracle's HYP,
and the plain FP64 path.

Bar: eigenvalues within 1e-10 relative of the FP64 oracle, per-pair
residual |D v - lambda v| <= tol, tr(V^H V) = nev, phase-fixed eigenvectors
within 1e-9 per component where the eigenvalue gap exceeds 1e-4."""
import os

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


def smeared(L, seed, eps, T):
    cfg = N.synth_config(seed, eps, L, T)
    out = []
    for t in range(T):
        ts = cfg[t * 72 * L ** 3:(t + 1) * 72 * L ** 3]
        out.append((ts, O.hyp(O.map_timeslice(ts, L), L, 0.62, 0.62, 3)))
    return out


def check_pairs(ev, V, res, D, nev, tol):
    w = np.sort(np.linalg.eigvalsh(D))[::-1][:nev]
    np.testing.assert_allclose(ev, w, rtol=1e-10, atol=1e-12)
    assert res.max() <= tol
    R = [np.linalg.norm(D @ V[i] - ev[i] * V[i]) for i in range(nev)]
    assert max(R) <= 1.5 * tol, max(R)
    G = V.conj() @ V.T
    np.testing.assert_allclose(G, np.eye(nev), atol=1e-12)
    assert np.all(np.abs(V[:, 0].imag) < 1e-14) and np.all(V[:, 0].real >= 0)
    return w


@pytest.mark.parametrize("L,nts,nev,deg", [(6, 3, 20, 8), (8, 2, 20, 12)])
def test_batch_against_dense(lp, gpu, L, nts, nev, deg):
    tol = 1e-10
    data = smeared(L, 1227, 0.4, nts)
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=5, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    print(f"L={L} nts={nts}: cycles {st.iterations} fp32 vecs {st.stencil_vecs} max res {st.max_residual:.2e}")
    assert st.converged == nts
    # first-cycle window: 3 x the K-th (K = 64) level of the free Laplacian,
    # whose levels are 4 sum_i sin^2(pi k_i / L), each three-fold; at most 9
    s2 = 4 * np.sin(np.pi * np.arange(L) / L) ** 2
    free = np.sort(np.repeat((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel(), 3))
    assert abs(st.start_window - min(3 * free[63], 9.0)) < 1e-12
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
        check_pairs(ev, V, res, D, nev, tol)
    b.close()


def test_batch_set_links_matches_set_timeslice(lp, gpu):
    L, nev = 6, 12
    (ts, U), = smeared(L, 77, 0.3, 1)
    b1 = lp.Batch(L, 8, 12.0, 0.0, 1, nev, nsub=64, tol=1e-11, seed=3, device=gpu)
    b1.set_links(0, U)
    b1.solve()
    ev1, V1, _, _ = b1.get(0)
    b2 = lp.Batch(L, 8, 12.0, 0.0, 1, nev, nsub=64, tol=1e-11, seed=3, device=gpu)
    b2.set_timeslice(0, ts, 0.62, 0.62, 3)
    b2.solve()
    ev2, V2, _, _ = b2.get(0)
    # the eigenvectors into a caller's buffer, twice (the batch's staging
    # buffer is reused): the same values, in the caller's array
    out = np.full_like(V2, np.nan)
    for _ in range(2):
        ev3, V3, _, _ = b2.get(0, out=out)
        assert V3 is out
        np.testing.assert_array_equal(V3, V2)
        np.testing.assert_array_equal(ev3, ev2)
    np.testing.assert_allclose(ev1, ev2, rtol=1e-12)


def test_batch_matches_plain_fp64(lp, gpu):
    """Mixed batched vs plain FP64 Krylov-Schur on one 12^3 timeslice."""
    L, nev, deg = 12, 40, 8
    (ts, U), = smeared(L, 4242, 0.35, 1)
    b = lp.Batch(L, deg, 12.0, 0.0, 1, nev, nsub=64, tol=1e-10, seed=9, device=gpu)
    b.set_timeslice(0, ts, 0.62, 0.62, 3)
    b.solve()
    ev, V, ph, res = b.get(0)
    lamC = -ev[-1] * 1.3
    with lp.Tslice(L, deg, 12.0, lamC, device=gpu, ts_config=ts) as t:
        t.smearing_hyp(0.62, 0.62, 3)
        theta, Vp, _ = t.eps_solve(nev, tol=1e-12)
    evp = lp.recover_spectrum(theta, deg, 12.0, lamC)
    np.testing.assert_allclose(ev, evp, rtol=1e-10, atol=1e-12)
    Vpf, php = lp.fix_phase(Vp)
    gaps = np.abs(np.diff(ev))
    for i in range(1, nev - 1):
        if gaps[i - 1] > 1e-4 and gaps[i] > 1e-4:
            assert np.abs(V[i] - Vpf[i]).max() < 1e-9, i


@pytest.mark.parametrize("strip", [8, 16])
def test_batch_matches_plain_fp64_l16(lp, gpu, strip, monkeypatch):
    """Mixed batched vs plain FP64 Krylov-Schur on two 16^3 timeslices with
    the y-row work-group stencils over y-strips of 8 (two groups per strip
    and x-plane) and over whole x-planes."""
    L, nts, nev, deg = 16, 2, 60, 8
    monkeypatch.setenv("LAPH_STRIP", str(strip))
    data = smeared(L, 5151, 0.35, nts)
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=128, tol=1e-10, seed=2, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.converged == nts and st.filter_strip == strip
    for t, (ts, _) in enumerate(data):
        ev, V, ph, res = b.get(t)
        lamC = -ev[-1] * 1.3
        with lp.Tslice(L, deg, 12.0, lamC, device=gpu, ts_config=ts) as tsl:
            tsl.smearing_hyp(0.62, 0.62, 3)
            theta, Vp, _ = tsl.eps_solve(nev, tol=1e-12)
        evp = lp.recover_spectrum(theta, deg, 12.0, lamC)
        np.testing.assert_allclose(ev, evp, rtol=1e-10, atol=1e-12)
    b.close()


@pytest.mark.parametrize("window0", [None, "0.3"])
def test_rough_field_start_window(lp, gpu, window0, monkeypatch):
    """An unsmeared, strongly disordered field (eps 1.2, no HYP; lambda_K =
    2.08 for K = 64 at L = 8, the free Laplacian's 1.76) from the default
    first-cycle window (three times the free K-th level) and from a window
    far below the field's lambda_K (LAPH_WINDOW0 = 0.3), where the first
    filter damps wanted directions: the next windows come from the Ritz
    values and both give the dense oracle's spectrum."""
    L, nts, nev, deg, tol = 8, 2, 24, 8, 1e-10
    if window0:
        monkeypatch.setenv("LAPH_WINDOW0", window0)
    cfg = N.synth_config(77, 1.2, L, nts)
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=9, device=gpu)
    data = []
    for t in range(nts):
        ts = cfg[t * 72 * L ** 3:(t + 1) * 72 * L ** 3]
        b.set_timeslice(t, ts, 0.62, 0.62, 0)
        data.append(O.map_timeslice(ts, L))
    st = b.solve()
    print(f"window0 {window0}: cycles {st.iterations} max res {st.max_residual:.2e}")
    assert st.converged == nts
    for t, U in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


@pytest.mark.parametrize("storage,width", [(0, 32), (0, 64), (0, 128), (1, 128), (1, 256)])
def test_filter_block_widths(lp, gpu, storage, width, monkeypatch):
    """Every layout of the filter buffers (FP32 storage: two sites per
    wavefront in 32-vector blocks, one or two vectors per lane in 64- /
    128-vector blocks; BF16 storage: two or four per lane in 128- /
    256-vector blocks) gives the dense oracle's spectrum, and the timing
    autotuner picks one of the widths of the storage format."""
    L, nts, nev, deg, tol = 8, 2, 60, 10, 1e-10
    data = smeared(L, 99, 0.35, nts)
    monkeypatch.setenv("LAPH_BW", str(width))
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=256, tol=tol, seed=5, device=gpu, filter_storage=storage)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.filter_block == width and st.converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()
    monkeypatch.delenv("LAPH_BW")
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=256, tol=tol, seed=5, device=gpu, filter_storage=storage)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    assert b.solve().filter_block in ((32, 64, 128) if storage == 0 else (128, 256))
    b.close()


@pytest.mark.parametrize("L,zrun", [(9, 0), (12, 4), (12, 0)])
def test_filter_block_32_shapes(lp, gpu, L, zrun, monkeypatch):
    """32-vector filter blocks on a lattice that is not a multiple of 4 (one
    site per wavefront in the FP64 residual kernel) and with and without the
    y-row work-groups of the FP64 stencils (a 64-column group of the residual
    spans two 32-vector blocks), against the dense oracle."""
    tol, nev = 1e-10, 40
    data = smeared(L, 17, 0.35, 2)
    monkeypatch.setenv("LAPH_BW", "32")
    monkeypatch.setenv("LAPH_ZRUN", str(zrun))
    b = lp.Batch(L, 10, 12.0, 0.0, 2, nev, nsub=96 if L == 9 else 128, tol=tol, seed=5, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.filter_block == 32 and st.converged == 2
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


@pytest.mark.parametrize("width", [64, 128])
@pytest.mark.parametrize("zrun", [0, 1, 2, 4, 8])
def test_filter_z_runs(lp, gpu, zrun, width, monkeypatch):
    """The FP32 64- and 128-vector filter steps in work-groups of 4 y-rows,
    each wavefront walking zrun sites of its z-line (z-neighbour rows shifted
    through registers, y-neighbours of the inner rows through LDS; at L = 8
    the runs and the groups meet the periodic edges in z and y), and the
    one-site-per-wavefront steps (zrun 0), give the dense oracle's spectrum in
    both strip orders."""
    L, nts, nev, deg, tol = 8, 2, 40, 8, 1e-10
    data = smeared(L, 57, 0.4, nts)
    monkeypatch.setenv("LAPH_BW", str(width))
    monkeypatch.setenv("LAPH_ZRUN", str(zrun))
    for strip in (4, 8):
        monkeypatch.setenv("LAPH_STRIP", str(strip))
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=128, tol=tol, seed=4, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        assert st.converged == nts and st.filter_zrun == zrun and st.filter_strip == strip
        for t, (_, U) in enumerate(data):
            ev, V, ph, res = b.get(t)
            check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
        b.close()


def test_filter_z_run_rejected(lp, gpu, monkeypatch):
    """A z-run on a lattice whose L is not a multiple of 4 is refused at creation."""
    monkeypatch.setenv("LAPH_ZRUN", "4")
    with pytest.raises(lp.LaphError):
        lp.Batch(6, 8, 12.0, 0.0, 1, 10, nsub=64, tol=1e-10, seed=1, device=gpu)


@pytest.mark.parametrize("form,passes", [(2, 1), (2, 2), (2, 3), (2, 4), (1, 1), (1, 4), (0, 1), (0, 4)])
def test_filter_passes(lp, gpu, form, passes, monkeypatch):
    """T_m(B)^P per cycle (correction pass + P-1 plain passes, each adding
    gamma^(p-1) f_m in its last step) gives the dense oracle's spectrum: in
    the product form over the roots of T_m with the correction pass as a
    product over the complex roots of (T_m(y) - T_m(theta)) / (y - theta)
    applied to r (form 2, the default), with the correction pass reading r in
    every step (form 1), and in the three-term recurrence (form 0)."""
    L, nts, nev, deg, tol = 8, 2, 40, 8, 1e-10
    data = smeared(L, 31, 0.4, nts)
    monkeypatch.setenv("LAPH_FILTER_FORM", str(form))
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=2, device=gpu, filter_passes=passes)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


@pytest.mark.parametrize("fdeg,wdeg,safety", [(24, 0, None), (72, 0, None), (0, 0, None), (72, 24, None),
                                               (72, 0, "0"), (72, 0, "3")])
def test_filter_degree_per_cycle(lp, gpu, fdeg, wdeg, safety, monkeypatch):
    """Automatic passes (filter_passes = 0): after the warm-up cycles one
    filter per cycle of the highest degree (a multiple of 4 up to
    LAPH_FILTER_DEGREE) whose amplification stays under the cap, lowered to
    what the unconverged pairs need (margin LAPH_DEGREE_SAFETY: 30, 0 = never
    lowered, 3 = a thin margin that may cost cycles but not accuracy); 0
    selects T_DEG(B) applied up to 5 times per cycle; LAPH_WARM_DEG sets the
    degree of the warm-up cycle.  All give the dense oracle's spectrum."""
    L, nts, nev, deg, tol = 8, 2, 40, 8, 1e-10
    data = smeared(L, 313, 0.4, nts)
    monkeypatch.setenv("LAPH_FILTER_DEGREE", str(fdeg))
    if wdeg:
        monkeypatch.setenv("LAPH_WARM_DEG", str(wdeg))
    if safety is not None:
        monkeypatch.setenv("LAPH_DEGREE_SAFETY", safety)
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=7, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    assert b.solve().converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


def test_shifted_cholqr_path(lp, gpu, monkeypatch):
    """Every Cholesky QR pass preceded by the shifted pass (the fallback for
    numerically singular Grams): same spectrum as the dense oracle."""
    L, nts, nev, deg, tol = 8, 2, 40, 8, 1e-10
    data = smeared(L, 57, 0.4, nts)
    monkeypatch.setenv("LAPH_SHIFTED_CHOLQR", "1")
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=4, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    assert b.solve().converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


@pytest.mark.parametrize("L,nts,nev,nsub,lanes", [(5, 3, 16, 64, 2), (7, 1, 30, 64, 1), (9, 2, 60, 64, 2),
                                                   (6, 4, 20, 64, 3), (12, 1, 40, 128, 1), (8, 3, 40, 128, 2)])
def test_batch_edge_shapes(lp, gpu, L, nts, nev, nsub, lanes, monkeypatch):
    """Odd L and L not divisible by 4 (plain sweep order, one site per
    wavefront, partial last site-block), L = 12 (y-row work-groups walking
    z-lines in runs of 4, three strips of 4 or x-major planes), a single
    timeslice, timeslices split over two or three lanes (host threads and
    streams, LAPH_LANES; at L = 8 with the y-row work-group kernels on each
    lane's share), and nev close to the search-space width."""
    tol = 1e-10
    data = smeared(L, 4000 + L, 0.4, nts)
    monkeypatch.setenv("LAPH_LANES", str(lanes))
    b = lp.Batch(L, 10, 12.0, 0.0, nts, nev, nsub=nsub, tol=tol, seed=3, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.converged == nts and st.filter_strip in ((L,) if L % 4 else (L, 4, 8))
    assert st.filter_zrun == (0 if L % 4 or st.filter_block != 128 else (8 if L % 8 == 0 else 4))
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


@pytest.mark.parametrize("start_rr,trsm", [(0, 0), (1, 0), (0, 1)])
def test_start_block(lp, gpu, start_rr, trsm, monkeypatch):
    """The first filter cycle straight from the Gaussian start block (every
    column at the window edge, theta = 1) and after a Rayleigh-Ritz step on it
    (LAPH_START_RR=1) span the same subspace: both give the dense oracle's
    spectrum; so does Rayleigh-Ritz through triangular solves (LAPH_TRSM=1)
    instead of the explicit inverse of the Gram's Cholesky factor."""
    L, nts, nev, deg, tol = 8, 2, 40, 10, 1e-10
    data = smeared(L, 77, 0.4, nts)
    monkeypatch.setenv("LAPH_START_RR", str(start_rr))
    monkeypatch.setenv("LAPH_TRSM", str(trsm))
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=9, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    assert b.solve().converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


def test_recovery_from_a_non_finite_block(lp, gpu, monkeypatch):
    """A timeslice whose unlocked columns turn non-finite (NaN written into one
    row by the LAPH_TEST_POISON control) makes every Cholesky fail; its active
    columns are re-seeded, the batch still converges, and every pair matches
    the dense oracle."""
    L, nts, nev, tol = 6, 2, 20, 1e-10
    data = smeared(L, 606, 0.4, nts)
    monkeypatch.setenv("LAPH_TEST_POISON", "1")
    b = lp.Batch(L, 8, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=6, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    assert st.recoveries >= 1 and st.converged == nts
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


def test_batch_solved_twice(lp, gpu):
    """One batch solved twice in a row (the second solve reuses the captured
    filter graphs, the layout chosen by the first and every work buffer):
    both spectra match dense diagonalisation, and the second solve repeats
    the first bit for bit -- the column-norm reductions add fixed-order
    partials, so nothing depends on the order atomics would land in."""
    L, nts, nev, tol = 8, 3, 24, 1e-10
    data = smeared(L, 31, 0.4, nts)
    b = lp.Batch(L, 8, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=5, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    runs = []
    for rep in range(2):
        st = b.solve()
        print(f"solve {rep}: cycles {st.iterations} recoveries {st.recoveries} rr fallbacks "
              f"{st.rr_cholqr_fallbacks} shifted {st.shifted_cholqr}")
        assert st.converged == nts and st.recoveries == 0
        runs.append([b.get(t) for t in range(nts)])
    for t, (_, U) in enumerate(data):
        D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
        for r in runs:
            ev, V, ph, res = r[t]
            check_pairs(ev, V, res, D, nev, tol)
        for a, c in zip(runs[0][t], runs[1][t]):
            np.testing.assert_array_equal(a, c)
    b.close()


@pytest.mark.parametrize("split,mfma", [("1", "0"), ("2", "0"), ("1,4", "0"), ("2,2", "0"), ("2,4", "0"),
                                        ("1", "1"), ("1,2", "1"), ("1,4", "1"), ("1,8", "1")])
def test_gram_row_blocks_and_site_chunks(lp, gpu, split, mfma, monkeypatch):
    """Every Gram layout -- rocBLAS zgemm on the full product or its upper
    triangle in two row blocks, or the hand-written FP64 MFMA kernel on the
    64 x 64 upper tiles; the site sum in 1, 2 or 4 chunks added in order --
    gives the dense oracle's spectrum (the Rayleigh quotient's lower triangle
    is filled from its upper one where full products need it)."""
    L, nts, nev, tol = 8, 3, 40, 1e-10
    data = smeared(L, 23, 0.35, nts)
    monkeypatch.setenv("LAPH_GRAM_SPLIT", split)
    monkeypatch.setenv("LAPH_GRAM_MFMA", mfma)
    b = lp.Batch(L, 10, 12.0, 0.0, nts, nev, nsub=128, tol=tol, seed=5, device=gpu)
    for t, (ts, _) in enumerate(data):
        b.set_timeslice(t, ts, 0.62, 0.62, 3)
    st = b.solve()
    want = [int(x) for x in split.split(",")] + [1]
    assert st.converged == nts and (st.gram_blocks, st.gram_chunks, st.gram_mfma) == (want[0], want[1], int(mfma))
    for t, (_, U) in enumerate(data):
        ev, V, ph, res = b.get(t)
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
    b.close()


def test_layout_cache(lp, gpu, tmp_path, monkeypatch):
    """LAPH_TUNE_FILE: the first solve times the layouts and appends them
    under the shape's key; a later batch of the same shape reads them instead
    of timing (layout_from_cache), skipping lines whose values are not a
    valid layout, and solves to the same eigenvalues bit for bit; a batch of
    another shape misses and appends its own line."""
    L, nts, nev, deg, tol = 8, 2, 40, 8, 1e-10
    data = smeared(L, 23, 0.4, nts)
    cache = tmp_path / "layouts.txt"
    monkeypatch.setenv("LAPH_TUNE_FILE", str(cache))

    def run(nsub=128):
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=nsub, tol=tol, seed=3, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        assert st.converged == nts
        ev = [b.get(t, want_vectors=False)[0] for t in range(nts)]
        b.close()
        return st, ev

    st1, ev1 = run()
    assert st1.layout_from_cache == 0
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and " : bw " in lines[0]
    key, val = lines[0].split(" : ")
    # an invalid entry for the same key (a width that does not divide K) comes first and is skipped
    cache.write_text(f"{key} : bw 48 strip 3 split 1 chunks 1 mfma 0 warm {deg}\n{lines[0]}\n")
    st2, ev2 = run()
    assert st2.layout_from_cache == 1
    for f in ("filter_block", "filter_strip", "gram_blocks", "gram_chunks", "gram_mfma", "warmup_degree"):
        assert getattr(st2, f) == getattr(st1, f), f
    for a, b in zip(ev1, ev2):
        assert np.array_equal(a, b)
    st3, _ = run(nsub=192)  # another search space: a miss, timed and appended
    assert st3.layout_from_cache == 0
    assert len(cache.read_text().splitlines()) == 3


def test_rayleigh_ritz_start_keeps_the_start_window(lp, gpu, monkeypatch):
    """A warm-up cycle that follows a Rayleigh-Ritz step on the start block
    (LAPH_START_RR=1; a network proposal takes the same path) keeps its
    window's lower end at the start window, as the start straight from the
    block does: both take the same cycles (and within 5% the same stencil
    applications) to the plain FP64 drop-in's spectrum.  With the window taken
    from the unfiltered block's own Ritz values instead
    (LAPH_WARMUP_WINDOW_CAP=0) the warm-up filter damps only the top of the
    spectrum and the solve needs more work.  The warm-up degree is held at
    DEG in all three (the first solve's timed choice could differ between
    them)."""
    L, nts, nev, deg, tol = 16, 2, 60, 12, 1e-10
    data = smeared(L, 31, 0.4, nts)
    monkeypatch.setenv("LAPH_WARM_DEG", str(deg))
    runs = {}
    for name, env in (("block", {"LAPH_START_RR": "0"}), ("rr", {"LAPH_START_RR": "1"}),
                      ("rr_uncapped", {"LAPH_START_RR": "1", "LAPH_WARMUP_WINDOW_CAP": "0"})):
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=128, tol=tol, seed=4, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        assert st.converged == nts
        runs[name] = (st.iterations, st.stencil_vecs, [b.get(t, want_vectors=False)[0] for t in range(nts)])
        b.close()
        monkeypatch.delenv("LAPH_WARMUP_WINDOW_CAP", raising=False)
    print({k: v[:2] for k, v in runs.items()})  # (cycles, FP32 stencil applications)
    assert runs["rr"][0] == runs["block"][0] and abs(runs["rr"][1] - runs["block"][1]) <= 0.05 * runs["block"][1]
    assert runs["rr_uncapped"][:2] >= runs["rr"][:2]  # measured at 16^3: 4 cycles, 31744 against 30720
    for t, (ts, _) in enumerate(data):  # reference: the plain FP64 drop-in
        lamC = -runs["block"][2][t][-1] * 1.3
        with lp.Tslice(L, deg, 12.0, lamC, device=gpu, ts_config=ts) as tsl:
            tsl.smearing_hyp(0.62, 0.62, 3)
            theta, _, _ = tsl.eps_solve(nev, tol=1e-12, want_vectors=False)
        evp = lp.recover_spectrum(theta, deg, 12.0, lamC)
        for name in runs:
            np.testing.assert_allclose(runs[name][2][t], evp, rtol=1e-10, atol=1e-12, err_msg=name)


def test_warmup_degree_timed_on_first_solve(lp, gpu, monkeypatch):
    """The first solve from the start block solves the batch once per
    warm-up degree candidate (DEG, then 24, 32 and 40 above it within the
    amplification bound at the start window) and keeps the fastest; its pairs
    are those of the chosen degree, so a second solve of the batch repeats
    them bit for bit.  LAPH_WARM_DEG fixes the degree and skips the timing."""
    L, nts, nev, deg, tol = 12, 2, 40, 8, 1e-10
    data = smeared(L, 41, 0.4, nts)

    def batch():
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=7, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        return b

    b = batch()
    st1 = b.solve()
    ev1 = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    st2 = b.solve()
    ev2 = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    b.close()
    print(f"warm-up degree {st1.warmup_degree}: first solve {st1.seconds:.2f}s, second {st2.seconds:.2f}s")
    assert st1.converged == st2.converged == nts and st1.warmup_degree == st2.warmup_degree
    assert st1.warmup_degree in (deg, 24, 32, 40) or (deg < st1.warmup_degree < 40 and st1.warmup_degree % 4 == 0)
    for a, c in zip(ev1, ev2):
        assert np.array_equal(a, c)
    monkeypatch.setenv("LAPH_WARM_DEG", "24")
    b = batch()
    st3 = b.solve()
    ev3 = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    b.close()
    assert st3.warmup_degree == 24 and st3.converged == nts
    for a, c in zip(ev1, ev3):
        np.testing.assert_allclose(c, a, rtol=1e-10, atol=1e-12)
    # a candidate above DEG whose solve fails is dropped: the first solve
    # keeps another degree and returns its converged pairs
    monkeypatch.delenv("LAPH_WARM_DEG")
    monkeypatch.setenv("LAPH_TEST_FAIL_WARM", str(st1.warmup_degree if st1.warmup_degree != deg else 24))
    b = batch()
    st4 = b.solve()
    ev4 = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    b.close()
    print(f"failed candidate {int(os.environ['LAPH_TEST_FAIL_WARM'])}: kept {st4.warmup_degree}")
    assert st4.converged == nts and st4.warmup_degree != int(os.environ["LAPH_TEST_FAIL_WARM"])
    for a, c in zip(ev1, ev4):
        np.testing.assert_allclose(c, a, rtol=1e-10, atol=1e-12)


def test_block_degrees(lp, gpu, monkeypatch):
    """Per-block filter degrees (each 64-column block filtered only to the
    degree its unconverged wanted columns need, LAPH_BLOCK_DEGREE=1, the
    default) against one degree per cycle (LAPH_BLOCK_DEGREE=0): the same
    spectrum to 1e-10 as the plain FP64 drop-in, every residual <= tol, with
    no more FP32 stencil work."""
    L, nts, nev, deg, tol = 16, 3, 100, 8, 1e-10
    data = smeared(L, 97, 0.35, nts)
    monkeypatch.setenv("LAPH_WARM_DEG", str(deg))
    runs = {}
    for on in (0, 1):
        monkeypatch.setenv("LAPH_BLOCK_DEGREE", str(on))
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=128, tol=tol, seed=12, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        assert st.converged == nts and st.max_residual <= tol
        runs[on] = (st.iterations, st.stencil_vecs, [b.get(t, want_vectors=False)[0] for t in range(nts)])
        b.close()
    print({k: v[:2] for k, v in runs.items()})  # (cycles, FP32 stencil applications)
    assert runs[1][1] <= runs[0][1]
    for t, (ts, _) in enumerate(data):
        lamC = -runs[0][2][t][-1] * 1.3
        with lp.Tslice(L, deg, 12.0, lamC, device=gpu, ts_config=ts) as tsl:
            tsl.smearing_hyp(0.62, 0.62, 3)
            theta, _, _ = tsl.eps_solve(nev, tol=1e-12, want_vectors=False)
        evp = lp.recover_spectrum(theta, deg, 12.0, lamC)
        for on in (0, 1):
            np.testing.assert_allclose(runs[on][2][t], evp, rtol=1e-10, atol=1e-12, err_msg=str(on))


def test_dense_steps_on_unconverged_timeslices(lp, gpu, monkeypatch):
    """Once some timeslices have converged, the dense steps (projection,
    Rayleigh-Ritz, residuals) run on the runs of unconverged timeslices
    only (LAPH_DENSE_SUBSET=1, the default).  Six 8^3 timeslices, two of
    them (1 and 4) from a rougher field, with the per-cycle degree held at
    16 and the warm-up degree at DEG so that they converge a cycle after the
    others (6 cycles; a warm-up degree of 16 converges all six in 5): every
    timeslice's pairs against the dense eigensystem of its D (eigenvalues
    1e-10, residuals, orthonormality, phases), and the same spectrum as the
    dense steps on every timeslice (LAPH_DENSE_SUBSET=0)."""
    L, nev, deg, tol = 8, 40, 8, 1e-10
    eps = [0.2, 0.5, 0.2, 0.2, 0.5, 0.2]
    fields = {e: smeared(L, 1227, e, len(eps)) for e in set(eps)}
    data = [fields[e][t] for t, e in enumerate(eps)]
    monkeypatch.setenv("LAPH_FILTER_DEGREE", "16")
    monkeypatch.setenv("LAPH_WARM_DEG", str(deg))
    out = {}
    for on in (1, 0):
        monkeypatch.setenv("LAPH_DENSE_SUBSET", str(on))
        b = lp.Batch(L, deg, 12.0, 0.0, len(eps), nev, nsub=64, tol=tol, seed=5, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        print(f"LAPH_DENSE_SUBSET={on}: cycles {st.iterations} subset cycles {st.subset_cycles} "
              f"max res {st.max_residual:.2e}")
        assert st.converged == len(eps)
        assert (st.subset_cycles >= 1) if on else (st.subset_cycles == 0)
        out[on] = [b.get(t) for t in range(len(eps))]
        b.close()
    for t, (_, U) in enumerate(data):
        ev, V, _, res = out[1][t]
        check_pairs(ev, V, res, N.dense_D(U.reshape(L, L, L, 3, 3, 3)), nev, tol)
        np.testing.assert_allclose(ev, out[0][t][0], rtol=1e-12, atol=1e-14)


def test_pairs_meet_the_eps_criterion_at_the_derived_bound(lp, gpu):
    """SLEPc's relative test on T_DEG(B), |T_DEG(B) v - theta v| / |theta| <=
    1e-10 (the plain drop-in's -eps_tol), follows for every pair from a bound
    on |D v - lambda v| of 1e-10 / max_j c |T'_DEG(b_j)| / |T_DEG(b_j)|,
    b_j = c lambda_j + a (first order in the residual): the batched solver run
    at that bound, every pair's relative residual computed through the
    MatShell entry point."""
    L, nts, nev, deg = 12, 2, 40, 8
    data = smeared(L, 4242, 0.35, nts)

    def solve(tol):
        b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=64, tol=tol, seed=3, device=gpu)
        for t, (ts, _) in enumerate(data):
            b.set_timeslice(t, ts, 0.62, 0.62, 3)
        st = b.solve()
        assert st.converged == nts
        out = [b.get(t) for t in range(nts)]
        b.close()
        return out

    first = solve(1e-10)
    lam_c = 1.3 * max(-ev[-1] for ev, _, _, _ in first)
    c = 2.0 / (12.0 - lam_c)
    a = 1.0 + c * lam_c

    def factor(lam):  # c |T'_m(b)| / |T_m(b)|, b = cosh(psi) > 1
        psi = np.arccosh(c * lam + a)
        return c * deg * np.tanh(deg * psi) / np.sinh(psi)

    fmax = max(float(np.max(factor(ev))) for ev, _, _, _ in first)
    tol_d = 1e-10 / fmax
    print(f"lambda_c {lam_c:.4f}: largest factor {fmax:.2f}, bound on |D v - lambda v| {tol_d:.2e}")
    for t, (ev, V, _, res) in enumerate(solve(tol_d)):
        assert res.max() <= tol_d
        theta = np.cosh(deg * np.arccosh(c * ev + a))
        with lp.Tslice(L, deg, 12.0, lam_c, device=gpu, ts_config=data[t][0]) as tsl:
            tsl.smearing_hyp(0.62, 0.62, 3)
            rel = [np.linalg.norm(tsl.MatMult_Laplacian2D(V[j]) - theta[j] * V[j]) / abs(theta[j]) for j in range(nev)]
        print(f"timeslice {t}: largest relative EPS residual {max(rel):.2e}")
        assert max(rel) <= 1e-10
