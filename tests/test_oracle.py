"""Pins the C oracle (oracle/lapheig_oracle.c) without the reference binary:
closed forms, dense diagonalisation, invariants, and the independent numpy
restatement (oracle/lapheig_np.py).  CPU only."""
import numpy as np
import pytest

from oracle import lapheig_np as N
from oracle import oracle as O

RNG = np.random.default_rng(7)


def field(L, seed=1227, eps=0.4):
    cfg = N.synth_config(seed, eps, L, 1)
    return O.map_timeslice(cfg, L)


def rand_vec(n, rng=RNG):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def rand_su3(rng, k=1):
    H = rng.standard_normal((k, 3, 3)) + 1j * rng.standard_normal((k, 3, 3))
    H = H + np.conj(np.swapaxes(H, -1, -2))
    w, Q = np.linalg.eigh(H)
    U = np.einsum("kij,kj,klj->kil", Q, np.exp(1j * w), Q.conj())
    return U / np.linalg.det(U)[:, None, None] ** (1 / 3)


# ---------------------------------------------------------------- navigation
@pytest.mark.parametrize("L", [3, 4, 5])
def test_navigation_closed_form(L):
    for s in range(L ** 3):
        x, y, z = s // (L * L), (s // L) % L, s % L
        assert O.up(L, s, 0) == ((x + 1) % L) * L * L + y * L + z
        assert O.dn(L, s, 1) == x * L * L + ((y - 1) % L) * L + z
        assert O.up(L, s, 2) == x * L * L + y * L + (z + 1) % L


# ---------------------------------------------------------------- LIME + map
def coded_config(T, L):
    """Entry (t,x,y,z,slot mu',a,b) = t*1e4 + x*1e3 + y*100 + z*10 + mu' + (3a+b)/10, imag = -that."""
    t, x, y, z, m, e = np.meshgrid(np.arange(T), np.arange(L), np.arange(L), np.arange(L), np.arange(4),
                                   np.arange(9), indexing="ij")
    v = t * 1e4 + x * 1e3 + y * 100 + z * 10 + m + e / 10
    out = np.stack([v, -v], axis=-1)
    return out.reshape(-1)


def test_lime_known_answer_oracle_reader(tmp_path):
    T, L = 3, 4
    cfg = coded_config(T, L)
    p = str(tmp_path / "conf.lime")
    N.write_lime(p, cfg, T, L)
    got = O.read_lime(p, T, L, 1, 2)
    # io.cpp:114-131: position ((t-ti)*L^3 + x L^2 + y L + z)*72 + slot*18 + 6a + 2b
    for (t, x, y, z, slot, a, b) in [(1, 0, 0, 0, 0, 0, 0), (2, 3, 1, 2, 2, 1, 2), (1, 1, 2, 3, 3, 2, 0)]:
        pos = (((t - 1) * L ** 3 + x * L * L + y * L + z) * 72) + slot * 18 + 6 * a + 2 * b
        want = t * 1e4 + x * 1e3 + y * 100 + z * 10 + slot + (3 * a + b) / 10
        assert got[pos] == want and got[pos + 1] == -want
    np.testing.assert_array_equal(got, cfg[L ** 3 * 72:])


def test_lime_errors(tmp_path):
    T, L = 2, 3
    p = str(tmp_path / "c.lime")
    N.write_lime(p, coded_config(T, L), T, L)
    with pytest.raises(IOError):
        O.read_lime(p, T, 4, 0, 0)  # wrong lattice size (io.cpp:63-71)
    with pytest.raises(IOError):
        O.read_lime(str(tmp_path / "missing"), T, L, 0, 0)


def test_map_timeslice_known_answer():
    L = 3
    cfg = coded_config(1, L)
    U = O.map_timeslice(cfg, L)
    for s in range(L ** 3):
        x, y, z = s // 9, (s // 3) % 3, s % 3
        for d in range(3):
            for a in range(3):
                for b in range(3):
                    v = x * 1e3 + y * 100 + z * 10 + (d + 1) + (3 * a + b) / 10
                    assert U[s, d, a, b] == complex(v, -v)


def test_synth_config_is_su3():
    cfg = N.synth_config(99, 0.45, 3, 2)
    c = cfg.reshape(-1, 4, 3, 3, 2)
    U = c[..., 0] + 1j * c[..., 1]
    np.testing.assert_allclose(U @ np.conj(np.swapaxes(U, -1, -2)), np.broadcast_to(np.eye(3), U.shape), atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(U), 1.0, atol=1e-14)


# ---------------------------------------------------------------- projection
def test_herm3_eig():
    for _ in range(20):
        A = RNG.standard_normal((3, 3)) + 1j * RNG.standard_normal((3, 3))
        A = A + A.conj().T
        w, Q = O.herm3_eig(A)
        np.testing.assert_allclose(np.sort(w), np.linalg.eigvalsh(A), atol=1e-13)
        np.testing.assert_allclose(Q @ np.diag(w) @ Q.conj().T, A, atol=1e-13)


def test_proj_su3_polar_uniqueness():
    W = rand_su3(RNG, 10)
    for k in range(10):
        np.testing.assert_allclose(O.proj_su3(W[k]), W[k], atol=1e-14)           # P[W] = W
        np.testing.assert_allclose(O.proj_su3(2.7 * W[k]), W[k], atol=1e-14)     # P[sW] = W
        X = RNG.standard_normal((3, 3)) + 1j * RNG.standard_normal((3, 3))
        Hp = X @ X.conj().T + 0.5 * np.eye(3)                                  # Hermitian > 0
        np.testing.assert_allclose(O.proj_su3(W[k] @ Hp), W[k], atol=1e-12)   # P[W H] = W
        M = RNG.standard_normal((3, 3)) + 1j * RNG.standard_normal((3, 3))
        P = O.proj_su3(M)
        np.testing.assert_allclose(P @ P.conj().T, np.eye(3), atol=1e-13)
        # det is 1 unless arg det(W) sits at the principal-root branch cut
        d = np.linalg.det(P)
        assert abs(d - 1) < 1e-13 or abs(abs(d) - 1) < 1e-13


def test_proj_matches_numpy():
    M = RNG.standard_normal((50, 3, 3)) + 1j * RNG.standard_normal((50, 3, 3))
    M = M + 3 * np.eye(3)  # near-identity regime of HYP
    ref = N.proj_su3(M)
    for k in range(50):
        np.testing.assert_allclose(O.proj_su3(M[k]), ref[k], atol=1e-13)


# ---------------------------------------------------------------- HYP
def test_hyp_unit_field_fixed_point():
    L = 4
    U = np.broadcast_to(np.eye(3, dtype=complex), (L ** 3, 3, 3, 3)).copy()
    np.testing.assert_allclose(O.hyp(U, L, 0.62, 0.62, 3), U, atol=1e-15)


def test_hyp_constant_abelian_fixed_point():
    L = 4
    U = np.empty((L ** 3, 3, 3, 3), dtype=complex)
    for d in range(3):
        th = 0.3 * (d + 1)
        U[:, d] = np.diag(np.exp(1j * np.array([th, 0.7 * th, -1.7 * th])))
    np.testing.assert_allclose(O.hyp(U, L, 0.62, 0.62, 3), U, atol=1e-14)


def test_hyp_alpha_zero_identity():
    L = 4
    U = field(L)
    np.testing.assert_allclose(O.hyp(U, L, 0.0, 0.0, 2), U, atol=1e-14)


def test_hyp_matches_numpy_restatement():
    L = 4
    U = field(L)
    got = O.hyp(U, L, 0.62, 0.62, 3).reshape(L, L, L, 3, 3, 3)
    ref = N.hyp(U.reshape(L, L, L, 3, 3, 3), 0.62, 0.62, 3)
    np.testing.assert_allclose(got, ref, atol=1e-13)


def test_hyp_gauge_covariance():
    L = 4
    U = field(L)
    Om = rand_su3(RNG, L ** 3).reshape(L, L, L, 3, 3)
    Ug = N.gauge_transform(U.reshape(L, L, L, 3, 3, 3), Om)
    a = O.hyp(Ug.reshape(-1, 3, 3, 3), L, 0.62, 0.62, 3).reshape(L, L, L, 3, 3, 3)
    b = N.gauge_transform(O.hyp(U, L, 0.62, 0.62, 3).reshape(L, L, L, 3, 3, 3), Om)
    np.testing.assert_allclose(a, b, atol=1e-13)


# ---------------------------------------------------------------- stencil
def test_free_field_plane_waves():
    L = 6
    U = np.broadcast_to(np.eye(3, dtype=complex), (L ** 3, 3, 3, 3)).copy()
    x, y, z = np.meshgrid(np.arange(L), np.arange(L), np.arange(L), indexing="ij")
    for p in [(0, 0, 0), (1, 0, 0), (1, 2, 3), (5, 5, 1)]:
        k = 2 * np.pi * np.array(p) / L
        ph = np.exp(1j * (k[0] * x + k[1] * y + k[2] * z))
        psi = np.zeros((L, L, L, 3), dtype=complex)
        psi[..., 1] = ph
        lam = sum(2 * (1 - np.cos(ki)) for ki in k)
        np.testing.assert_allclose(O.apply_D(U, L, psi.reshape(-1)), -lam * psi.reshape(-1), atol=1e-13)
        # T_n(B) psi = T_n(b) psi with b = 1 + c (lamC - lam)
        lamL, lamC, deg = 12.0, 0.5, 8
        b = 1 + 2 / (lamL - lamC) * (lamC - lam)
        tn = np.cos(deg * np.arccos(b)) if abs(b) <= 1 else np.cosh(deg * np.arccosh(abs(b))) * np.sign(b) ** deg
        np.testing.assert_allclose(O.tv_iter(U, L, lamL, lamC, deg, psi.reshape(-1)), tn * psi.reshape(-1),
                                   atol=1e-11)


def test_stencil_hermitian_and_dense():
    L = 4
    U = O.hyp(field(L), L, 0.62, 0.62, 3)
    n = 3 * L ** 3
    u, v = rand_vec(n), rand_vec(n)
    lhs = np.vdot(u, O.apply_D(U, L, v))
    rhs = np.vdot(O.apply_D(U, L, u), v)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
    np.testing.assert_allclose(D, D.conj().T, atol=1e-14)
    np.testing.assert_allclose(O.apply_D(U, L, v), D @ v, atol=1e-12)
    lamL, lamC = 12.0, 0.7
    c = 2 / (lamL - lamC)
    B = c * D + (1 + c * lamC) * np.eye(n)
    np.testing.assert_allclose(O.tv2(U, L, lamL, lamC, v), B @ v, atol=1e-12)
    T0, T1 = np.eye(n), B
    for _ in range(2, 7):
        T0, T1 = T1, 2 * B @ T1 - T0
    np.testing.assert_allclose(O.tv_iter(U, L, lamL, lamC, 6, v), T1 @ v, rtol=1e-12, atol=1e-10)
    with pytest.raises(ValueError):
        O.tv_iter(U, L, lamL, lamC, 0, v)


# ---------------------------------------------------------------- spectrum
def test_recover_spectrum_round_trip():
    lamL, lamC, deg = 12.0, 0.8, 12
    lam = np.linspace(-0.79, 0.0, 17)  # eigenvalues of D above -lamC
    b = 1 + 2 / (lamL - lamC) * (lamC + lam)
    theta = np.cosh(deg * np.arccosh(b))
    np.testing.assert_allclose(O.recover_spectrum(theta, deg, lamL, lamC), lam, atol=1e-13)


def test_fix_phase_and_trace():
    n, nev = 60, 7
    Q, _ = np.linalg.qr(RNG.standard_normal((n, nev)) + 1j * RNG.standard_normal((n, nev)))
    V = Q.T.copy()
    Vf, ph = O.fix_phase(V)
    np.testing.assert_allclose(ph, np.angle(V[:, 0]))
    assert np.all(np.abs(Vf[:, 0].imag) < 1e-15) and np.all(Vf[:, 0].real > 0)
    np.testing.assert_allclose(np.linalg.norm(Vf, axis=1), 1.0)
    ok, tr, sm = O.check_trace(Vf)
    assert ok and abs(tr - nev) < 1e-12
    assert not O.check_trace(2 * Vf)[0]
    Vr = O.reset_phase(Vf, ph)
    np.testing.assert_allclose(Vr, V, atol=1e-14)


def test_dsyev_matches_numpy():
    for m in (1, 2, 5, 40):
        A = RNG.standard_normal((m, m))
        A = A + A.T
        w, Q = O.dsyev(A)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(A), atol=1e-12)
        np.testing.assert_allclose(A @ Q, Q * w, atol=1e-11)


@pytest.mark.parametrize("L,nev,deg,tol", [(4, 10, 8, 1e-12), (8, 20, 8, 1e-12)])
def test_eps_solve_against_dense_eigh(L, nev, deg, tol):
    U = O.hyp(field(L), L, 0.62, 0.62, 3)
    D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
    w = np.sort(np.linalg.eigvalsh(D))[::-1]  # 0^- first
    lamC = -w[nev + 4] * 1.05
    nconv, theta, V, _ = O.eps_solve(U, L, 12.0, lamC, deg, nev, tol=tol)
    assert nconv >= nev
    ev = O.recover_spectrum(theta, deg, 12.0, lamC)
    np.testing.assert_allclose(ev, w[:nev], rtol=1e-10, atol=1e-12)
    for i in range(nev):
        r = np.linalg.norm(D @ V[i] - ev[i] * V[i])
        assert r < 1e-9
    assert O.check_trace(V)[0]


def test_sparse_assembly_matches_stencil():
    """The entry-by-entry sparse -D (BuildLaplacian layout, source of the C2
    fixture) equals the numpy stencil, and is Hermitian."""
    L = 5
    U = N.hyp(N.links_from_config(N.synth_config(3, 0.4, L, 1), L), 0.62, 0.62, 1)
    A = N.sparse_minus_D(U)
    rng = np.random.default_rng(0)
    v = rng.normal(size=(L, L, L, 3)) + 1j * rng.normal(size=(L, L, L, 3))
    assert np.abs(A @ v.reshape(-1) + N.apply_D(U, v).reshape(-1)).max() < 1e-13
    assert abs(A - A.conj().T).max() == 0


def test_source_shape_oracle_brute_force():
    """The numpy source shape (roll + einsum) equals the reference's loops
    written out literally (dll_op per site, sourceshape_funcs.cpp:31-81) on a
    3^3 lattice, and avg_radii averages equal radii in first-appearance order."""
    L, nev = 3, 4
    n = 3 * L ** 3
    rng = np.random.default_rng(1)
    V = rng.standard_normal((nev, n)) + 1j * rng.standard_normal((nev, n))
    got = N.source_shape(V, L)
    idx = lambda x, y, z: ((x % L) * L + y % L) * L + z % L  # noqa: E731
    k = 0
    for r1 in range(0, L, 2):
        for r2 in range(0, L, 2):
            for r3 in range(0, L, 2):
                psi = 0.0
                for x in range(L):
                    for y in range(L):
                        for z in range(L):
                            d, s = idx(x + r1, y + r2, z + r3), idx(x, y, z)
                            M = V[:, 3 * d:3 * d + 3].T @ V[:, 3 * s:3 * s + 3].conj()
                            psi += np.sqrt(np.trace(M @ M.conj().T).real)
                assert got[k][0] == (r1, r2, r3)
                assert abs(got[k][2] - psi) < 1e-12 * psi
                k += 1
    assert k == len(got) == 8
    res = [((0, 0, 0), 0.0, 1.0), ((0, 0, 2), 2.0, 2.0), ((0, 2, 0), 2.0, 4.0), ((2, 2, 2), 12 ** 0.5, 5.0)]
    assert N.avg_radii(res) == [(0.0, 1.0), (2.0, 3.0), (12 ** 0.5, 5.0)]
