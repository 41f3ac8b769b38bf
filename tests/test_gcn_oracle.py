"""Pins the numpy restatement of the warm-start network (oracle/gcn_np.py):
the hand-derived backward pass against central finite differences, and
gauge covariance of the forward pass.  CPU only."""
import numpy as np

from oracle import gcn_np as G
from oracle import lapheig_np as N
from oracle import oracle as O


def setup(L=3, K=32, nev=5, layers=3, seed=0):
    rng = np.random.default_rng(seed)
    U = O.map_timeslice(N.synth_config(11, 0.4, L, 1), L).reshape(L, L, L, 3, 3, 3)
    n = 3 * L ** 3
    x = rng.standard_normal((n, K)) + 1j * rng.standard_normal((n, K))
    V, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    p = G.init_params(layers, seed)
    return U, x, V, p, layers


def test_backward_matches_finite_differences():
    U, x, V, p, layers = setup()
    loss, y, grad = G.loss_grad(p, layers, U, x, V)
    rng = np.random.default_rng(1)
    idx = list(rng.choice(len(p) - layers * 16, 12, replace=False)) + list(len(p) - 1 - np.arange(6))
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (G.loss_grad(p + e, layers, U, x, V)[0] - G.loss_grad(p - e, layers, U, x, V)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1e-3, abs(fd)) + 1e-9, (i, fd, grad[i])


def test_forward_gauge_covariant():
    U, x, V, p, layers = setup(L=4, K=16)
    L = 4
    rng = np.random.default_rng(5)
    H = rng.standard_normal((L, L, L, 3, 3)) + 1j * rng.standard_normal((L, L, L, 3, 3))
    H = H + np.conj(np.swapaxes(H, -1, -2))
    w, Q = np.linalg.eigh(H)
    Om = np.einsum("...ik,...k,...jk->...ij", Q, np.exp(1j * w), Q.conj())
    Ug = N.gauge_transform(U, Om)
    # psi'(x) = Om(x)^H psi(x) transforms like U'(x) = Om(x)^H U Om(x+mu)
    xf = G.to_field(x, L)
    xg = np.einsum("...ba,...bgc->...agc", np.conj(Om), xf).reshape(x.shape)
    y, _ = G.forward(p, layers, U, x)
    yg, _ = G.forward(p, layers, Ug, xg)
    want = np.einsum("...ba,...bgc->...agc", np.conj(Om), G.to_field(y, L)).reshape(y.shape)
    np.testing.assert_allclose(yg, want, atol=1e-11)


def test_loss_is_phase_and_rotation_invariant():
    U, x, V, p, layers = setup()
    l0 = G.loss_fn(x, V)[0]
    R, _ = np.linalg.qr(np.random.default_rng(2).standard_normal((V.shape[1],) * 2))
    assert abs(G.loss_fn(x, V @ R * np.exp(0.3j))[0] - l0) < 1e-13
    assert abs(G.loss_fn(x * np.exp(1.1j), V)[0] - l0) < 1e-13
