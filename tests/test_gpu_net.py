"""GPU kernels of the warm-start network against the numpy oracle
(oracle/gcn_np.py, itself pinned by finite differences): loss, output and
parameter gradients in FP32-activation mode (relative 1e-4), BF16 mode
(loss within 3e-2), gauge covariance of the device forward pass, and a short
train/propose/solve round trip."""
import numpy as np
import pytest

from oracle import gcn_np as G
from oracle import lapheig_np as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lp():
    from laph_eigsys_amd import lapheig
    lapheig.lib()
    return lapheig


def case(L=4, K=64, nev=6, seed=0):
    rng = np.random.default_rng(seed)
    U = O.map_timeslice(N.synth_config(21, 0.4, L, 1), L)
    n = 3 * L ** 3
    x = (rng.standard_normal((n, K)) + 1j * rng.standard_normal((n, K))).astype(np.complex64).astype(np.complex128)
    V, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    return U, x, V


def test_net_gradients_match_oracle(lp, gpu):
    L, layers = 4, 4
    U, x, V = case(L)
    net = lp.Net(layers=layers, seed=3, device=gpu)
    p = net.get_params().astype(np.float64)
    p[-layers * 16:] = np.linspace(-0.2, 0.2, layers * 16)  # exercise the gate
    net.set_params(p)
    p = net.get_params().astype(np.float64)
    loss, y, grad = net.loss_grad(L, U, x, V, fp32_act=True)
    loss0, y0, grad0 = G.loss_grad(p, layers, U.reshape(L, L, L, 3, 3, 3), x, V)
    print(f"loss gpu {loss:.8f} oracle {loss0:.8f}; max|dgrad|/max|grad| "
          f"{abs(grad - grad0).max() / abs(grad0).max():.2e}")
    assert abs(loss - loss0) < 1e-5 * abs(loss0)
    assert abs(y - y0).max() < 1e-5 * abs(y0).max()
    assert abs(grad - grad0).max() < 1e-4 * abs(grad0).max()
    lb, _, gb = net.loss_grad(L, U, x, V, fp32_act=False)
    assert abs(lb - loss0) < 3e-2 * abs(loss0)
    net.close()


def test_net_gauge_covariance(lp, gpu):
    L = 4
    U, x, V = case(L, seed=4)
    net = lp.Net(layers=4, seed=7, device=gpu)
    rng = np.random.default_rng(9)
    H = rng.standard_normal((L, L, L, 3, 3)) + 1j * rng.standard_normal((L, L, L, 3, 3))
    H = H + np.conj(np.swapaxes(H, -1, -2))
    w, Q = np.linalg.eigh(H)
    Om = np.einsum("...ik,...k,...jk->...ij", Q, np.exp(1j * w), Q.conj())
    Ug = N.gauge_transform(U.reshape(L, L, L, 3, 3, 3), Om).reshape(-1, 3, 3, 3)
    xf = G.to_field(x, L)
    xg = np.einsum("...ba,...bgc->...agc", np.conj(Om), xf).reshape(x.shape)
    _, y, _ = net.loss_grad(L, U, x, V)
    _, yg, _ = net.loss_grad(L, Ug, xg, V)
    want = np.einsum("...ba,...bgc->...agc", np.conj(Om), G.to_field(y, L)).reshape(y.shape)
    assert abs(yg - want).max() < 1e-5 * abs(y).max()
    net.close()


def test_train_propose_solve(lp, gpu):
    """Training lowers the loss; a proposed start converges to the same spectrum."""
    L, nts, nev = 8, 2, 20
    cfg = N.synth_config(31, 0.35, L, nts)
    b = lp.Batch(L, 12, 12.0, 0.0, nts, nev, nsub=64, tol=1e-10, seed=2, device=gpu, warm_start=1)
    for t in range(nts):
        b.set_timeslice(t, cfg[t * 72 * L ** 3:(t + 1) * 72 * L ** 3], 0.62, 0.62, 3)
    st0 = b.solve()
    ev0 = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    net = lp.Net(layers=4, seed=5, device=gpu)
    losses = [net.train_step(b, seed=100 + i, lr=3e-3) for i in range(40)]
    print("loss", losses[0], "->", losses[-1])
    assert losses[-1] < losses[0]
    net.propose(b, seed=999)
    st1 = b.solve()
    print(f"cycles random {st0.iterations} warm {st1.iterations}")
    assert st1.converged == nts
    for t in range(nts):
        np.testing.assert_allclose(b.get(t, want_vectors=False)[0], ev0[t], rtol=1e-10, atol=1e-12)
    net.close()
    b.close()
