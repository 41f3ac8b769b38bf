"""The FP32 filter does not make the filtered block's FP64 Gram singular:
numpy emulation (tests/fp32_filter_emulation.py) of the batched solver's
filter cycles at 8^3, K = 64, in complex64 and in complex128 arithmetic.  In
every cycle whose amplification stays bounded the equilibrated Gram's
smallest eigenvalue is the same in both to a few percent and every Gram
factors; in the uncapped schedule's four-pass cycle (amplification above
1e20) the Gram is singular at working precision with the FP64 filter as much
as with the FP32 one: the loss of rank comes from the amplification, not from
FP32 rounding.  The current schedule's per-cycle amplification respects the
1e10 bound after the warm-up cycle.  The emulated filters themselves are
pinned against T_m(B) by the plain three-term recurrence in complex128."""
import numpy as np
import pytest

import fp32_filter_emulation as E


@pytest.mark.parametrize("schedule", ["uncapped", "current"])
def test_fp32_filter_keeps_gram_conditioning(schedule):
    r32 = E.run(schedule, L=8, K=64, nev=48, cycles=3, dtype=np.complex64)
    r64 = E.run(schedule, L=8, K=64, nev=48, cycles=3, dtype=np.complex128)
    bounded = 3 if schedule == "current" else 2
    for a, b in zip(r32[:bounded], r64[:bounded]):
        assert a["cholesky_ok"] and b["cholesky_ok"]
        assert abs(np.log(a["gram_min_eig"] / b["gram_min_eig"])) < 0.1, (a, b)
    if schedule == "current":
        assert all(r["amplification"] <= 1.0001e10 for r in r32[1:])
    else:
        for r in (r32[2], r64[2]):
            assert r["amplification"] > 1e20 and r["gram_min_eig"] < 1e-12, r


def _dense_filter(U, X, c, a, m):
    """T_m(B) X by the plain three-term recurrence in complex128."""
    B = lambda v: c * E.block_D(U, v, np.complex128) + a * v  # noqa: E731
    t_old, t = X, B(X)
    for _ in range(2, m + 1):
        t_old, t = t, 2.0 * B(t) - t_old
    return t


def _setup(L=4, K=6, seed=3):
    cfg = E.N.synth_config(seed, 0.3, L, 1)
    U = E.N.hyp(E.N.links_from_config(cfg, L), 0.62, 0.62, 3)
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((3 * L ** 3, K)) + 1j * rng.standard_normal((3 * L ** 3, K))
    lc = 1.0
    c = 2.0 / (E.LAM_L - lc)
    return U, X, c, 1.0 + c * lc


@pytest.mark.parametrize("m", [4, 8, 12])
def test_emulated_correction_pass_is_the_chebyshev_filter(m):
    """The emulated correction pass (product over the roots of q_theta applied
    to r = (B - theta) X) equals T_m(B) X / s_m for any theta, in complex128:
    T_m(y) - T_m(theta) = q_theta(y) (y - theta) is a polynomial identity."""
    U, X, c, a = _setup()
    theta = np.array([-0.9, -0.3, 0.2, 0.7, 1.0, 1.04])
    g, w = E.filter_q_form(U, X, theta, c, a, m, np.complex128)
    ref = _dense_filter(U, X, c, a, m) / np.maximum(np.abs(E.cheb(m, theta)), 1.0)[None, :]
    got = g[None, :] * X + w
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-12


@pytest.mark.parametrize("passes", [1, 2])
def test_emulated_three_term_passes_are_powers_of_the_filter(passes):
    """The emulated three-term correction pass and plain passes give
    T_m(B)^P X / s_m^P (coefficients rounded to FP32 as the kernels take them,
    so to 1e-6 relative)."""
    U, X, c, a = _setup()
    theta = np.array([-0.9, -0.3, 0.2, 0.7, 1.0, 1.04])
    m = 8
    g, w = E.filter_three_term(U, X, theta, c, a, m, passes, np.complex128)
    sm = np.maximum(np.abs(E.cheb(m, theta)), 1.0)[None, :]
    ref = X
    for _ in range(passes):
        ref = _dense_filter(U, ref, c, a, m) / sm
    got = g[None, :] * X + w
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-5
