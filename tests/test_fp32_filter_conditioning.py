"""The FP32 filter does not make the filtered block's FP64 Gram singular:
numpy emulation (tests/fp32_filter_emulation.py) of the batched solver's
filter cycles at 8^3, K = 64, in complex64 and in complex128 arithmetic.  The
equilibrated Gram's smallest eigenvalue is the same in both to a few percent
in every cycle (its drop in the multi-pass cycles of the uncapped schedule
comes from the amplification, not from FP32 rounding), and every Gram
factors.  The current schedule's per-cycle amplification respects the 1e10
bound after the warm-up cycle."""
import numpy as np
import pytest

import fp32_filter_emulation as E


@pytest.mark.parametrize("schedule", ["uncapped", "current"])
def test_fp32_filter_keeps_gram_conditioning(schedule):
    r32 = E.run(schedule, L=8, K=64, nev=48, cycles=3, dtype=np.complex64)
    r64 = E.run(schedule, L=8, K=64, nev=48, cycles=3, dtype=np.complex128)
    for a, b in zip(r32, r64):
        assert a["cholesky_ok"] and b["cholesky_ok"]
        assert abs(np.log(a["gram_min_eig"] / b["gram_min_eig"])) < 0.1, (a, b)
    if schedule == "current":
        assert all(r["amplification"] <= 1.0001e10 for r in r32[1:])
    else:
        assert r32[2]["amplification"] > 1e20 and r32[2]["gram_min_eig"] < 1e-6
