"""laph_shell_mult_noalloc (the filtered matvec with no context and no
allocation, recurrence buffers in the library's fixed device array) against
laph_shell_mult_device on the same links and vector: bitwise equal, on two
calls in a row, the second at a smaller lattice; the limit L <= 48 is
rejected before any device work.  Both run through the C-ABI."""
import ctypes

import numpy as np
import pytest

from oracle import lapheig_np as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(lp, torch, L, deg, lam_c, seed, stream):
    cfg = N.synth_config(seed, 0.4, L, 1)
    U = O.hyp(O.map_timeslice(cfg, L), L, 0.62, 0.62, 3)
    rng = np.random.default_rng(seed)
    n = 3 * L ** 3
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    dev = torch.device("cuda", 0)
    dU = torch.from_numpy(np.ascontiguousarray(U.reshape(-1))).to(dev)
    dx = torch.from_numpy(x).to(dev)
    dy_ref = torch.zeros(n, dtype=torch.complex128, device=dev)
    dy = torch.full((n,), float("nan"), dtype=torch.complex128, device=dev)
    torch.cuda.synchronize()
    with lp.Tslice(L, deg, 12.0, lam_c, device=0) as t:
        t.set_links(U)
        lp._check(lp.lib().laph_shell_mult_device(t.ctx, ctypes.c_void_p(dx.data_ptr()),
                                                  ctypes.c_void_p(dy_ref.data_ptr())), "laph_shell_mult_device")
    p = lp.laph_params(L, deg, 12.0, lam_c, 0)
    h = ctypes.c_void_p(stream.cuda_stream if stream is not None else 0)
    print(f"noalloc L={L} deg={deg} stream={'torch' if stream is not None else 'null'}", flush=True)
    lp._check(lp.lib().laph_shell_mult_noalloc(ctypes.byref(p), ctypes.c_void_p(dU.data_ptr()),
                                               ctypes.c_void_p(dx.data_ptr()), ctypes.c_void_p(dy.data_ptr()), h),
              "laph_shell_mult_noalloc")
    torch.cuda.synchronize()
    y, y_ref = dy.cpu().numpy(), dy_ref.cpu().numpy()
    np.testing.assert_array_equal(y, y_ref)
    ref = O.tv_iter(U, L, 12.0, lam_c, deg, x)
    assert abs(y - ref).max() < 1e-13 * abs(ref).max()
    np.testing.assert_array_equal(dx.cpu().numpy(), x)  # input untouched


def test_noalloc_matches_shell_mult_device(gpu):
    import torch

    from laph_eigsys_amd import lapheig as lp
    _case(lp, torch, 16, 12, 0.6, 3, None)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        pass
    _case(lp, torch, 8, 7, 0.9, 4, s)  # second call, smaller lattice, odd degree, a non-null stream
    _case(lp, torch, 6, 1, 0.9, 5, None)  # DEG 1: the single step writes y


def test_noalloc_rejects_beyond_scratch(gpu):
    from laph_eigsys_amd import lapheig as lp
    p = lp.laph_params(49, 8, 12.0, 0.6, 0)
    one = ctypes.c_void_p(16)  # never dereferenced: the size check comes first
    rc = lp.lib().laph_shell_mult_noalloc(ctypes.byref(p), one, ctypes.c_void_p(32), ctypes.c_void_p(48), None)
    assert rc == -1 and b"48" in lp.lib().laph_last_error()
