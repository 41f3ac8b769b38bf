"""ctypes front end of the C oracle (oracle/lapheig_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Arrays are numpy complex128 in the layouts documented in lapheig_oracle.h:
links U[V3, 3, 3, 3] (site, dir, row, col), vectors v[3*V3] (3*site+colour).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liblapheig_oracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_up.argtypes = [ctypes.c_int] * 3
        L.orc_dn.argtypes = [ctypes.c_int] * 3
        L.orc_read_lime.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 4 + [_dp]
        L.orc_map_timeslice.argtypes = [_dp, ctypes.c_int, _dp]
        L.orc_herm3_eig.argtypes = [_dp, _dp, _dp]
        L.orc_proj_su3.argtypes = [_dp]
        L.orc_hyp.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_tv2_ca.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_tv2.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_tv_iter.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_int, _dp, _dp]
        L.orc_eps_solve.argtypes = [_dp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_ulonglong, _dp, _dp,
                                    ctypes.POINTER(ctypes.c_longlong)]
        L.orc_invert_B.restype = ctypes.c_double
        L.orc_invert_B.argtypes = [ctypes.c_double] * 3
        L.orc_invert_Tn.restype = ctypes.c_double
        L.orc_invert_Tn.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_recover_spectrum.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                           ctypes.c_double, _dp]
        L.orc_fix_phase.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_reset_phase.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_check_trace.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
        L.orc_dsyev.argtypes = [ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def up(L, s, d):
    return lib().orc_up(L, s, d)


def dn(L, s, d):
    return lib().orc_dn(L, s, d)


def read_lime(path, T, L, ti, tf):
    out = np.zeros((tf - ti + 1) * L ** 3 * 72, dtype=np.float64)
    rc = lib().orc_read_lime(path.encode(), T, L, ti, tf, _p(out))
    if rc != 0:
        raise IOError(f"orc_read_lime failed with {rc}")
    return out


def map_timeslice(ts_config, L):
    ts_config = np.ascontiguousarray(ts_config, dtype=np.float64)
    U = np.zeros((L ** 3, 3, 3, 3), dtype=np.complex128)
    lib().orc_map_timeslice(_p(ts_config), L, _p(U))
    return U


def herm3_eig(A):
    A = c128(A)
    w = np.zeros(3)
    Q = np.zeros((3, 3), dtype=np.complex128)
    lib().orc_herm3_eig(_p(A), _p(w), _p(Q))
    return w, Q


def proj_su3(M):
    M = c128(M).copy()
    lib().orc_proj_su3(_p(M))
    return M


def hyp(U, L, a1, a2, it):
    U = c128(U).copy()
    lib().orc_hyp(_p(U), L, a1, a2, it)
    return U


def tv2(U, L, lamL, lamC, x):
    U = c128(U)
    x = c128(x)
    y = np.zeros_like(x)
    lib().orc_tv2(_p(U), L, lamL, lamC, _p(x), _p(y))
    return y


def apply_D(U, L, x):
    """D x = sum_mu [U x(x+mu) + U^H x(x-mu)] - 6 x   (tv2 with c=1, a=0)."""
    U = c128(U)
    x = c128(x)
    y = np.zeros_like(x)
    lib().orc_tv2_ca(_p(U), L, 1.0, 0.0, _p(x), _p(y))
    return y


def tv_iter(U, L, lamL, lamC, deg, x):
    U = c128(U)
    x = c128(x)
    y = np.zeros_like(x)
    rc = lib().orc_tv_iter(_p(U), L, lamL, lamC, deg, _p(x), _p(y))
    if rc != 0:
        raise ValueError("deg < 1")
    return y


def eps_solve(U, L, lamL, lamC, deg, nev, ncv=0, tol=1e-10, max_it=0, seed=1227):
    U = c128(U)
    n = 3 * L ** 3
    theta = np.zeros(nev)
    V = np.zeros((nev, n), dtype=np.complex128)  # row i = eigenvector i (column-major n x nev)
    stats = (ctypes.c_longlong * 2)()
    nconv = lib().orc_eps_solve(_p(U), L, lamL, lamC, deg, nev, ncv, tol, max_it, seed,
                                _p(theta), _p(V), stats)
    return nconv, theta, V, (stats[0], stats[1])


def recover_spectrum(theta, deg, lamL, lamC):
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.zeros_like(theta)
    lib().orc_recover_spectrum(len(theta), _p(theta), deg, lamL, lamC, _p(out))
    return out


def fix_phase(V):
    """V: (nev, n) rows are eigenvectors (column-major n x nev)."""
    V = c128(V)
    nev, n = V.shape
    Vf = np.zeros_like(V)
    ph = np.zeros(nev)
    lib().orc_fix_phase(n, nev, _p(V), _p(Vf), _p(ph))
    return Vf, ph


def reset_phase(V, want):
    V = c128(V)
    nev, n = V.shape
    want = np.ascontiguousarray(want, dtype=np.float64)
    Vr = np.zeros_like(V)
    lib().orc_reset_phase(n, nev, _p(V), _p(want), _p(Vr))
    return Vr


def check_trace(V):
    V = c128(V)
    nev, n = V.shape
    tr = np.zeros(2)
    sm = np.zeros(2)
    ok = lib().orc_check_trace(n, nev, _p(V), _p(tr), _p(sm))
    return bool(ok), complex(tr[0], tr[1]), complex(sm[0], sm[1])


def dsyev(A):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    m = A.shape[0]
    w = np.zeros(m)
    rc = lib().orc_dsyev(m, _p(A), _p(w))
    if rc != 0:
        raise RuntimeError("orc_dsyev did not converge")
    return w, A
