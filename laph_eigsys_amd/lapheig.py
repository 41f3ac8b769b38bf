"""ctypes front end of liblapheig.so (include/laph.h).

Mirrors the reference's host API for the per-timeslice eigensystem path
(names and argument meaning of ev_ts.cpp / timeslice.cpp / shell_matop.cpp /
eigensystem.cpp / recover_spec.cpp / read_write.cpp / io.cpp), with numpy
arrays in place of Eigen/PETSc objects:

  Tslice(params).map_timeslice_to_eigen(ts_config)   timeslice.cpp:34-79
  Tslice.smearing_hyp(alpha_1, alpha_2, iter)          timeslice.cpp:243-369
  Tslice.MatMult_Laplacian2D(x)                        shell_matop.cpp:226-244
  Tslice.MatGetDiagonal_Laplacian2D()                  shell_matop.cpp:248-255
  Tslice.eps_solve(nev, ...)                           ev_ts.cpp:231-303
  recover_spectrum, fix_phase, reset_phase, check_trace, write_eig_sys_bin,
  write_eigenvalues_bin, read_lime_gauge_field_doubleprec_timeslices.

Errors raise LaphError (the C code never exits, unlike the reference).  The
library must be built (laph_eigsys_amd/build.py); nothing here falls back to
another implementation.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblapheig.so")

_dp = ctypes.POINTER(ctypes.c_double)


class LaphError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed with {code}: {msg}")
        self.code = code


class laph_params(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int), ("deg", ctypes.c_int), ("lambda_l", ctypes.c_double),
                ("lambda_c", ctypes.c_double), ("device", ctypes.c_int)]


class laph_eps_opts(ctypes.Structure):
    _fields_ = [("nev", ctypes.c_int), ("ncv", ctypes.c_int), ("tol", ctypes.c_double),
                ("max_it", ctypes.c_int), ("seed", ctypes.c_ulonglong)]


class laph_eps_stats(ctypes.Structure):
    _fields_ = [("restarts", ctypes.c_int), ("nconv", ctypes.c_int), ("matvecs", ctypes.c_longlong),
                ("seconds", ctypes.c_double)]


class laph_batch_opts(ctypes.Structure):
    _fields_ = [("nts", ctypes.c_int), ("nev", ctypes.c_int), ("nsub", ctypes.c_int),
                ("tol", ctypes.c_double), ("max_iter", ctypes.c_int), ("seed", ctypes.c_ulonglong),
                ("warm_start", ctypes.c_int), ("filter_passes", ctypes.c_int),
                ("filter_storage", ctypes.c_int)]


class laph_batch_stats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("stencil_vecs", ctypes.c_longlong),
                ("stencil64_vecs", ctypes.c_longlong), ("converged", ctypes.c_int),
                ("max_residual", ctypes.c_double), ("seconds", ctypes.c_double),
                ("seconds_filter", ctypes.c_double), ("seconds_dense", ctypes.c_double),
                ("filter_launches", ctypes.c_longlong), ("filter_bytes", ctypes.c_double),
                ("filter_block", ctypes.c_int), ("filter_strip", ctypes.c_int), ("recoveries", ctypes.c_int),
                ("filter_zrun", ctypes.c_int), ("start_window", ctypes.c_double),
                ("rr_cholqr_fallbacks", ctypes.c_int), ("shifted_cholqr", ctypes.c_int), ("gram_blocks", ctypes.c_int),
                ("gram_chunks", ctypes.c_int), ("gram_mfma", ctypes.c_int),
                ("layout_from_cache", ctypes.c_int), ("warmup_degree", ctypes.c_int),
                ("subset_cycles", ctypes.c_int)]


# (name, argtypes) of every entry point declared in include/laph.h
_VP = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_S = ctypes.c_char_p
SIGNATURES = {
    "laph_error_string": ([_I], ctypes.c_char_p),
    "laph_last_error": ([], ctypes.c_char_p),
    "laph_version": ([], ctypes.c_char_p),
    "laph_ctx_create": ([ctypes.POINTER(_VP), ctypes.POINTER(laph_params), _dp], _I),
    "laph_ctx_destroy": ([_VP], _I),
    "laph_map_timeslice": ([_VP, _dp], _I),
    "laph_set_links": ([_VP, _dp], _I),
    "laph_get_links": ([_VP, _dp], _I),
    "laph_hyp_smear": ([_VP, _D, _D, _I], _I),
    "laph_shell_mult": ([_VP, _dp, _dp], _I),
    "laph_shell_get_diagonal": ([_VP, _dp], _I),
    "laph_shell_mult_device": ([_VP, _VP, _VP], _I),
    "laph_shell_mult_noalloc": ([ctypes.POINTER(laph_params), _VP, _VP, _VP, _VP], _I),
    "laph_laplacian": ([_VP, _dp, _dp], _I),
    "laph_eps_solve": ([_VP, ctypes.POINTER(laph_eps_opts), _dp, _dp, ctypes.POINTER(laph_eps_stats)], _I),
    "laph_recover_spectrum": ([_I, _dp, _I, _D, _D, _dp], _I),
    "laph_fix_phase": ([_I, _I, _dp, _dp, _dp], _I),
    "laph_reset_phase": ([_I, _I, _dp, _dp, _dp], _I),
    "laph_check_trace": ([_I, _I, _dp, _dp, _dp], _I),
    "laph_write_eig_sys_bin": ([_S, _S, _I, _I, _I, _I, _dp], _I),
    "laph_write_eigenvalues_bin": ([_S, _S, _I, _I, _I, _dp], _I),
    "laph_read_evectors_bin_ts": ([_S, _S, _I, _I, _I, _I, _dp], _I),
    "laph_read_eigenvalues_bin": ([_S, _S, _I, _I, _I, _dp], _I),
    "laph_read_eigenvalues_ascii": ([_S, _I, _I, _I, _dp], _I),
    "laph_read_lime": ([_S, _I, _I, _I, _I, _dp], _I),
    "laph_write_lime": ([_S, _I, _I, _dp], _I),
    "laph_synth_config": ([ctypes.c_ulonglong, _D, _I, _I, _dp], _I),
    "laph_batch_create": ([ctypes.POINTER(_VP), ctypes.POINTER(laph_params), ctypes.POINTER(laph_batch_opts)], _I),
    "laph_batch_destroy": ([_VP], _I),
    "laph_batch_set_timeslice": ([_VP, _I, _dp, _D, _D, _I], _I),
    "laph_batch_set_links": ([_VP, _I, _dp], _I),
    "laph_batch_solve": ([_VP, ctypes.POINTER(laph_batch_stats)], _I),
    "laph_batch_get": ([_VP, _I, _dp, _dp, _dp, _dp], _I),
    "laph_batch_get_device": ([_VP, _I, _VP, _dp], _I),
    "laph_batch_time_filter": ([_VP, _I, _dp, _dp], _I),
    "laph_batch_time_gram": ([_VP, _I, _dp, _dp], _I),
    "laph_batch_write": ([_VP, _I, _S, _I, _I, _dp], _I),
    "laph_batch_source_shape": ([_VP, _I, _dp], _I),
    "laph_source_shape": ([_I, _I, _dp, _I, _dp], _I),
    "laph_net_create": ([ctypes.POINTER(_VP), _I, ctypes.c_ulonglong, _I], _I),
    "laph_net_destroy": ([_VP], _I),
    "laph_net_num_params": ([_VP], _I),
    "laph_net_get_params": ([_VP, ctypes.POINTER(ctypes.c_float)], _I),
    "laph_net_set_params": ([_VP, ctypes.POINTER(ctypes.c_float)], _I),
    "laph_net_propose": ([_VP, _VP, ctypes.c_ulonglong], _I),
    "laph_net_train_step": ([_VP, _VP, ctypes.c_ulonglong, _D, _dp], _I),
    "laph_net_loss_grad": ([_VP, _I, _I, _I, _dp, ctypes.POINTER(ctypes.c_float), _dp, _I, _dp, _dp,
                            ctypes.POINTER(ctypes.c_float)], _I),
}

_lib = None


def lib():
    """Loads liblapheig.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run laph_eigsys_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc, func):
    if rc != 0:
        msg = lib().laph_last_error().decode(errors="replace")
        raise LaphError(rc, func, msg)


def _p(a):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(_dp)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


# ---------------------------------------------------------------------------
def synth_config(seed, eps, L, T):
    """io.cpp-layout config of the seeded field U = exp(i eps H): T*72*L^3 doubles."""
    out = np.empty(T * 72 * L ** 3, dtype=np.float64)
    _check(lib().laph_synth_config(seed, eps, L, T, _p(out)), "laph_synth_config")
    return out


def read_lime_gauge_field_doubleprec_timeslices(path, T, L, ti, tf):
    """io.cpp:22-146."""
    out = np.empty((tf - ti + 1) * 72 * L ** 3, dtype=np.float64)
    _check(lib().laph_read_lime(os.fsencode(path), T, L, ti, tf, _p(out)), "laph_read_lime")
    return out


def write_lime(path, config, T, L):
    config = np.ascontiguousarray(config, dtype=np.float64)
    _check(lib().laph_write_lime(os.fsencode(path), T, L, _p(config)), "laph_write_lime")


def recover_spectrum(theta, deg, lambda_l, lambda_c):
    """recover_spec.cpp:45-54."""
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    out = np.empty_like(theta)
    _check(lib().laph_recover_spectrum(len(theta), _p(theta), deg, lambda_l, lambda_c, _p(out)),
           "laph_recover_spectrum")
    return out


def fix_phase(V):
    """eigensystem.cpp:4-30.  V: (nev, n), row i = eigenvector i."""
    V = _c128(V)
    nev, n = V.shape
    Vf = np.empty_like(V)
    ph = np.empty(nev)
    _check(lib().laph_fix_phase(n, nev, _p(V), _p(Vf), _p(ph)), "laph_fix_phase")
    return Vf, ph


def reset_phase(V, phase_want):
    """eigensystem.cpp:32-57."""
    V = _c128(V)
    nev, n = V.shape
    want = np.ascontiguousarray(phase_want, dtype=np.float64)
    Vr = np.empty_like(V)
    _check(lib().laph_reset_phase(n, nev, _p(V), _p(want), _p(Vr)), "laph_reset_phase")
    return Vr


def check_trace(V):
    """read_write.cpp:20-35 -> (ok, tr(V^H V), sum(V^H V))."""
    V = _c128(V)
    nev, n = V.shape
    tr = np.zeros(2)
    sm = np.zeros(2)
    rc = lib().laph_check_trace(n, nev, _p(V), _p(tr), _p(sm))
    if rc not in (0, -7):
        _check(rc, "laph_check_trace")
    return rc == 0, complex(tr[0], tr[1]), complex(sm[0], sm[1])


def write_eig_sys_bin(result_path, prefix, config, t, V):
    """read_write.cpp:176-201 (column-major n x nev complex128)."""
    V = _c128(V)
    nev, n = V.shape
    _check(lib().laph_write_eig_sys_bin(os.fsencode(result_path), prefix.encode(), config, t, n, nev, _p(V)),
           "laph_write_eig_sys_bin")


def write_eigenvalues_bin(result_path, prefix, config, t, values):
    """read_write.cpp:237-250."""
    values = np.ascontiguousarray(values, dtype=np.float64)
    _check(lib().laph_write_eigenvalues_bin(os.fsencode(result_path), prefix.encode(), config, t,
                                            len(values), _p(values)), "laph_write_eigenvalues_bin")


def read_evectors_bin_ts(result_path, prefix, config, t, n, nev):
    """read_write.cpp:40-96 -> V (nev, n), row i = eigenvector i; LaphError on
    a missing or short file and when check_trace fails."""
    V = np.empty((nev, n), dtype=np.complex128)
    _check(lib().laph_read_evectors_bin_ts(os.fsencode(result_path), prefix.encode(), config, t, n, nev, _p(V)),
           "laph_read_evectors_bin_ts")
    return V


def read_eigenvalues_bin(result_path, prefix, config, t, nev):
    """read_write.cpp:137-152."""
    out = np.empty(nev)
    _check(lib().laph_read_eigenvalues_bin(os.fsencode(result_path), prefix.encode(), config, t, nev, _p(out)),
           "laph_read_eigenvalues_bin")
    return out


def read_eigenvalues_ascii(prefix, config, t, nev):
    """read_write.cpp:98-115: "<prefix>.%03d.%04d" % (t, config)."""
    out = np.empty(nev)
    _check(lib().laph_read_eigenvalues_ascii(os.fsencode(prefix), config, t, nev, _p(out)),
           "laph_read_eigenvalues_ascii")
    return out


def source_shape(L, V, device=0):
    """sourceshape_funcs.cpp:31-81 on the device: psi for r in range(0, L, 2)^3
    (r3 fastest) of eigenvectors V (nev, n)."""
    V = _c128(V)
    nev = V.shape[0]
    assert V.shape[1] == 3 * L ** 3
    half = (L + 1) // 2
    psi = np.empty(half ** 3)
    _check(lib().laph_source_shape(L, nev, _p(V), device, _p(psi)), "laph_source_shape")
    return psi


# ---------------------------------------------------------------------------
class Tslice:
    """One timeslice on the device: links, HYP, the MatShell operator and the
    plain FP64 Krylov-Schur solve (the reference's Tslice singleton plus the
    MatShell context of ev_ts.cpp:197)."""

    def __init__(self, L, deg, lambda_l, lambda_c, device=0, ts_config=None):
        self.L, self.deg, self.lambda_l, self.lambda_c = L, deg, lambda_l, lambda_c
        self.n = 3 * L ** 3
        self.params = laph_params(L, deg, lambda_l, lambda_c, device)
        self.ctx = ctypes.c_void_p()
        cfg = None if ts_config is None else np.ascontiguousarray(ts_config, dtype=np.float64)
        _check(lib().laph_ctx_create(ctypes.byref(self.ctx), ctypes.byref(self.params),
                                     None if cfg is None else _p(cfg)), "laph_ctx_create")

    def close(self):
        if self.ctx:
            lib().laph_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def map_timeslice_to_eigen(self, ts_config):
        cfg = np.ascontiguousarray(ts_config, dtype=np.float64)
        _check(lib().laph_map_timeslice(self.ctx, _p(cfg)), "laph_map_timeslice")

    def set_links(self, U):
        U = _c128(U).reshape(-1)
        assert U.size == self.n * 9
        _check(lib().laph_set_links(self.ctx, _p(U)), "laph_set_links")

    def get_links(self):
        U = np.empty((self.L ** 3, 3, 3, 3), dtype=np.complex128)
        _check(lib().laph_get_links(self.ctx, _p(U)), "laph_get_links")
        return U

    def smearing_hyp(self, alpha_1, alpha_2, iter):
        _check(lib().laph_hyp_smear(self.ctx, alpha_1, alpha_2, iter), "laph_hyp_smear")

    def MatMult_Laplacian2D(self, x):
        x = _c128(x)
        assert x.size == self.n
        y = np.empty_like(x)
        _check(lib().laph_shell_mult(self.ctx, _p(x), _p(y)), "laph_shell_mult")
        return y

    def MatGetDiagonal_Laplacian2D(self):
        d = np.empty(self.n, dtype=np.complex128)
        _check(lib().laph_shell_get_diagonal(self.ctx, _p(d)), "laph_shell_get_diagonal")
        return d

    def laplacian(self, x):
        x = _c128(x)
        y = np.empty_like(x)
        _check(lib().laph_laplacian(self.ctx, _p(x), _p(y)), "laph_laplacian")
        return y

    def eps_solve(self, nev, ncv=0, tol=1e-10, max_it=0, seed=1227, want_vectors=True):
        """EPS (HEP, LARGEST_MAGNITUDE) on T_DEG(B) -> (theta, V (nev, n) or None, stats)."""
        opts = laph_eps_opts(nev, ncv, tol, max_it, seed)
        theta = np.empty(nev)
        V = np.empty((nev, self.n), dtype=np.complex128) if want_vectors else None
        st = laph_eps_stats()
        _check(lib().laph_eps_solve(self.ctx, ctypes.byref(opts), _p(theta),
                                    None if V is None else _p(V), ctypes.byref(st)), "laph_eps_solve")
        return theta, V, st


class Batch:
    """All timeslices of a configuration in one mixed-precision batched solve."""

    def __init__(self, L, deg, lambda_l, lambda_c, nts, nev, nsub=0, tol=1e-10, max_iter=0, seed=1227,
                 device=0, warm_start=0, filter_passes=0, filter_storage=0):
        self.L, self.nts, self.nev = L, nts, nev
        self.n = 3 * L ** 3
        self.params = laph_params(L, deg, lambda_l, lambda_c, device)
        self.opts = laph_batch_opts(nts, nev, nsub, tol, max_iter, seed, int(warm_start), int(filter_passes),
                                    int(filter_storage))
        self.h = ctypes.c_void_p()
        _check(lib().laph_batch_create(ctypes.byref(self.h), ctypes.byref(self.params),
                                       ctypes.byref(self.opts)), "laph_batch_create")

    def close(self):
        if self.h:
            lib().laph_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timeslice(self, ts, ts_config, alpha_1, alpha_2, iter):
        cfg = np.ascontiguousarray(ts_config, dtype=np.float64)
        _check(lib().laph_batch_set_timeslice(self.h, ts, _p(cfg), alpha_1, alpha_2, iter),
               "laph_batch_set_timeslice")

    def set_links(self, ts, U):
        U = _c128(U).reshape(-1)
        _check(lib().laph_batch_set_links(self.h, ts, _p(U)), "laph_batch_set_links")

    def solve(self):
        st = laph_batch_stats()
        _check(lib().laph_batch_solve(self.h, ctypes.byref(st)), "laph_batch_solve")
        return st

    def time_filter(self, reps=20):
        """(ms per filter-step launch, algorithmic bytes per launch), kernel alone on the device."""
        ms, by = ctypes.c_double(), ctypes.c_double()
        _check(lib().laph_batch_time_filter(self.h, reps, ctypes.byref(ms), ctypes.byref(by)), "laph_batch_time_filter")
        return ms.value, by.value

    def time_gram(self, reps=10):
        """(ms per Gram, flops it executes): the FP64 matrix-core product alone on the device."""
        ms, fl = ctypes.c_double(), ctypes.c_double()
        _check(lib().laph_batch_time_gram(self.h, reps, ctypes.byref(ms), ctypes.byref(fl)), "laph_batch_time_gram")
        return ms.value, fl.value

    def write(self, ts, result_path, config, t):
        """eigenvectors / eigenvalues / phases files of timeslice ts (file index t)
        straight from the device -> (tr(V^H V), sum(V^H V))."""
        ts_ = np.zeros(4)
        _check(lib().laph_batch_write(self.h, ts, os.fsencode(result_path), config, t, _p(ts_)), "laph_batch_write")
        return complex(ts_[0], ts_[1]), complex(ts_[2], ts_[3])

    def source_shape(self, ts):
        """psi(r) of timeslice ts of the solved batch, r in range(0, L, 2)^3."""
        half = (self.L + 1) // 2
        psi = np.empty(half ** 3)
        _check(lib().laph_batch_source_shape(self.h, ts, _p(psi)), "laph_batch_source_shape")
        return psi

    def get_device(self, ts, d_V):
        """Phase-fixed eigenvectors of timeslice ts (file layout, n x nev) into
        the device buffer at address d_V; returns the phases."""
        ph = np.empty(self.nev)
        _check(lib().laph_batch_get_device(self.h, ts, ctypes.c_void_p(d_V), _p(ph)), "laph_batch_get_device")
        return ph

    def get(self, ts, want_vectors=True, out=None):
        """Eigenvalues, phase-fixed eigenvectors (nev x n, each eigenvector
        contiguous; written into `out` when given), phases and residuals of
        timeslice ts."""
        ev = np.empty(self.nev)
        ph = np.empty(self.nev)
        res = np.empty(self.nev)
        if out is not None:
            assert out.shape == (self.nev, self.n) and out.dtype == np.complex128 and out.flags["C_CONTIGUOUS"]
        V = out if out is not None else (np.empty((self.nev, self.n), dtype=np.complex128) if want_vectors else None)
        _check(lib().laph_batch_get(self.h, ts, _p(ev), None if V is None else _p(V), _p(ph), _p(res)),
               "laph_batch_get")
        return ev, V, ph, res


def _fp(a):
    assert a.flags["C_CONTIGUOUS"] and a.dtype == np.float32
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Net:
    """Gauge-covariant warm-start network (16 complex channels, BF16
    activations): proposes the start block of a Batch."""

    def __init__(self, layers=4, seed=1, device=0):
        self.layers = layers
        self.h = ctypes.c_void_p()
        _check(lib().laph_net_create(ctypes.byref(self.h), layers, seed, device), "laph_net_create")

    def close(self):
        if self.h:
            lib().laph_net_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def num_params(self):
        return lib().laph_net_num_params(self.h)

    def get_params(self):
        p = np.empty(self.num_params, dtype=np.float32)
        _check(lib().laph_net_get_params(self.h, _fp(p)), "laph_net_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.num_params
        _check(lib().laph_net_set_params(self.h, _fp(p)), "laph_net_set_params")

    def propose(self, batch, seed):
        _check(lib().laph_net_propose(self.h, batch.h, seed), "laph_net_propose")

    def train_step(self, batch, seed, lr):
        loss = ctypes.c_double()
        _check(lib().laph_net_train_step(self.h, batch.h, seed, lr, ctypes.byref(loss)), "laph_net_train_step")
        return loss.value

    def loss_grad(self, L, links, x, V, fp32_act=True):
        """links (V3,3,3,3) complex; x (n, K) complex; V (n, nev) complex."""
        links = _c128(links).reshape(-1)
        n, K = x.shape
        xin = np.ascontiguousarray(x, dtype=np.complex64)
        xf = xin.view(np.float32)
        Vc = _c128(V)
        nev = Vc.shape[1]
        loss = ctypes.c_double()
        y = np.empty((n, K), dtype=np.complex128)
        grad = np.empty(self.num_params, dtype=np.float32)
        _check(lib().laph_net_loss_grad(self.h, L, K, int(bool(fp32_act)), _p(links), _fp(xf.reshape(-1)), _p(Vc),
                                        nev, ctypes.byref(loss), _p(y), _fp(grad)), "laph_net_loss_grad")
        return loss.value, y, grad
