"""The reference-side binding of the MatShell callbacks (signatures of
modules/shell_matop.h:25-26) compiled as C11 with gcc against
include/laph.h, over stand-ins for the PETSc handles
(tests/binding/matshell_binding.c).

CPU: the two callback bodies in the C file are the ones the repository's
integration notes show, the file compiles warning-free as C11 and links against
liblapheig.so.  GPU: MatMult_Laplacian2D, called through the mock (Mat, Vec,
Vec) handles as EPSSolve's MatMult would, equals the oracle's tv_iter
(T_DEG(B) x, shell_matop.cpp:136-168) to 1e-13 relative, and
MatGetDiagonal_Laplacian2D returns -6 (shell_matop.cpp:253)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "binding", "matshell_binding.c")
LIB = os.path.join(HERE, "binding", "libmatshell_binding.so")


def _bodies(text):
    out = {}
    for name in ("MatMult_Laplacian2D", "MatGetDiagonal_Laplacian2D"):
        m = re.search(r"PetscErrorCode " + name + r"\(.*?\n}\n", text, re.S)
        assert m, name
        out[name] = [ln.split("//")[0].split("/*")[0].rstrip() for ln in m.group(0).splitlines()]
    return out


def test_binding_matches_integration_notes():
    doc = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    assert _bodies(doc) == _bodies(open(SRC).read())


def test_binding_compiles_as_c11(tmp_path):
    out = tmp_path / "libb.so"
    r = subprocess.run(["gcc", "-std=c11", "-O2", "-fPIC", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-I" + os.path.join(ROOT, "include"), "-shared", "-o", str(out), SRC,
                        "-L" + os.path.join(ROOT, "laph_eigsys_amd"), "-llapheig",
                        "-Wl,-rpath," + os.path.join(ROOT, "laph_eigsys_amd")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lib = ctypes.CDLL(str(out))
    for s in ("MatMult_Laplacian2D", "MatGetDiagonal_Laplacian2D", "binding_mat_mult", "binding_get_diagonal"):
        assert hasattr(lib, s), s


@pytest.mark.gpu
def test_binding_mat_mult_against_oracle(gpu):
    from laph_eigsys_amd import lapheig as lp
    from oracle import lapheig_np as N
    from oracle import oracle as O
    assert os.path.exists(LIB), "build the binding first (__graft_entry__.build)"
    b = ctypes.CDLL(LIB)
    b.binding_mat_mult.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    b.binding_get_diagonal.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L, deg, lam_c = 12, 16, 0.7
    cfg = N.synth_config(1227, 0.4, L, 1)
    U = O.hyp(O.map_timeslice(cfg, L), L, 0.62, 0.62, 3)
    n = 3 * L ** 3
    x = np.random.default_rng(12).standard_normal(n) + 1j * np.random.default_rng(13).standard_normal(n)
    y = np.empty(n, dtype=np.complex128)
    d = np.empty(n, dtype=np.complex128)
    with lp.Tslice(L, deg, 12.0, lam_c, device=gpu, ts_config=cfg) as t:
        t.smearing_hyp(0.62, 0.62, 3)
        assert b.binding_mat_mult(t.ctx, x.ctypes.data, y.ctypes.data) == 0
        assert b.binding_get_diagonal(t.ctx, d.ctypes.data) == 0
    ref = O.tv_iter(U, L, 12.0, lam_c, deg, x)
    err = abs(y - ref).max() / abs(ref).max()
    print(f"MatMult_Laplacian2D through the C binding: rel err {err:.2e}")
    assert err < 1e-13
    assert np.all(d == -6.0)
