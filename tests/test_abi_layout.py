"""The ctypes mirrors of the C ABI (laph_eigsys_amd/lapheig.py) against
include/laph.h: every struct's size and field offsets as a C compiler lays
them out, and every entry point's argument kinds.  The library writes whole structs into the
caller's memory (laph_batch_solve fills laph_batch_stats), so a field added
on one side only would write past the Python buffer.  No GPU: the C side is
a host program compiled from the header alone."""
import ctypes
import os
import shutil
import subprocess

import pytest

from laph_eigsys_amd import lapheig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = ("laph_params", "laph_eps_opts", "laph_eps_stats", "laph_batch_opts", "laph_batch_stats")


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "laph.h"', "int main(void) {"]
    for s in STRUCTS:
        lines.append(f'  printf("{s} size %zu\\n", sizeof({s}));')
        for name, _ in getattr(lapheig, s)._fields_:
            lines.append(f'  printf("{s} {name} %zu\\n", offsetof({s}, {name}));')
    lines += ["  return 0;", "}"]
    src = d / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = d / "layout"
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    layout = {}
    for line in out.splitlines():
        s, key, val = line.split()
        layout[(s, key)] = int(val)
    return layout


@pytest.mark.parametrize("name", STRUCTS)
def test_ctypes_mirror_matches_the_header(c_layout, name):
    cls = getattr(lapheig, name)
    assert ctypes.sizeof(cls) == c_layout[(name, "size")], name
    for field, _ in cls._fields_:
        assert getattr(cls, field).offset == c_layout[(name, field)], (name, field)


def test_header_fields_all_mirrored():
    """Every member the header declares in these structs has a ctypes field
    (a member missing on the Python side would shift nothing, but its value
    would be unreadable and the sizes above would differ)."""
    import re
    hdr = open(os.path.join(ROOT, "include", "laph.h")).read()
    for s in STRUCTS:
        body = re.search(r"typedef struct " + s + r" \{(.*?)\} " + s + ";", hdr, re.S).group(1)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        members = re.findall(r"\b(\w+)\s*;", body)
        assert members == [f for f, _ in getattr(lapheig, s)._fields_], s


def _kind_c(decl):
    decl = " ".join(decl.replace("const ", "").split())
    if "*" in decl:
        return "ptr"
    base = decl.rsplit(" ", 1)[0]
    return {"int": "int", "double": "double", "unsigned long long": "u64"}[base]


def _kind_ctypes(t):
    if t in (ctypes.c_int,):
        return "int"
    if t in (ctypes.c_double,):
        return "double"
    if t in (ctypes.c_ulonglong,):
        return "u64"
    if t in (ctypes.c_void_p, ctypes.c_char_p) or issubclass(t, ctypes._Pointer):
        return "ptr"
    raise AssertionError(t)


def test_ctypes_signatures_match_the_prototypes():
    """Every prototype of include/laph.h against lapheig.SIGNATURES: the same
    entry points, the same number of arguments and, per argument, the same
    kind (int, double, 64-bit unsigned, pointer); int or string results."""
    import re
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "laph.h")).read(), flags=re.S)
    protos = re.findall(r"\b(int|const char\s*\*)\s+(laph_\w+)\s*\(([^)]*)\)\s*;", hdr)
    assert len(protos) >= 40
    assert {n for _, n, _ in protos} == set(lapheig.SIGNATURES)
    for ret, name, args in protos:
        argtypes, restype = lapheig.SIGNATURES[name]
        c_args = [] if args.strip() == "void" else [_kind_c(a) for a in args.split(",")]
        assert c_args == [_kind_ctypes(t) for t in argtypes], name
        assert restype is (ctypes.c_char_p if "char" in ret else ctypes.c_int), name
