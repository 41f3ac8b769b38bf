"""CPU tests of liblapheig.so: every symbol declared in include/laph.h is
exported, and the host-side entry points (synthetic field, LIME ingest and
egress, spectrum recovery, phase fixing, trace guard, writers) agree with the
oracle.  No device compute is called here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import lapheig_np as N
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    from laph_eigsys_amd import lapheig
    return lapheig


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "laph.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(laph_\w+)\s*\(", txt, re.M)))


def test_all_declared_symbols_exported(L):
    names = declared_symbols()
    assert len(names) >= 25
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [s for s in names if not hasattr(lib, s)]
    assert not missing, missing
    assert set(names) == set(L.SIGNATURES), set(names) ^ set(L.SIGNATURES)


def test_error_strings(L):
    lib = L.lib()
    assert lib.laph_error_string(0) == b"ok"
    assert b"LIME" in lib.laph_error_string(-5)
    assert lib.laph_version().startswith(b"lapheig")


def test_synth_matches_numpy_definition(L):
    got = L.synth_config(1227, 0.4, 3, 2)
    ref = N.synth_config(1227, 0.4, 3, 2)
    np.testing.assert_allclose(got, ref, atol=2e-14)
    assert not np.allclose(L.synth_config(1228, 0.4, 3, 2), got)


def test_lime_write_read_round_trip(L, tmp_path):
    T, Ls = 3, 4
    cfg = L.synth_config(5, 0.3, Ls, T)
    p = str(tmp_path / "conf.0001")
    L.write_lime(p, cfg, T, Ls)
    # product reader and oracle reader agree with the source, slices [1, 2]
    np.testing.assert_array_equal(L.read_lime_gauge_field_doubleprec_timeslices(p, T, Ls, 1, 2),
                                  cfg[Ls ** 3 * 72:])
    np.testing.assert_array_equal(O.read_lime(p, T, Ls, 0, 2), cfg)
    # numpy writer -> product reader
    q = str(tmp_path / "np.lime")
    N.write_lime(q, cfg, T, Ls)
    np.testing.assert_array_equal(L.read_lime_gauge_field_doubleprec_timeslices(q, T, Ls, 0, 0),
                                  cfg[:Ls ** 3 * 72])


def test_lime_errors(L, tmp_path):
    p = str(tmp_path / "c.lime")
    L.write_lime(p, L.synth_config(1, 0.3, 3, 2), 2, 3)
    with pytest.raises(L.LaphError) as e:
        L.read_lime_gauge_field_doubleprec_timeslices(p, 2, 4, 0, 0)
    assert e.value.code == -5
    with pytest.raises(L.LaphError) as e:
        L.read_lime_gauge_field_doubleprec_timeslices(str(tmp_path / "nope"), 2, 3, 0, 0)
    assert e.value.code == -4
    with pytest.raises(L.LaphError):
        L.read_lime_gauge_field_doubleprec_timeslices(p, 2, 3, 1, 0)
    trunc = str(tmp_path / "t.lime")
    open(trunc, "wb").write(open(p, "rb").read()[:-100])
    with pytest.raises(L.LaphError) as e:
        L.read_lime_gauge_field_doubleprec_timeslices(trunc, 2, 3, 0, 1)
    assert e.value.code == -4


def test_post_processing_matches_oracle(L):
    rng = np.random.default_rng(3)
    n, nev = 96, 9
    Q, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    V = Q.T.copy()
    Vf, ph = L.fix_phase(V)
    Vf0, ph0 = O.fix_phase(V)
    np.testing.assert_array_equal(ph, ph0)
    np.testing.assert_allclose(Vf, Vf0, atol=1e-15)
    np.testing.assert_allclose(L.reset_phase(Vf, ph), O.reset_phase(Vf0, ph0), atol=1e-15)
    ok, tr, sm = L.check_trace(Vf)
    ok0, tr0, sm0 = O.check_trace(Vf0)
    assert ok and ok0 and abs(tr - tr0) < 1e-12 and abs(sm - sm0) < 1e-12
    assert not L.check_trace(1.01 * Vf)[0]
    theta = np.cosh(12 * np.arccosh(np.linspace(1.0, 1.05, 11)))
    np.testing.assert_allclose(L.recover_spectrum(theta, 12, 12.0, 0.6),
                               O.recover_spectrum(theta, 12, 12.0, 0.6), rtol=0, atol=1e-15)


def test_writers_byte_format(L, tmp_path):
    rng = np.random.default_rng(4)
    n, nev = 24, 3
    Q, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    V = Q.T.copy()
    L.write_eig_sys_bin(str(tmp_path), "eigenvectors", 1300, 7, V)
    raw = np.fromfile(tmp_path / "eigenvectors.1300.007", dtype="<c16")
    np.testing.assert_array_equal(raw, V.reshape(-1))  # column-major n x nev
    ev = np.array([-0.1, -0.2, -0.3])
    L.write_eigenvalues_bin(str(tmp_path), "eigenvalues", 1300, 7, ev)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "eigenvalues.1300.007", dtype="<f8"), ev)
    with pytest.raises(L.LaphError) as e:
        L.write_eig_sys_bin(str(tmp_path), "bad", 1, 1, 2 * V)  # read_write.cpp:183-186 guard
    assert e.value.code == -7
    assert not os.path.exists(tmp_path / "bad.0001.001")


def test_readers_round_trip_and_errors(L, tmp_path):
    """read_evectors_bin_ts / read_eigenvalues_bin / read_eigenvalues_ascii
    (read_write.cpp:40-152): what the writers wrote comes back bit-exact; a
    missing file, a short file and a damaged basis are errors (the reference
    exits on the last two)."""
    rng = np.random.default_rng(9)
    n, nev = 30, 4
    Q, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    V = Q.T.copy()
    d = str(tmp_path)
    L.write_eig_sys_bin(d, "eigenvectors", 1300, 5, V)
    np.testing.assert_array_equal(L.read_evectors_bin_ts(d, "eigenvectors", 1300, 5, n, nev), V)
    ev = -np.arange(1, nev + 1) / 7.0
    L.write_eigenvalues_bin(d, "eigenvalues", 1300, 5, ev)
    np.testing.assert_array_equal(L.read_eigenvalues_bin(d, "eigenvalues", 1300, 5, nev), ev)
    with pytest.raises(L.LaphError) as e:
        L.read_eigenvalues_bin(d, "eigenvalues", 1300, 6, nev)
    assert e.value.code == -4
    with pytest.raises(L.LaphError) as e:
        L.read_eigenvalues_bin(d, "eigenvalues", 1300, 5, nev + 1)
    assert e.value.code == -5
    with pytest.raises(L.LaphError) as e:
        L.read_evectors_bin_ts(d, "eigenvectors", 1300, 5, n, nev + 1)  # eof_check
    assert e.value.code == -5
    (2 * V).astype("<c16").tofile(tmp_path / "eigenvectors.1300.006")  # not orthonormal
    with pytest.raises(L.LaphError) as e:
        L.read_evectors_bin_ts(d, "eigenvectors", 1300, 6, n, nev)
    assert e.value.code == -7
    with open(tmp_path / "ev.005.1300", "w") as f:  # timeslice before config
        f.write("\n".join(repr(float(x)) for x in ev) + "\n")
    np.testing.assert_array_equal(L.read_eigenvalues_ascii(str(tmp_path / "ev"), 1300, 5, nev), ev)
    with pytest.raises(L.LaphError) as e:
        L.read_eigenvalues_ascii(str(tmp_path / "ev"), 1300, 5, nev + 1)
    assert e.value.code == -5


@pytest.mark.parametrize("kw", [dict(L=2), dict(deg=1), dict(lambda_l=0.0), dict(nts=0), dict(nev=0),
                                dict(tol=0.0), dict(filter_passes=7), dict(filter_storage=2),
                                dict(L=3, nev=20, nsub=128), dict(nev=64, nsub=64)])
def test_batch_rejects_bad_arguments(L, kw):
    """laph_batch_create validates its arguments before touching a device:
    each bad one is LAPH_ERR_INVALID with a message (no GPU needed)."""
    a = dict(L=6, deg=8, lambda_l=12.0, nts=2, nev=10, nsub=0, tol=1e-10, filter_passes=0, filter_storage=0)
    a.update(kw)
    with pytest.raises(L.LaphError) as e:
        L.Batch(a["L"], a["deg"], a["lambda_l"], 0.0, a["nts"], a["nev"], nsub=a["nsub"], tol=a["tol"],
                filter_passes=a["filter_passes"], filter_storage=a["filter_storage"])
    assert e.value.code == -1 and "laph_batch_create" in str(e.value)


@pytest.mark.parametrize("value", ["1", "2", "3", "300"])
def test_batch_rejects_filter_degree_below_4(L, monkeypatch, value):
    """LAPH_FILTER_DEGREE must be 0 or a highest per-cycle degree in 4..256
    (the per-cycle degree is a multiple of 4, at least 4, and the root and
    coefficient tables are sized by it): other values are rejected before
    any device work."""
    monkeypatch.setenv("LAPH_FILTER_DEGREE", value)
    with pytest.raises(L.LaphError) as e:
        L.Batch(8, 8, 12.0, 0.0, 2, 10)
    assert e.value.code == -1 and "LAPH_FILTER_DEGREE" in str(e.value)
