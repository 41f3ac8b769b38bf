"""Host logic of the per-configuration driver (laph_eigsys_amd/ev_ts.py):
the parameter file of par_io.cpp:8-133 and the grouping of timeslices into
batched solves that fit the device memory."""
from laph_eigsys_amd import ev_ts


PARAMS = """# comment line
config_path = /data/confs
result_path = ./out
phase_path = ./old
#
LT = 48
LX = 24
NEV = 120
alpha_1 = 0.62
alpha_2 = 0.62
iter = 3
LAM_L = 12.0
LAM_C = 0.9
DEG = 12
OMP_THRDS = 2
"""


def test_parameter_file(tmp_path):
    f = tmp_path / "parameters.txt"
    f.write_text(PARAMS)
    p = ev_ts.IO().set_values(str(f))
    assert p.get_int("LT") == 48 and p.get_int("LX") == 24 and p.get_int("LZ") == 24
    assert p.get_int("V3") == 24 ** 3 and p.get_int("MAT_ENTRIES") == 3 * 24 ** 3
    assert p.get_int("V_TS") == 72 * 24 ** 3 and p.get_int("V4_LIME") == 48 * 72 * 24 ** 3
    assert p.get_int("NEV") == 120 and p.get_int("DEG") == 12 and p.get_int("iter") == 3
    assert p.get_float("alpha_1") == 0.62 and p.get_float("LAM_C") == 0.9
    assert p.get_path("conf") == "/data/confs" and p.get_path("res") == "./out"
    assert p.get_path("phase_path") == "./old"


def test_tune_file_default_is_in_the_result_directory(tmp_path):
    """The layout cache of a run: --tune-file wins ('none' turns it off),
    then LAPH_TUNE_FILE, else laph_layouts.txt beside the results."""
    f = tmp_path / "parameters.txt"
    f.write_text(PARAMS)
    p = ev_ts.IO().set_values(str(f))
    assert ev_ts.tune_file_path("", p, {}) == "./out/laph_layouts.txt"
    assert ev_ts.tune_file_path("", p, {"LAPH_TUNE_FILE": "/x/t.txt"}) == "/x/t.txt"
    assert ev_ts.tune_file_path("/y/u.txt", p, {"LAPH_TUNE_FILE": "/x/t.txt"}) == "/y/u.txt"
    assert ev_ts.tune_file_path("none", p, {"LAPH_TUNE_FILE": "/x/t.txt"}) is None


def test_groups_cover_range_and_fit():
    # 48^3 x 96, nev 200: 6.8 GB per timeslice -> 4 groups of 24 under 200 GB
    g = ev_ts.plan_groups(96, 48, 200, 200.0)
    assert [n for _, n in g] == [24, 24, 24, 24]
    assert ev_ts.batch_bytes_per_timeslice(48, 200) * 24 < 200e9
    # 32^3 x 64, nev 220, search space 256: one group
    assert ev_ts.plan_groups(64, 32, 220, 200.0, nsub=256) == [(0, 64)]
    # explicit group size, ragged tail
    assert ev_ts.plan_groups(10, 8, 20, 1.0, group=4) == [(0, 4), (4, 4), (8, 2)]
    assert ev_ts.plan_groups(0, 8, 20, 1.0) == []
    for todo in range(1, 40):
        g = ev_ts.plan_groups(todo, 24, 120, 5.0)
        assert sum(n for _, n in g) == todo and all(a == sum(n for _, n in g[:i]) for i, (a, _) in enumerate(g))


def test_set_phase_resets_to_stored_phases(tmp_path):
    """set_phase (set_phase.cpp:13-58): stored eigenvectors come back with
    arg V(0, n) = the stored old phase n, otherwise unchanged up to that
    phase; the old phases are written beside them."""
    import numpy as np
    from laph_eigsys_amd import lapheig as lp
    from laph_eigsys_amd import set_phase
    L, nev, conf, t = 3, 5, 1300, 7
    n = 3 * L ** 3
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((n, nev)) + 1j * rng.standard_normal((n, nev)))
    V = Q.T.copy()
    old = rng.uniform(-np.pi, np.pi, nev)
    for d in ("in", "old", "out"):
        (tmp_path / d).mkdir()
    lp.write_eig_sys_bin(str(tmp_path / "in"), "eigenvectors", conf, t, V)
    lp.write_eigenvalues_bin(str(tmp_path / "old"), "phases", conf, t, old)
    f = tmp_path / "p.txt"
    f.write_text(f"in_path = {tmp_path / 'in'}\nphase_path = {tmp_path / 'old'}\nresult_path = {tmp_path / 'out'}\n"
                 f"LT = 8\nLX = {L}\nNEV = {nev}\n")
    set_phase.main(["--conf", str(conf), "--t", str(t), "--params", str(f)])
    Vr = lp.read_evectors_bin_ts(str(tmp_path / "out"), "eigenvectors_reset", conf, t, n, nev)
    np.testing.assert_allclose(np.angle(Vr[:, 0]), old, atol=1e-12)
    ratio = Vr / V
    np.testing.assert_allclose(np.abs(ratio), 1.0, atol=1e-12)
    np.testing.assert_allclose(ratio, ratio[:, :1] * np.ones((1, n)), atol=1e-12)
    np.testing.assert_array_equal(lp.read_eigenvalues_bin(str(tmp_path / "out"), "eigenvectors_old_phase",
                                                          conf, t, nev), old)


def test_alpha_grid_follows_the_accumulating_loop():
    """tune_hyp.cpp:116-117: a += 0.05 in double precision, a <= upper bound."""
    from laph_eigsys_amd.tune_hyp import alpha_grid
    g = alpha_grid(0.3, 0.7)
    a, ref = 0.3, []
    while a <= 0.7:
        ref.append(a)
        a += 0.05
    assert g == ref
    assert len(g) == 8 and abs(g[-1] - 0.65) < 1e-12  # 0.3 + 8 * 0.05 rounds above 0.7
    assert alpha_grid(0.62, 0.62) == [0.62]
    assert alpha_grid(0.5, 0.4) == []


def test_sourceshape_host_logic(tmp_path, monkeypatch):
    """Loop order, radius averaging and the two writers of the source-shape
    program (read_write.cpp:206-235) against the oracle's restatement."""
    import struct
    import numpy as np
    from laph_eigsys_amd import sourceshape
    from oracle import lapheig_np as N
    L = 4
    psi = np.arange(8, dtype=float) + 0.5
    res = sourceshape.results(L, psi)
    assert [r for r, _, _ in res] == [(a, b, c) for a in (0, 2) for b in (0, 2) for c in (0, 2)]
    avg = sourceshape.avg_radii(res)
    assert avg == N.avg_radii(res)
    assert avg[0] == (0.0, 0.5) and avg[1] == (2.0, (1.5 + 2.5 + 4.5) / 3)
    monkeypatch.chdir(tmp_path)
    fa = sourceshape.write_sourceshape_ascii("sourceshape", 1300, 7, 120, avg)
    assert fa == "sourceshape_nev120.1300.7"
    lines = open(fa).read().splitlines()
    assert lines[0] == "#r\tpsi(r)" and lines[1] == "0 0.5" and len(lines) == 1 + len(avg)
    fb = sourceshape.write_sourceshape_bin("sourceshape", 1300, 7, 120, avg)
    assert fb == "sourceshape_nev120.1300.007.bin"
    raw = open(fb, "rb").read()
    assert len(raw) == 16 * len(avg) and struct.unpack("<dd", raw[16:32]) == avg[1]
