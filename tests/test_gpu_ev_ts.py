"""The per-configuration driver end to end on the GPU: LIME file -> HYP ->
plain drop-in and batched solver -> phase-fixed eigensystem files, checked
against the dense diagonalisation of the oracle's smeared Laplacian."""
import numpy as np
import pytest

from laph_eigsys_amd import ev_ts
from oracle import lapheig_np as N
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_driver_files(gpu, tmp_path):
    from laph_eigsys_amd import lapheig as lp
    L, LT, nev, deg, conf = 6, 3, 16, 8, 1227
    cfg = N.synth_config(77, 0.4, L, LT)
    (tmp_path / "confs").mkdir()
    lp.write_lime(str(tmp_path / "confs" / f"conf.{conf:04d}"), cfg, LT, L)
    dense = []
    for t in range(LT):
        U = O.hyp(O.map_timeslice(cfg[t * 72 * L ** 3:(t + 1) * 72 * L ** 3], L), L, 0.62, 0.62, 3)
        D = N.dense_D(U.reshape(L, L, L, 3, 3, 3))
        dense.append((D, np.sort(np.linalg.eigvalsh(D))[::-1][:nev]))
    lam_c = 1.3 * max(-d[1][-1] for d in dense)
    results = {}
    for solver in ("plain", "batched"):
        out = tmp_path / solver
        (tmp_path / f"{solver}.txt").write_text(
            f"config_path = {tmp_path / 'confs'}\nresult_path = {out}\nLT = {LT}\nLX = {L}\nNEV = {nev}\n"
            f"alpha_1 = 0.62\nalpha_2 = 0.62\niter = 3\nLAM_L = 12.0\nLAM_C = {lam_c}\nDEG = {deg}\n")
        pars = ev_ts.IO().set_values(str(tmp_path / f"{solver}.txt"))
        tol = 1e-12 if solver == "plain" else 1e-10
        info = ev_ts.run(pars, conf, solver=solver, tol=tol, device=gpu, log=lambda m: None)
        assert info["timeslices"] == LT
        for t in range(LT):
            ev = np.fromfile(out / f"eigenvalues.{conf:04d}.{t:03d}")
            ph = np.fromfile(out / f"phases.{conf:04d}.{t:03d}")
            V = np.fromfile(out / f"eigenvectors.{conf:04d}.{t:03d}", dtype=np.complex128)
            assert ev.shape == (nev,) and ph.shape == (nev,) and V.size == nev * 3 * L ** 3
            V = V.reshape(nev, 3 * L ** 3)  # column-major n x nev: each eigenvector contiguous
            D, w = dense[t]
            np.testing.assert_allclose(ev, w, rtol=1e-10, atol=1e-12)
            R = max(np.linalg.norm(D @ V[i] - ev[i] * V[i]) for i in range(nev))
            assert R < 1e-9, (solver, t, R)
            assert np.all(np.abs(V[:, 0].imag) < 1e-13) and np.all(V[:, 0].real >= 0)
            np.testing.assert_allclose(V.conj() @ V.T, np.eye(nev), atol=1e-11)
            results[(solver, t)] = ev
    for t in range(LT):
        np.testing.assert_allclose(results[("plain", t)], results[("batched", t)], rtol=1e-10, atol=1e-12)


def test_driver_layout_cache_in_result_directory(gpu, tmp_path, monkeypatch):
    """The per-configuration program keeps the batched solver's layouts and
    warm-up degree in <result_path>/laph_layouts.txt by default: a second
    process-like run of the same shape reads them instead of timing, and
    writes the same eigensystem files bit for bit."""
    import json
    from laph_eigsys_amd import lapheig as lp
    monkeypatch.setenv("LAPH_TUNE_FILE", "unset-below")  # restored after the test
    monkeypatch.delenv("LAPH_TUNE_FILE")
    L, LT, nev, deg, conf = 8, 2, 20, 8, 31
    (tmp_path / "confs").mkdir()
    lp.write_lime(str(tmp_path / "confs" / f"conf.{conf:04d}"), N.synth_config(5, 0.4, L, LT), LT, L)
    out = tmp_path / "res"
    pf = tmp_path / "parameters.txt"
    pf.write_text(f"config_path = {tmp_path / 'confs'}\nresult_path = {out}\nLT = {LT}\nLX = {L}\nNEV = {nev}\n"
                  f"alpha_1 = 0.62\nalpha_2 = 0.62\niter = 3\nLAM_L = 12.0\nLAM_C = 1.0\nDEG = {deg}\n")
    runs = []
    for k in range(2):
        js = tmp_path / f"run{k}.json"
        assert ev_ts.main(["--conf", str(conf), "--params", str(pf), "--json", str(js), "--device", str(gpu)]) == 0
        info = json.loads(js.read_text())
        files = {t: (out / f"eigenvectors.{conf:04d}.{t:03d}").read_bytes() for t in range(LT)}
        runs.append((info["groups"][0], files))
    assert (out / "laph_layouts.txt").exists()
    assert runs[0][0]["layout_from_cache"] == 0 and runs[1][0]["layout_from_cache"] == 1
    assert runs[0][0]["warmup_degree"] == runs[1][0]["warmup_degree"]
    for t in range(LT):
        assert runs[0][1][t] == runs[1][1][t]


def test_batch_write_matches_host_path(gpu, tmp_path):
    """laph_batch_write (device phase fix, device V^H V guard, chunked
    pinned streaming) writes the same bytes as laph_batch_get followed by the
    host writers."""
    from laph_eigsys_amd import lapheig as lp
    L, nts, nev, conf = 6, 2, 16, 42
    cfg = N.synth_config(5, 0.4, L, nts)
    b = lp.Batch(L, 8, 12.0, 0.0, nts, nev, nsub=64, tol=1e-10, seed=1, device=gpu)
    vol = 72 * L ** 3
    for k in range(nts):
        b.set_timeslice(k, cfg[k * vol:(k + 1) * vol], 0.62, 0.62, 3)
    b.solve()
    (tmp_path / "dev").mkdir()
    (tmp_path / "host").mkdir()
    for k in range(nts):
        trc, sm = b.write(k, str(tmp_path / "dev"), conf, 10 + k)
        assert abs(trc - nev) < 1e-10 and abs(sm - nev) < 1e-9
        ev, V, ph, _ = b.get(k)
        lp.write_eig_sys_bin(str(tmp_path / "host"), "eigenvectors", conf, 10 + k, V)
        lp.write_eigenvalues_bin(str(tmp_path / "host"), "eigenvalues", conf, 10 + k, ev)
        lp.write_eigenvalues_bin(str(tmp_path / "host"), "phases", conf, 10 + k, ph)
        for name in ("eigenvectors", "eigenvalues", "phases"):
            f = f"{name}.{conf:04d}.{10 + k:03d}"
            assert (tmp_path / "dev" / f).read_bytes() == (tmp_path / "host" / f).read_bytes(), f
    b.close()


def test_driver_phase_reset(gpu, tmp_path):
    """ev_ts --phase_reset: the files hold the eigenvectors with the stored
    phases (arg V(0, n) = stored phase n) and the stored phases."""
    from laph_eigsys_amd import lapheig as lp
    L, LT, nev, conf = 6, 2, 12, 7
    cfg = N.synth_config(11, 0.4, L, LT)
    (tmp_path / "confs").mkdir()
    lp.write_lime(str(tmp_path / "confs" / f"conf.{conf:04d}"), cfg, LT, L)
    (tmp_path / "old").mkdir()
    rng = np.random.default_rng(0)
    old = {t: rng.uniform(-np.pi, np.pi, nev) for t in range(LT)}
    for t in range(LT):
        lp.write_eigenvalues_bin(str(tmp_path / "old"), "phases", conf, t, old[t])
    runs = {}
    for mode in (0, 1):
        out = tmp_path / f"out{mode}"
        (tmp_path / f"p{mode}.txt").write_text(
            f"config_path = {tmp_path / 'confs'}\nresult_path = {out}\nphase_path = {tmp_path / 'old'}\n"
            f"LT = {LT}\nLX = {L}\nNEV = {nev}\nalpha_1 = 0.62\nalpha_2 = 0.62\niter = 3\nLAM_L = 12.0\n"
            f"LAM_C = 1.0\nDEG = 8\n")
        pars = ev_ts.IO().set_values(str(tmp_path / f"p{mode}.txt"))
        ev_ts.run(pars, conf, solver="batched", phase_reset=bool(mode), device=gpu, log=lambda m: None)
        runs[mode] = out
    n = 3 * L ** 3
    for t in range(LT):
        Vf = lp.read_evectors_bin_ts(str(runs[0]), "eigenvectors", conf, t, n, nev)
        Vr = lp.read_evectors_bin_ts(str(runs[1]), "eigenvectors", conf, t, n, nev)
        np.testing.assert_allclose(np.angle(Vr[:, 0]), old[t], atol=1e-10)
        np.testing.assert_array_equal(lp.read_eigenvalues_bin(str(runs[1]), "phases", conf, t, nev), old[t])
        ev = lp.read_eigenvalues_bin(str(runs[1]), "eigenvalues", conf, t, nev)
        iso = [i for i in range(1, nev - 1) if min(ev[i - 1] - ev[i], ev[i] - ev[i + 1]) > 1e-4]
        for i in iso:  # the same eigenvector, phase-fixed once and reset once
            np.testing.assert_allclose(Vr[i] * np.exp(-1j * old[t][i]), Vf[i], atol=1e-8)


def test_tune_hyp_scan(gpu, tmp_path):
    """tune_hyp: every grid point through the batched solver (one batch) and
    the plain drop-in gives the dense oracle's spectrum of that smearing; the
    per-point files carry the reference's names."""
    from laph_eigsys_amd import lapheig as lp
    from laph_eigsys_amd import tune_hyp
    L, LT, nev, conf, ts = 6, 3, 12, 9, 1
    cfg = N.synth_config(21, 0.4, L, LT)
    (tmp_path / "confs").mkdir()
    lp.write_lime(str(tmp_path / "confs" / f"conf.{conf:04d}"), cfg, LT, L)
    vol = 72 * L ** 3
    grid = [(a1, a2) for a1 in tune_hyp.alpha_grid(0.5, 0.6) for a2 in tune_hyp.alpha_grid(0.5, 0.55)]
    dense = {}
    for a1, a2 in grid:
        U = O.hyp(O.map_timeslice(cfg[ts * vol:(ts + 1) * vol], L), L, a1, a2, 3)
        dense[(a1, a2)] = np.sort(np.linalg.eigvalsh(N.dense_D(U.reshape(L, L, L, 3, 3, 3))))[::-1][:nev]
    lam_c = 1.3 * max(-d[-1] for d in dense.values())
    out = {}
    for solver in ("batched", "plain"):
        res = tmp_path / solver
        (tmp_path / f"{solver}.txt").write_text(
            f"config_path = {tmp_path / 'confs'}\nresult_path = {res}\nLT = {LT}\nLX = {L}\nNEV = {nev}\n"
            f"alpha_1_i = 0.5\nalpha_1_f = 0.6\nalpha_2_i = 0.5\nalpha_2_f = 0.55\niter = 3\nLAM_L = 12.0\n"
            f"LAM_C = {lam_c}\nDEG = 8\n")
        res.mkdir()
        pars = ev_ts.IO().set_values(str(tmp_path / f"{solver}.txt"))
        rows = tune_hyp.tune_hyp(pars, conf, ts=ts, solver=solver, tol=1e-12 if solver == "plain" else 1e-10,
                                 device=gpu, log=lambda m: None)
        assert [(a1, a2) for a1, a2, _ in rows] == grid
        for a1, a2, lo in rows:
            ev = lp.read_eigenvalues_bin(str(res), f"eigenvalues_hyp-3_a1-{a1:.2f}_a2-{a2:.2f}", conf, ts, nev)
            np.testing.assert_allclose(ev, dense[(a1, a2)], rtol=1e-10, atol=1e-12)
            assert lo == ev[0]
            ph = lp.read_eigenvalues_bin(str(res), f"phases_hyp-3_a1-{a1:.2f}_a2-{a2:.2f}", conf, ts, nev)
            assert np.all(np.abs(ph) <= np.pi)
        out[solver] = rows


def test_source_shape_matches_oracle(gpu):
    """psi(r) on the device (host eigenvectors and a solved batch) against the
    numpy restatement of sourceshape_funcs.cpp:31-81."""
    from laph_eigsys_amd import lapheig as lp
    L, nev = 6, 10
    rng = np.random.default_rng(2)
    V = rng.standard_normal((nev, 3 * L ** 3)) + 1j * rng.standard_normal((nev, 3 * L ** 3))
    ref = np.array([p for _, _, p in N.source_shape(V, L)])
    np.testing.assert_allclose(lp.source_shape(L, V, device=gpu), ref, rtol=1e-12)
    L, nev = 6, 16
    cfg = N.synth_config(8, 0.4, L, 1)
    b = lp.Batch(L, 8, 12.0, 0.0, 1, nev, nsub=64, tol=1e-10, seed=1, device=gpu)
    b.set_timeslice(0, cfg, 0.62, 0.62, 3)
    b.solve()
    _, Vb, _, _ = b.get(0)
    psi = b.source_shape(0)
    b.close()
    np.testing.assert_allclose(psi, np.array([p for _, _, p in N.source_shape(Vb, L)]), rtol=1e-12)
