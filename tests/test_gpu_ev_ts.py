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
