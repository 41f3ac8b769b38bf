"""bench.py's contract and its profile summarizers, on synthetic inputs (no
GPU): the one-GPU guard, the HBM-traffic calibration of the --pmc runs, the
FP64 matrix-core rate and the per-phase split of a kernel trace."""
import csv
import importlib.util
import json
import os
import sqlite3
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU only" in r.stderr


def _pmc_csv(path, counter, rows):
    cols = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp",
            "End_Timestamp"]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols)
        w.writeheader()
        for i, (name, value) in enumerate(rows):
            w.writerow({"Correlation_Id": i + 1, "Dispatch_Id": i + 1, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": value, "Start_Timestamp": 1000 * i, "End_Timestamp": 1000 * i + 500})


def test_traffic_calibration(bench, tmp_path, capsys):
    # FETCH_SIZE reads half of a copy's bytes, WRITE_SIZE all of them: the
    # 1 GiB copy calibrates both, so 3 MiB fetched and 2 MiB written per
    # filter dispatch become 6 MiB + 2 MiB of traffic
    gib_kib = (1 << 30) / 1024
    dense = [("_ZN4laph8k_lap64yILb0ELi64E", 1024.0), ("_ZN4laph8k_lap64yILb0ELi64E", 4096.0)]
    fetch = dense + [("k_filter_step32x2", 3 * 1024.0)] * 4 + [("__amd_rocclr_copyBuffer", gib_kib / 2)]
    write = dense + [("k_filter_step32x2", 2 * 1024.0)] * 4 + [("__amd_rocclr_copyBuffer", gib_kib)]
    _pmc_csv(tmp_path / "f.csv", "FETCH_SIZE", fetch)
    _pmc_csv(tmp_path / "w.csv", "WRITE_SIZE", write)
    bench.summarize_pmc(str(tmp_path / "f.csv"), str(tmp_path / "w.csv"), 4)
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert out["filter_dispatches"] == [4, 4]
    assert out["fetch_scale"] == pytest.approx(2.0) and out["write_scale"] == pytest.approx(1.0)
    assert out["traffic_bytes_per_launch"] == pytest.approx(8 * 1024 * 1024)
    # the dense phase's kernels of the same runs, scaled like the filter's
    dx = out["dense_kernels"]["k_lap64y (D X)"]
    assert dx["read"]["dispatches"] == 2 and dx["read"]["max_bytes"] == pytest.approx(2 * 4096 * 1024)
    assert dx["write"]["max_bytes"] == pytest.approx(4096 * 1024)


def test_mfma_rate(bench, tmp_path, capsys):
    # two dispatches of 500 ns each with 512e6 / 512 MOPS: 1.024e9 flop in 1 us
    rows = [("Cijk_gram", 1e6), ("Cijk_gram", 1e6)]
    _pmc_csv(tmp_path / "m.csv", "SQ_INSTS_VALU_MFMA_MOPS_F64", rows)
    bench.summarize_mfma(str(tmp_path / "m.csv"))
    line = [ln for ln in capsys.readouterr().out.splitlines() if ln.startswith("Cijk_gram")][0].split()
    assert int(line[1]) == 2 and float(line[4]) == pytest.approx(1024.0, rel=1e-3)


def test_trace_phases(bench, tmp_path, capsys):
    db = tmp_path / "t.db"
    c = sqlite3.connect(db)
    c.execute("create table rocpd_info_kernel_symbol (id integer, kernel_name text)")
    c.execute("create table rocpd_kernel_dispatch (kernel_id integer, start integer, end integer)")
    names = ["laph::k_random_block", "laph::k_filter_step32x2", "Cijk_Ailk_BjlkC_MT64x64x8", "rocsolver_potf2",
             "laph::k_lap_residual64", "laph::k_residual_norms"]
    for i, n in enumerate(names):
        c.execute("insert into rocpd_info_kernel_symbol values (?, ?)", (i, n))
    t = 0
    for i, dur in enumerate([1e6, 4e6, 2e6, 1e5, 1e6, 1e5]):  # ns
        c.execute("insert into rocpd_kernel_dispatch values (?, ?, ?)", (i, t, t + int(dur)))
        t += int(dur) + 10_000
    c.commit()
    c.close()
    bench.summarize_trace(str(db), 1)
    out = capsys.readouterr().out
    filt = [ln for ln in out.splitlines() if ln.strip().startswith("filter")][0]
    assert "4.0 ms wall" in filt and "1 dispatches" in filt
    assert "large GEMMs" in out and "rocSOLVER" in out


def test_plain_lambda_c_is_above_the_wanted_free_levels(bench):
    """lambda_c of the plain drop-in: 1.35 x the nev-th free-field level
    (levels 4 sum_i sin^2(pi k_i / L), each three-fold)."""
    import numpy as np
    lam = bench.plain_lambda_c(32, 220)
    s2 = 4 * np.sin(np.pi * np.arange(32) / 32) ** 2
    lev = np.sort((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel())
    counts = np.repeat(lev, 3)
    assert abs(lam - 1.35 * counts[219]) < 1e-15
    assert counts[219] < lam < 12.0


def test_spread_and_traffic_tie(bench, monkeypatch):
    assert bench.spread([]) is None
    assert bench.spread([3.0, 1.0, 2.0]) == {"min": 1.0, "median": 2.0, "max": 3.0}
    import hashlib
    src = open(os.path.join(ROOT, bench.TRAFFIC["source"]), "rb").read()
    monkeypatch.setitem(bench.TRAFFIC, "sha16", hashlib.sha256(src).hexdigest()[:16])
    assert bench.traffic_for_build() == bench.TRAFFIC["bytes_per_launch"]
    monkeypatch.setitem(bench.TRAFFIC, "sha16", "0" * 16)
    assert bench.traffic_for_build() is None


def test_host_info(bench):
    h = bench.host_info()
    assert h["nproc"] == os.cpu_count()


def test_measurement_programs_built_and_guarded():
    """The measurement programs bench.py runs (--stream-ceiling,
    --filter-bound, --peak-rates) are built with the library; filter_bound
    rejects a lattice its sweep cannot tile before it touches a device."""
    for name in ("stream_ceiling", "filter_bound", "peak_rates"):
        assert os.access(os.path.join(ROOT, "laph_eigsys_amd", name), os.X_OK), name
    r = subprocess.run([os.path.join(ROOT, "laph_eigsys_amd", "filter_bound"), "12"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2 and "multiple of 8" in r.stderr
