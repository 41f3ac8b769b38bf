"""Benchmark: eigensolved timeslices/s at 32^3 x 64, N_ev = 220, residual <= 1e-10,
mixed-precision batched solver vs the plain FP64 drop-in, on one MI355X.

A step is one batched eigensolve of all 64 timeslices of the seeded
synthetic configuration (U = exp(i 0.3 H), HYP 0.62/0.62/3 applied on the
device beforehand; smeared links resident in HBM when the timed region
starts): FP32 Chebyshev filter in residual-correction form (one warm-up
cycle of degree 16 or the higher degree the first, untimed solve found
faster, then one filter per cycle of the degree the window allows, up to 128) on 64- or 128-vector blocks, FP64 Rayleigh-Ritz / residual
locking, until every timeslice has 220 locked pairs with
|D v - lambda v| <= 1e-10, then fix_phase of every eigenvector: the timed
region ends with phase-fixed eigenpairs resident in HBM.  A timed step that
needed a recovery (a re-seeded block) fails the bench.

Printed JSON (one line) also carries:
  roofline      the FP32 filter stencil (dominant kernel): algorithmic bytes
                per launch / average launch duration, from a replay of the
                captured filter alone on the device (HIP events on the
                library's stream), vs 8 TB/s; `traffic` = measured HBM bytes
                per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE on those
                replays, each scaled by the same counter on a 1 GiB copy),
                reported only while the kernel source is the one measured
  plain_fp64    the plain FP64 drop-in (device-resident Krylov-Schur on
                T_16(B), lambda_c = 1.35 x the 220th free-field level, a rule
                that needs no solve) on every timeslice of the same field,
                timeslices/s, with the per-timeslice spread and the largest
                relative eigenvalue difference to the batched solver
  cpu_baseline  the oracle's Krylov-Schur (C, OpenMP over the host cores)
                solving one timeslice to 1e-10, median of 3 solves;
                --cpu-full runs the one-timeslice-per-core plan instead
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

L, LT, EPS, NEV, DEG, TOL, NSUB = 32, 64, 0.3, 220, 16, 1e-10, 256
SEED = 1227
LAMBDA_L = 12.0
PEAK_HBM = 8.0e12
PEAK_FP64_MFMA = 78.6e12  # dense FP64 matrix peak of the MI355X
# HBM bytes per filter-step launch (all columns active, the 64 timeslices of
# the default single lane) from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on the
# isolated replays of --profile-filter, each scaled by the same counter on a
# 1 GiB device copy (python bench.py --summarize-pmc ...), with the sha256
# (first 16 hex digits) of the source file of the kernel it was measured on:
# the figure is reported only while that file is unchanged.
TRAFFIC = {"bytes_per_launch": 43441229619, "source": "laph_eigsys_amd/csrc/kernels_block.hip",
           "sha16": "dd16b5a5b4ebc6a2"}  # profiles/filter_step_32x32x32x64_k256_pmc.md


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nts", type=int, default=LT)
    ap.add_argument("--plain-ts", type=int, default=LT, help="timeslices solved by the plain FP64 drop-in")
    ap.add_argument("--batch-tol", type=float, default=TOL,
                    help="bound on ||D v - lambda v|| for the batched solver (the plain drop-in keeps the "
                         "reference's -eps_tol 1e-10 on T_DEG(B)); 5.8e-12 meets that criterion for every "
                         "wanted pair of the headline field")
    ap.add_argument("--pcie-ts", type=int, default=8,
                    help="timeslices whose eigenpairs are copied to host memory for the PCIe-inclusive rate")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-full", action="store_true",
                    help="only the CPU baseline as planned: one timeslice per host core, each solved to 1e-10 by "
                         "the oracle (single-threaded), --cpu-runs times")
    ap.add_argument("--cpu-cores", type=int, default=0, help="with --cpu-full: cores (0 = OMP_NUM_THREADS or all)")
    ap.add_argument("--cpu-threads-per-ts", type=int, default=1,
                    help="with --cpu-full: OpenMP threads per timeslice (ranks = cores / threads)")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--cpu-worker", type=str, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--filter-reps", type=int, default=20)
    ap.add_argument("--profile-filter", action="store_true",
                    help="counter-collection workload: filter replays + a calibration copy, no solve timing")
    ap.add_argument("--summarize-trace", metavar="DB",
                    help="per-kernel durations from a rocprofv3 --kernel-trace database (no GPU use)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="HBM bytes per filter launch from two rocprofv3 --pmc runs of --profile-filter")
    ap.add_argument("--summarize-mfma", metavar="PMC_CSV",
                    help="FP64 matrix-core rate per kernel from a rocprofv3 --pmc run with "
                         "SQ_INSTS_VALU_MFMA_MOPS_F64")
    ap.add_argument("--workload", choices=("c3", "c2", "c4", "fields"), default="c3",
                    help="c3: the headline (32^3x64, N_ev 220); c2: 24^3x48 N_ev 120, every timeslice through the "
                         "plain drop-in, the batched solver and the CPU oracle; c4: 48^3x96 N_ev 200 in HBM-sized groups; fields: "
                         "the headline solve on other seeds and field strengths")
    ap.add_argument("--stream-ceiling", action="store_true",
                    help="only run the streaming-bandwidth measurement (laph_eigsys_amd/stream_ceiling)")
    ap.add_argument("--no-stream", action="store_true", help="skip the streaming-bandwidth measurement")
    ap.add_argument("--filter-bound", action="store_true",
                    help="only run laph_eigsys_amd/filter_bound: the filter step beside variants with a part taken "
                         "away and copies with its load pattern (32^3, 64 timeslices, 256 vectors)")
    ap.add_argument("--peak-rates", action="store_true",
                    help="only run laph_eigsys_amd/peak_rates: FP32 and FP64 FMA and FP64 MFMA issue rates with no "
                         "memory traffic, at the shader clock measured during the run")
    ap.add_argument("--last", type=int, default=0,
                    help="with --summarize-*: the isolated filter replays are the last LAST filter dispatches")
    args = ap.parse_args()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.last)
        return
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, args.last)
        return
    if args.summarize_mfma:
        summarize_mfma(args.summarize_mfma)
        return
    if args.stream_ceiling:
        print(json.dumps(stream_ceiling()))
        return
    if args.filter_bound:
        print(json.dumps(run_measure("filter_bound")))
        return
    if args.peak_rates:
        print(json.dumps(run_measure("peak_rates")))
        return
    if args.cpu_worker:
        cpu_worker(args.cpu_worker)
        return
    if args.cpu_full:
        print(json.dumps(cpu_full(args.cpu_cores, args.cpu_runs, args.nts, args.cpu_threads_per_ts)))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); timeslices are "
              "independent, so N GPUs would each run their own Lt/N share with no exchange.", file=sys.stderr)
        sys.exit(2)

    from laph_eigsys_amd import lapheig as lp
    if args.workload == "c2":
        workload_c2(lp, args)
        return
    if args.workload == "c4":
        workload_c4(lp, args)
        return
    if args.workload == "fields":
        workload_fields(lp, args)
        return
    if args.profile_filter:
        profile_filter(lp, args)
        return
    nts = args.nts
    t0 = time.time()
    cfg = lp.synth_config(SEED, EPS, L, nts)
    vol = 72 * L ** 3
    btol = args.batch_tol
    b = lp.Batch(L, DEG, LAMBDA_L, 0.0, nts, NEV, nsub=NSUB, tol=btol, seed=SEED)
    t_field = time.time() - t0
    sync = device_sync()
    sync()
    t1 = time.perf_counter()
    for t in range(nts):
        b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
    sync()
    t_hyp = time.perf_counter() - t1  # host-to-device copy, map to links and HYP of every timeslice
    print(f"[bench] setup {time.time() - t0:.1f}s (field {t_field:.1f}s, upload + map + HYP of {nts} timeslices "
          f"{t_hyp:.2f}s)", file=sys.stderr, flush=True)

    for w in range(args.warmup):
        st = b.solve()
        print(f"[bench] warmup {w}: {st.seconds:.2f}s cycles={st.iterations}", file=sys.stderr, flush=True)
    sync()
    tot = {"seconds": 0.0, "filter": 0.0, "fbytes": 0.0, "flaunch": 0, "cycles": 0, "maxres": 0.0,
           "recoveries": 0, "rr_fallbacks": 0, "shifted": 0}
    t_start = time.perf_counter()
    for s in range(args.steps):
        st = b.solve()
        tot["seconds"] += st.seconds
        tot["filter"] += st.seconds_filter
        tot["fbytes"] += st.filter_bytes
        tot["flaunch"] += st.filter_launches
        tot["cycles"] = max(tot["cycles"], st.iterations)
        tot["maxres"] = max(tot["maxres"], st.max_residual)
        tot["recoveries"] += st.recoveries
        tot["rr_fallbacks"] += st.rr_cholqr_fallbacks
        tot["shifted"] += st.shifted_cholqr
        assert st.converged == nts, "not every timeslice converged"
        print(f"[bench] step {s}: {st.seconds:.2f}s cycles={st.iterations} filter={st.seconds_filter:.2f}s "
              f"dense={st.seconds_dense:.2f}s maxres={st.max_residual:.2e} rr_fallbacks={st.rr_cholqr_fallbacks} "
              f"shifted={st.shifted_cholqr} recoveries={st.recoveries}", file=sys.stderr, flush=True)
    sync()
    wall = time.perf_counter() - t_start
    if tot["recoveries"] > 0:
        sys.exit(f"bench.py: {tot['recoveries']} recoveries (re-seeded blocks) in the timed steps")
    ms_per_step = wall / args.steps * 1e3
    value = nts * args.steps / wall
    ev_batch = [b.get(t, want_vectors=False)[0] for t in range(nts)]
    # dominant kernel: FP32 filter stencil, timed alone on the device
    ms_launch, bytes_launch = b.time_filter(reps=args.filter_reps)
    gram = gram_rate(b, st)
    eps_res = eps_residual(lp, b, cfg, vol, sorted({0, nts - 1}), plain_lambda_c(L, NEV))
    # PCIe-inclusive rate (never `value`): the host configuration goes in
    # (upload, map and HYP, timed above) and every phase-fixed eigenpair comes
    # out to host memory (laph_batch_get into one reused buffer, its pages
    # touched first); the download is timed on a sample of timeslices
    n_dl = max(1, min(args.pcie_ts, nts))
    vbytes = NEV * 3 * L ** 3 * 16
    pcie = {"unit": "timeslices/s", "upload_map_hyp_seconds": t_hyp, "download_sample_timeslices": n_dl,
            "definition": "timeslices / (solve + upload, map and HYP of every timeslice + download of every "
                          "timeslice's eigenpairs to host memory)"}
    for kind in ("pageable", "pinned"):
        with host_buffer((NEV, 3 * L ** 3), pinned=kind == "pinned") as vbuf:
            sync()
            t0 = time.perf_counter()
            for t in range(n_dl):
                b.get(t, out=vbuf)
            t_dl = (time.perf_counter() - t0) / n_dl
            del vbuf  # the pinned memory is freed when the block ends
        pcie[kind] = {"value": nts / (wall / args.steps + t_hyp + nts * t_dl),
                      "download_seconds_per_timeslice": t_dl, "download_GBps": vbytes / t_dl / 1e9}
    b.close()
    achieved = bytes_launch / (ms_launch * 1e-3)
    roofline = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": achieved / PEAK_HBM, "traffic": traffic_for_build(), "traffic_unit": "bytes/launch",
                "kernel": (f"k_filter_step32{'x2y' if st.filter_block == 128 else 'y'} (block {st.filter_block}, "
                           f"y-row groups walking z-lines in runs of {st.filter_zrun})"
                           if st.filter_zrun else
                           f"k_filter_step32{'x2' if st.filter_block == 128 else ''} (block {st.filter_block})"),
                "avg_launch_us": ms_launch * 1e3, "bytes_per_launch": bytes_launch,
                "in_solve_GBps": tot["fbytes"] / tot["filter"] / 1e9 if tot["filter"] > 0 else None}
    if roofline["traffic"] is None:
        roofline["traffic_note"] = f"not measured for this build of {TRAFFIC['source']}"

    if not args.no_stream:
        # what a plain streaming kernel reaches for the timed step's access mix
        # (1 read + 1 write, non-temporal) on this device: the practical ceiling
        sc = stream_ceiling()
        roofline["stream_1r1w_GBps"] = sc.get("1r1w_nt_GBps")
        roofline["frac_of_stream"] = achieved / 1e9 / sc["1r1w_nt_GBps"] if sc.get("1r1w_nt_GBps") else None

    # plain FP64 drop-in on every timeslice of the same field, lambda_c from
    # the free-field levels (no solve of the field needed to choose it)
    lamC = plain_lambda_c(L, NEV)
    plain_times, plain_cycles, plain_dev = [], [], 0.0
    for t in range(min(args.plain_ts, nts)):
        with lp.Tslice(L, DEG, LAMBDA_L, lamC, ts_config=cfg[t * vol:(t + 1) * vol]) as ts:
            ts.smearing_hyp(0.62, 0.62, 3)
            sync()
            tp = time.perf_counter()
            theta, _, pst = ts.eps_solve(NEV, tol=TOL, want_vectors=False)
            plain_times.append(time.perf_counter() - tp)
            plain_cycles.append(pst.restarts + 1)
        ev = lp.recover_spectrum(theta, DEG, LAMBDA_L, lamC)
        plain_dev = max(plain_dev, float(np.max(np.abs(ev - ev_batch[t]) / np.abs(ev_batch[t]))))
    plain_value = len(plain_times) / sum(plain_times) if plain_times else None
    print(f"[bench] plain FP64: {len(plain_times)} timeslices, {plain_value:.2f} ts/s, batched-vs-plain max rel "
          f"eigenvalue difference {plain_dev:.2e}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(cfg[:vol], lamC)

    out = {
        "metric": "eigensolved timeslices/s, 32^3 N_ev=220 res<=1e-10: mixed+batched vs plain FP64",
        "value": value, "unit": "timeslices/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 filter + fp64 CholQR2/RR/lock", "data": "synthetic",
        "config": {"workload": f"{L}^3x{nts} eps={EPS} HYP 0.62/0.62/3, N_ev={NEV}, Chebyshev degree {DEG}, "
                               f"one filter pass per cycle of degree <= 128 after one of degree {st.warmup_degree}, "
                               f"vector block "
                               f"{st.filter_block}, sweep strip {st.filter_strip}, search space {NSUB}, "
                               f"residual <= {btol}, ending with phase-fixed eigenpairs in HBM",
                   "lattice": f"{L}^3x{nts}", "nev": NEV, "deg": DEG, "warmup_degree": st.warmup_degree, "tol": btol,
                   "parallelism": "none (1 GPU)"},
        "roofline": roofline,
        "gram": gram,
        "plain_fp64": {"value": plain_value, "unit": "timeslices/s", "timeslices": len(plain_times), "eps_tol": TOL,
                       "lambda_c": lamC, "cycles_max": max(plain_cycles) if plain_cycles else None,
                       "seconds_per_timeslice": spread(plain_times),
                       "max_rel_eigenvalue_difference_vs_batched": plain_dev},
        "speedup_vs_plain_fp64": (value / plain_value) if plain_value else None,
        "max_residual": tot["maxres"], "eps_relative_residual": eps_res, "cycles": tot["cycles"],
        "fallbacks": {"recoveries": tot["recoveries"], "rr_cholqr": tot["rr_fallbacks"],
                      "shifted_cholqr": tot["shifted"]},
        "upload_map_hyp_seconds": t_hyp,
        "pcie_inclusive": pcie,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def eps_residual(lp, b, cfg, vol, tslices, lam_c):
    """The residual that the reference's -eps_tol bounds, ||T_DEG(B) v - theta v|| / |theta|
    with theta = T_DEG(c lambda + a) (ev_ts.cpp:263-292), for the batched solver's
    phase-fixed eigenpairs (v, lambda) of the given timeslices, B taken at the
    plain drop-in's lambda_c; T_DEG(B) v through the MatShell entry point."""
    c = 2.0 / (LAMBDA_L - lam_c)
    a = 1.0 + c * lam_c
    worst = 0.0
    for t in tslices:
        ev, V, _, _ = b.get(t)
        theta = np.cos(DEG * np.arccos((c * ev + a).astype(np.complex128))).real
        with lp.Tslice(L, DEG, LAMBDA_L, lam_c, ts_config=cfg[t * vol:(t + 1) * vol]) as ts:
            ts.smearing_hyp(0.62, 0.62, 3)
            for j in range(V.shape[0]):
                y = ts.MatMult_Laplacian2D(V[j])
                worst = max(worst, float(np.linalg.norm(y - theta[j] * V[j]) / abs(theta[j])))
    return {"max": worst, "timeslices": list(tslices), "lambda_c": lam_c,
            "definition": f"||T_{DEG}(B) v - theta v|| / |theta| over the N_ev phase-fixed pairs"}


def profile_filter(lp, args):
    """Counter-collection workload: one batched solve of --nts timeslices (for
    the coefficient tables), then --filter-reps replays of the captured FP32
    filter alone, then a 1 GiB device-to-device copy whose FETCH_SIZE /
    WRITE_SIZE calibrate the counters (2 x 1 GiB of known traffic)."""
    import ctypes
    nts = args.nts
    cfg = lp.synth_config(SEED, EPS, L, nts)
    vol = 72 * L ** 3
    b = lp.Batch(L, DEG, LAMBDA_L, 0.0, nts, NEV, nsub=NSUB, tol=TOL, seed=SEED)
    for t in range(nts):
        b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
    t0 = time.perf_counter()
    b.solve()
    print(f"[bench] profile-filter: solve {time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    ms, by = b.time_filter(reps=args.filter_reps)
    print(f"[bench] profile-filter: {args.filter_reps} replays, {ms:.3f} ms per launch", file=sys.stderr, flush=True)
    b.close()
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 1 << 30
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"filter_ms_per_launch": ms, "filter_bytes_per_launch": by, "filter_GBps": by / ms / 1e6,
                      "lane_timeslices": nts // int(os.environ.get("LAPH_LANES", "1")), "copy_bytes": nbytes,
                      "isolated_filter_dispatches": (args.filter_reps + 1) * (DEG - 1)}))


def _filter_kernel(name):
    return "k_filter_step" in name


def summarize_trace(db, last):
    """Per-kernel totals of a rocprofv3 kernel-trace database, and the mean
    duration of the last `last` filter dispatches (the isolated replays)."""
    import sqlite3
    c = sqlite3.connect(db)
    rows = list(c.execute(
        "select s.kernel_name, d.start, d.end from rocpd_kernel_dispatch d "
        "join rocpd_info_kernel_symbol s on d.kernel_id = s.id order by d.start"))
    agg = {}
    for name, t0, t1 in rows:
        a = agg.setdefault(name, [0, 0.0, float("inf"), 0.0])
        dt = (t1 - t0) / 1e3
        a[0] += 1
        a[1] += dt
        a[2] = min(a[2], dt)
        a[3] = max(a[3], dt)
    tot = sum(a[1] for a in agg.values())
    print(f"{'kernel':72s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s} {'pct':>6s}")
    for name, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:25]:
        print(f"{name[:72]:72s} {a[0]:7d} {a[1] / 1e3:10.2f} {a[1] / a[0]:10.2f} {a[2]:9.2f} {a[3]:9.2f} "
              f"{100 * a[1] / tot:6.1f}")
    if last:
        fl = [(t1 - t0) / 1e3 for name, t0, t1 in rows if _filter_kernel(name)][-last:]
        print(f"isolated filter replays: last {len(fl)} filter dispatches, mean {sum(fl) / len(fl):.2f} us")
    _solve_phases(rows)


def _phase(name):
    if _filter_kernel(name):
        return "filter (FP32 stencil steps)"
    if "k_gram_mfma" in name or (name.startswith("Cijk") and ("MT64x64" in name or "MT128x64" in name)):
        return "large GEMMs (Gram, Rayleigh quotient, rotation, projection)"
    if "laph" in name:
        return "other library kernels (FP64 stencil, residual, combine, ...)"
    if "rocsolver" in name or "rocblas" in name or name.startswith("Cijk"):
        return "rocSOLVER / small rocBLAS (Cholesky, inverse, eigensolver)"
    return "runtime copies and fills"


def _solve_phases(rows):
    """Wall-clock split of the last batched solve in the trace (from its
    Gaussian start block to its last residual): per phase, the time from the
    first to the last dispatch of each run of consecutive same-phase
    dispatches (kernels plus the gaps between them) and the kernel time."""
    starts = [t0 for name, t0, t1 in rows if "k_random_block" in name]
    ends = [t1 for name, t0, t1 in rows if "k_residual" in name]
    if not starts or not ends:
        return
    sel = [r for r in rows if starts[-1] <= r[1] <= ends[-1]]
    runs = []
    for name, t0, t1 in sel:
        ph = _phase(name)
        if runs and runs[-1][0] == ph:
            runs[-1][2] = max(runs[-1][2], t1)
            runs[-1][3] += t1 - t0
            runs[-1][4] += 1
        else:
            runs.append([ph, t0, t1, t1 - t0, 1])
    agg = {}
    for ph, t0, t1, busy, cnt in runs:
        a = agg.setdefault(ph, [0.0, 0.0, 0])
        a[0] += (t1 - t0) / 1e6
        a[1] += busy / 1e6
        a[2] += cnt
    span = (sel[-1][2] - sel[0][1]) / 1e6
    print(f"last solve: {span:.1f} ms from its start block to its last residual")
    for ph, (wall, busy, cnt) in sorted(agg.items(), key=lambda kv: -kv[1][0]):
        print(f"  {ph:62s} {wall:8.1f} ms wall {busy:8.1f} ms kernels {cnt:6d} dispatches")


def summarize_mfma(path):
    """Per kernel: FP64 matrix flops (512 per SQ_INSTS_VALU_MFMA_MOPS_F64
    unit), summed dispatch time and the resulting TFLOP/s against the FP64
    matrix peak (other counters in the file are ignored)."""
    import csv
    disp = {}
    for r in csv.DictReader(open(path)):
        d = disp.setdefault(r["Dispatch_Id"], {"name": r["Kernel_Name"],
                                               "ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
        d[r["Counter_Name"]] = d.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    agg = {}
    for d in disp.values():
        a = agg.setdefault(d["name"][:70], [0, 0.0, 0.0, 0.0, 0.0])
        a[0] += 1
        a[1] += d["ns"]
        a[2] += 512.0 * d.get("SQ_INSTS_VALU_MFMA_MOPS_F64", 0.0)
    print(f"{'kernel':70s} {'calls':>6s} {'ms':>9s} {'TFLOP':>9s} {'TFLOP/s':>8s} {'of peak':>8s}")
    for name, (n, ns, fl, _, _) in sorted(agg.items(), key=lambda kv: -kv[1][2]):
        if fl <= 0:
            continue
        rate = fl / (ns * 1e-9) / 1e12
        print(f"{name:70s} {n:6d} {ns / 1e6:9.2f} {fl / 1e12:9.3f} {rate:8.2f} {rate / (PEAK_FP64_MFMA / 1e12):8.3f}")


# kernels of the dense phase whose counter traffic --summarize-pmc reports
# beside the filter's: label and the name fragments (demangled, as the
# counter files give them, or mangled) that identify them
_DENSE_KERNELS = {"k_lap64y (D X)": ("k_lap64y<false", "k_lap64yILb0"),
                  "k_lap64y (residual)": ("k_lap64y<true", "k_lap64yILb1"),
                  "k_combine_gamma": ("k_combine_gamma",), "k_residual_norms": ("k_residual_norms",),
                  "k_gram_mfma": ("k_gram_mfma",)}


def summarize_pmc(fetch_csv, write_csv, last):
    """HBM bytes per isolated filter launch: mean FETCH_SIZE and WRITE_SIZE
    (KiB per dispatch) over the last `last` filter dispatches, each scaled by
    the known bytes of the 1 GiB calibration copy over what the same counter
    reads for it (the copy dispatches after the last filter dispatch)."""
    import csv
    nbytes = 1 << 30

    def read(path, counter):
        rows = [r for r in csv.DictReader(open(path)) if r["Counter_Name"] == counter]
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        filt = [i for i, r in enumerate(rows) if _filter_kernel(r["Kernel_Name"])]
        sel = filt[-last:]
        val = sum(float(rows[i]["Counter_Value"]) for i in sel) / len(sel)
        copy = sum(float(r["Counter_Value"]) for r in rows[filt[-1] + 1:] if "copyBuffer" in r["Kernel_Name"])
        return val, copy, len(sel)

    f, fc, nf = read(fetch_csv, "FETCH_SIZE")
    w, wc, nw = read(write_csv, "WRITE_SIZE")
    kf, kw = (nbytes / 1024) / fc, (nbytes / 1024) / wc
    traffic = (f * kf + w * kw) * 1024
    # the dense phase's own kernels in the same runs (the solve before the
    # replays): scaled bytes read and written per dispatch, median and largest
    # (the largest are the full-width dispatches); the two runs are separate
    # processes, so reads and writes are summarised apart
    others = {}
    for counter, path, scale, key in (("FETCH_SIZE", fetch_csv, kf, "read"), ("WRITE_SIZE", write_csv, kw, "write")):
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = next((k for k, frags in _DENSE_KERNELS.items() if any(f in r["Kernel_Name"] for f in frags)),
                        None)
            if name:
                others.setdefault(name, {}).setdefault(key, []).append(float(r["Counter_Value"]) * scale * 1024)
    dense = {}
    for name, d in sorted(others.items()):
        dense[name] = {}
        for key, v in d.items():
            v.sort()
            dense[name][key] = {"dispatches": len(v), "median_bytes": v[len(v) // 2], "max_bytes": v[-1]}
    print(json.dumps({"filter_dispatches": [nf, nw], "fetch_KiB": f, "write_KiB": w,
                      "copy_fetch_KiB": fc, "copy_write_KiB": wc, "fetch_scale": kf, "write_scale": kw,
                      "traffic_bytes_per_launch": traffic, "dense_kernels": dense}))


def workload_fields(lp, args):
    """The headline solve (32^3 x 64, N_ev 220, DEG 16, K 256, tol 1e-10) on
    synthetic fields of other seeds and field strengths eps, one batch object
    reused (layouts timed on the first field): per field the timed solve's
    rate, cycles, largest residual and fallbacks, and the largest relative
    eigenvalue difference to the plain FP64 drop-in on timeslices 0 and 63
    (lambda_c 1.35 x the batched |lambda_220| of the timeslice)."""
    vol = 72 * L ** 3
    sync = device_sync()
    b = lp.Batch(L, DEG, LAMBDA_L, 0.0, LT, NEV, nsub=NSUB, tol=TOL, seed=SEED)
    rows = []
    try:
        for seed, eps in ((1227, 0.3), (7, 0.3), (2024, 0.3), (1227, 0.2), (1227, 0.4), (99, 0.45)):
            cfg = lp.synth_config(seed, eps, L, LT)
            for t in range(LT):
                b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
            b.solve()  # untimed (the first one also times the layouts)
            sync()
            t0 = time.perf_counter()
            st = b.solve()
            sync()
            dt = time.perf_counter() - t0
            dev = 0.0
            for t in (0, LT - 1):
                evb = b.get(t, want_vectors=False)[0]
                lam_c = 1.35 * abs(evb[-1])  # above the wanted part of this field's spectrum
                with lp.Tslice(L, DEG, LAMBDA_L, lam_c, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
                    sl.smearing_hyp(0.62, 0.62, 3)
                    theta, _, _ = sl.eps_solve(NEV, tol=TOL, want_vectors=False)
                ev = lp.recover_spectrum(theta, DEG, LAMBDA_L, lam_c)
                dev = max(dev, float(np.max(np.abs(ev - evb) / np.abs(evb))))
            rows.append({"seed": seed, "eps": eps, "ts_per_s": LT / dt, "cycles": st.iterations,
                         "converged": st.converged, "max_residual": st.max_residual, "recoveries": st.recoveries,
                         "rr_cholqr_fallbacks": st.rr_cholqr_fallbacks, "shifted_cholqr": st.shifted_cholqr,
                         "max_rel_eigenvalue_difference_vs_plain": dev})
            print(f"[bench] field seed {seed} eps {eps}: {LT / dt:.2f} ts/s, {st.iterations} cycles, "
                  f"max res {st.max_residual:.2e}, vs plain {dev:.1e}", file=sys.stderr, flush=True)
    finally:
        b.close()
    print(json.dumps({"workload": f"{L}^3x{LT} N_ev={NEV} DEG {DEG} K {NSUB} tol {TOL}, HYP 0.62/0.62/3",
                      "fields": rows}))


def workload_c2(lp, args):
    """24^3 x 48, eps 0.4, HYP 0.62/0.62/3, N_ev 120, DEG 12, tol 1e-10: every
    timeslice through the plain FP64 drop-in (the MatShell path) and through
    the batched solver, and (unless --no-cpu) through the oracle on the host
    cores; eigenvalues compared timeslice by timeslice."""
    L2, T2, NEV2, DEG2 = 24, 48, 120, 12
    cfg = lp.synth_config(1227, 0.4, L2, T2)
    vol = 72 * L2 ** 3
    sync = device_sync()
    b = lp.Batch(L2, DEG2, LAMBDA_L, 0.0, T2, NEV2, tol=TOL, seed=SEED)
    for t in range(T2):
        b.set_timeslice(t, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
    b.solve()  # warm-up (layout autotuning, graph capture)
    sync()
    t0 = time.perf_counter()
    st = b.solve()
    sync()
    t_batch = time.perf_counter() - t0
    evb = [b.get(t, want_vectors=False)[0] for t in range(T2)]
    sync()
    t0 = time.perf_counter()
    b.source_shape(0)  # source shape of one timeslice on the device (12^3 vectors r)
    sync()
    t_shape = time.perf_counter() - t0
    b.close()
    lam_c = 1.35 * max(-e[-1] for e in evb)
    t_plain, dev, restarts, evp = 0.0, 0.0, [], []
    for t in range(T2):
        with lp.Tslice(L2, DEG2, LAMBDA_L, lam_c, ts_config=cfg[t * vol:(t + 1) * vol]) as sl:
            sl.smearing_hyp(0.62, 0.62, 3)
            sync()
            tp = time.perf_counter()
            theta, _, pst = sl.eps_solve(NEV2, tol=TOL, want_vectors=False)
            t_plain += time.perf_counter() - tp
        restarts.append(pst.restarts)
        evp.append(lp.recover_spectrum(theta, DEG2, LAMBDA_L, lam_c))
        dev = max(dev, float(np.max(np.abs(evp[t] - evb[t]) / np.abs(evb[t]))))
    out = {"workload": "24^3x48 eps=0.4 N_ev=120 DEG 12 tol 1e-10", "timeslices": T2,
           "batched_ts_per_s": T2 / t_batch, "batched_cycles": st.iterations,
           "batched_warmup_degree": st.warmup_degree, "batched_max_residual": st.max_residual,
           "batched_fallbacks": {"recoveries": st.recoveries, "rr_cholqr": st.rr_cholqr_fallbacks,
                                 "shifted_cholqr": st.shifted_cholqr},
           "plain_fp64_ts_per_s": T2 / t_plain, "plain_lambda_c": lam_c,
           "plain_restarts_max": max(restarts), "max_rel_eigenvalue_difference": dev,
           "source_shape_seconds_per_timeslice": t_shape}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_c2(cfg, L2, DEG2, NEV2, lam_c, evp, evb)
    print(json.dumps(out))


def cpu_c2(cfg, L2, deg, nev, lam_c, evp, evb):
    """The C2 comparison with the CPU run: the oracle (map, HYP, Krylov-Schur
    on T_deg(B) with the plain drop-in's lambda_c, recover_spectrum; C with
    OpenMP over the host cores) on every timeslice of the field, its
    eigenvalues against the plain FP64 drop-in's and the batched solver's,
    timeslice by timeslice; seconds per timeslice (HYP and solve)."""
    from oracle import oracle as O
    vol = 72 * L2 ** 3
    secs, d_plain, d_batch = [], 0.0, 0.0
    for t in range(len(evp)):
        t0 = time.perf_counter()
        U = O.hyp(O.map_timeslice(cfg[t * vol:(t + 1) * vol], L2), L2, 0.62, 0.62, 3)
        nconv, theta, _, _ = O.eps_solve(U, L2, LAMBDA_L, lam_c, deg, nev, tol=TOL)
        secs.append(time.perf_counter() - t0)
        if nconv < nev:
            raise RuntimeError(f"c2 cpu: oracle converged {nconv} of {nev} pairs on timeslice {t}")
        ev = O.recover_spectrum(theta[:nev], deg, LAMBDA_L, lam_c)
        d_plain = max(d_plain, float(np.max(np.abs(ev - evp[t]) / np.abs(ev))))
        d_batch = max(d_batch, float(np.max(np.abs(ev - evb[t]) / np.abs(ev))))
        if t % 8 == 7:
            print(f"[bench] c2 cpu: {t + 1} timeslices, {sum(secs):.0f}s", file=sys.stderr, flush=True)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": len(secs) / sum(secs), "unit": "timeslices/s", "cores": cores, "kind": "port",
            "sample": f"all {len(secs)} timeslices, one after another, HYP and Krylov-Schur to {TOL} with OpenMP "
                      f"over {cores} threads",
            "seconds_per_timeslice": spread(secs), "max_rel_eigenvalue_difference_vs_plain_fp64": d_plain,
            "max_rel_eigenvalue_difference_vs_batched": d_batch, **host_info()}


def workload_c4(lp, args):
    """48^3 x 96, eps 0.3, HYP 0.62/0.62/3, N_ev 200, DEG 16, tol 1e-10.  The
    phase-fixed eigensystems of all 96 timeslices stay resident in HBM (a
    102 GB device store, file layout) while the solver's work buffers serve
    groups of timeslices in turn, sized to what HBM has left beside the store
    (the per-configuration driver's plan); a device-to-file write of two
    timeslices and the FP64 Gram rate are measured after."""
    import ctypes
    from laph_eigsys_amd import ev_ts
    L4, T4, NEV4 = 48, 96, 200
    n4 = 3 * L4 ** 3
    per_ts = n4 * NEV4 * 16
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    store = ctypes.c_void_p()
    rc = hip.hipMalloc(ctypes.byref(store), ctypes.c_size_t(per_ts * T4))
    if rc != 0:
        raise RuntimeError(f"hipMalloc of the {per_ts * T4 / 1e9:.0f} GB eigensystem store failed ({rc})")
    budget = (free.value - per_ts * T4) / 1e9 - 12.0  # links, staging, library workspaces
    groups = ev_ts.plan_groups(T4, L4, NEV4, budget)
    print(f"[bench] c4: {free.value / 1e9:.0f} GB free, store {per_ts * T4 / 1e9:.0f} GB, work-buffer budget "
          f"{budget:.0f} GB -> groups {[g for _, g in groups]}", file=sys.stderr, flush=True)
    cfg = lp.synth_config(SEED, EPS, L4, T4)
    vol = 72 * L4 ** 3
    sync = device_sync()
    gn = groups[0][1]
    b = lp.Batch(L4, DEG, LAMBDA_L, 0.0, gn, NEV4, tol=TOL, seed=SEED)
    for k in range(gn):
        b.set_timeslice(k, cfg[k * vol:(k + 1) * vol], 0.62, 0.62, 3)
    b.solve()  # warm-up on the first group (layout autotuning, graph capture)
    sync()
    t_solve, t_set, t_store, cycles, maxres = 0.0, 0.0, 0.0, 0, 0.0
    fallbacks = {"recoveries": 0, "rr_cholqr": 0, "shifted_cholqr": 0}
    for g0, n in groups:
        ts0 = time.perf_counter()
        for k in range(n):
            b.set_timeslice(k, cfg[(g0 + k) * vol:(g0 + k + 1) * vol], 0.62, 0.62, 3)
        sync()
        t_set += time.perf_counter() - ts0
        t0 = time.perf_counter()
        st = b.solve()
        sync()
        t_solve += time.perf_counter() - t0
        cycles, maxres = max(cycles, st.iterations), max(maxres, st.max_residual)
        fallbacks["rr_cholqr"] += st.rr_cholqr_fallbacks
        fallbacks["shifted_cholqr"] += st.shifted_cholqr
        assert st.converged == n and st.recoveries == 0
        t0 = time.perf_counter()
        for k in range(n):
            b.get_device(k, store.value + (g0 + k) * per_ts)
        t_store += time.perf_counter() - t0
        print(f"[bench] c4 group {g0}..{g0 + n - 1}: {st.seconds:.2f}s cycles={st.iterations}", file=sys.stderr,
              flush=True)
    ms, by = b.time_filter(reps=5)
    gram = gram_rate(b, st)
    # streaming device -> file of two timeslices' eigensystems (1.06 GB of
    # eigenvectors each) into a scratch directory that is removed afterwards
    import shutil
    import tempfile
    tmp = tempfile.mkdtemp(prefix="laph_c4_", dir=os.environ.get("TMPDIR", "/tmp"))
    try:
        tw = time.perf_counter()
        for k in range(2):
            b.write(k, tmp, 1, k)
        t_write = time.perf_counter() - tw
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    wbytes = 2 * 3 * L4 ** 3 * NEV4 * 16
    b.close()
    # the store holds every timeslice: the first and the last are orthonormal
    traces = []
    for t in (0, T4 - 1):
        V = np.empty((NEV4, n4), dtype=np.complex128)
        assert hip.hipMemcpy(V.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(store.value + t * per_ts),
                             ctypes.c_size_t(per_ts), 2) == 0
        ok, tr, sm = lp.check_trace(V)
        assert ok, (t, tr, sm)
        traces.append(tr.real)
    hip.hipFree(store)
    print(json.dumps({"workload": "48^3x96 eps=0.3 N_ev=200 DEG 16 tol 1e-10", "timeslices": T4,
                      "resident_eigensystems_GB": per_ts * T4 / 1e9, "store_trace_ts0_ts95": traces,
                      "store_seconds": t_store,
                      "write_GBps": wbytes / t_write / 1e9, "write_seconds_per_timeslice": t_write / 2,
                      "groups": [n for _, n in groups], "ts_per_s_solve": T4 / t_solve,
                      "ts_per_s_with_upload_hyp": T4 / (t_solve + t_set), "cycles_max": cycles,
                      "warmup_degree": st.warmup_degree, "fallbacks": fallbacks,
                      "max_residual": maxres, "filter_block": st.filter_block, "filter_strip": st.filter_strip,
                      "filter_GBps": by / ms / 1e6, "filter_us_per_launch": ms * 1e3, "gram": gram}))


def run_measure(name, *argv):
    """JSON printed by one of the measurement programs built with the library
    (laph_eigsys_amd/<name>, from laph_eigsys_amd/csrc/measure/<name>.hip)."""
    import subprocess
    exe = os.path.join(ROOT, "laph_eigsys_amd", name)
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} missing: run __graft_entry__.build()")
    r = subprocess.run([exe, *map(str, argv)], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"{name} failed ({r.returncode}): {r.stderr.strip()}")
    return json.loads(r.stdout)


def stream_ceiling():
    """JSON of laph_eigsys_amd/stream_ceiling: best GB/s of float4 streaming
    kernels for 1-3 reads + 1 write and read-only."""
    return run_measure("stream_ceiling")


class host_buffer:
    """A complex128 host array of the given shape, page-locked
    (hipHostMalloc) when pinned, its pages touched before use."""

    def __init__(self, shape, pinned):
        self.shape, self.pinned, self.ptr = shape, pinned, None

    def __enter__(self):
        import ctypes
        n = int(np.prod(self.shape))
        if not self.pinned:
            a = np.zeros(self.shape, dtype=np.complex128)
            return a
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipHostMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(n * 16), 0) == 0
        a = np.ctypeslib.as_array((ctypes.c_double * (2 * n)).from_address(self.ptr.value))
        a = a.view(np.complex128).reshape(self.shape)
        a.fill(0)
        return a

    def __exit__(self, *exc):
        if self.ptr is not None:
            self.hip.hipHostFree(self.ptr)
            self.ptr = None
        return False


def device_sync():
    """hipDeviceSynchronize through the HIP runtime the library already loaded."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def sync():
        rc = hip.hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError(f"hipDeviceSynchronize failed with {rc}")
    return sync


def gram_rate(b, st, reps=10):
    """The FP64 matrix-core kernel of the dense phase: the Gram product of the
    first lane's timeslices at full width, timed alone on the device."""
    ms, flops = b.time_gram(reps=reps)
    rate = flops / (ms * 1e-3)
    kern = ("k_gram_mfma (hand-written, v_mfma_f64_16x16x4, 64x64 upper tiles)" if st.gram_mfma else
            f"rocBLAS strided-batched zgemm, {st.gram_blocks} row block(s)")
    return {"kernel": kern + f", {st.gram_chunks} site chunk(s)", "ms": ms, "tflop_executed": flops / 1e12,
            "TFLOPps": rate / 1e12, "frac_fp64_mfma_peak": rate / PEAK_FP64_MFMA}


def plain_lambda_c(L, nev, factor=1.35):
    """lambda_c of the plain drop-in: factor x the nev-th lowest level of -D
    on the free field (every link 1: 4 sum_i sin^2(pi k_i / L), each level
    three-fold).  HYP-smeared fields lie close above the free levels, so the
    wanted nev eigenvalues stay below lambda_c without solving the field."""
    s2 = 4 * np.sin(np.pi * np.arange(L) / L) ** 2
    free = np.sort(np.repeat((s2[:, None, None] + s2[None, :, None] + s2[None, None, :]).ravel(), 3))
    return factor * float(free[nev - 1])


def spread(times):
    if not times:
        return None
    return {"min": min(times), "median": float(np.median(times)), "max": max(times)}


def traffic_for_build():
    """The measured traffic constant, if the kernel source is the one measured."""
    import hashlib
    src = open(os.path.join(ROOT, TRAFFIC["source"]), "rb").read()
    return TRAFFIC["bytes_per_launch"] if hashlib.sha256(src).hexdigest()[:16] == TRAFFIC["sha16"] else None


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def _oracle_timeslice(ts_config):
    from oracle import oracle as O
    return O.hyp(O.map_timeslice(ts_config, L), L, 0.62, 0.62, 3)


def cpu_baseline(ts_config, lamC, runs=3):
    """The oracle's Krylov-Schur (C, OpenMP over the host cores) solving one
    32^3 timeslice to 1e-10 with the plain drop-in's lambda_c, `runs` times;
    the median solve time gives timeslices/s."""
    from oracle import oracle as O
    U = _oracle_timeslice(ts_config)
    times, conv, matvecs = [], [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        nconv, theta, V, stats = O.eps_solve(U, L, LAMBDA_L, lamC, DEG, NEV, tol=TOL)
        times.append(time.perf_counter() - t0)
        conv.append(nconv)
        matvecs.append(stats[1])
    if min(conv) < NEV:
        raise RuntimeError(f"cpu baseline: oracle solve converged {min(conv)} of {NEV} pairs")
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    med = float(np.median(times))
    return {"value": 1.0 / med, "unit": "timeslices/s", "cores": cores, "kind": "port",
            "sample": f"one 32^3 timeslice solved to {TOL} (N_ev {NEV}, T_{DEG}(B), ncv 440, {matvecs[0]} filtered "
                      f"matvecs) with OpenMP over {cores} threads, median of {runs} solves "
                      f"({', '.join(f'{t:.1f}' for t in times)} s); the one-timeslice-per-core plan over all 64 "
                      f"timeslices is bench.py --cpu-full (not in the default run: minutes of host time)",
            **host_info()}


def cpu_worker(spec):
    """One timeslice of --cpu-full: single-threaded oracle solve, times on stdout."""
    from oracle import oracle as O
    path, lam_c = spec.split(",")
    cfg = np.load(path)
    U = _oracle_timeslice(cfg)
    t0 = time.perf_counter()
    nconv, theta, V, stats = O.eps_solve(U, L, LAMBDA_L, float(lam_c), DEG, NEV, tol=TOL)
    print(json.dumps({"seconds": time.perf_counter() - t0, "nconv": nconv, "matvecs": stats[1]}))


def cpu_full(cores, runs, nts, threads=1):
    """The CPU baseline as planned: timeslices spread over the host cores
    like the reference's ranks (ev_ts.cpp:96-126), `threads` OpenMP threads
    each, every one solved to 1e-10 by the oracle; per run, cores/threads
    timeslices start together and the rate is their count / the time until
    the last finishes (process start and HYP included)."""
    import subprocess
    import tempfile
    from oracle import lapheig_np as N
    cores = cores or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ranks = max(1, min(nts, cores // max(1, threads)))
    lam_c = plain_lambda_c(L, NEV)
    vol = 72 * L ** 3
    cfg = N.synth_config(SEED, EPS, L, ranks)
    tmp = tempfile.mkdtemp(prefix="laph_cpu_", dir=os.environ.get("TMPDIR", "/tmp"))
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    rates, per_ts = [], []
    try:
        paths = []
        for t in range(ranks):
            p = os.path.join(tmp, f"ts{t}.npy")
            np.save(p, cfg[t * vol:(t + 1) * vol])
            paths.append(p)
        for r in range(runs):
            t0 = time.perf_counter()
            procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", f"{p},{lam_c}"],
                                      stdout=subprocess.PIPE, text=True, env=env) for p in paths]
            res, left = [None] * len(procs), set(range(len(procs)))
            while left:  # report each timeslice as it finishes
                for k in sorted(left):
                    if procs[k].poll() is not None:
                        res[k] = json.loads(procs[k].communicate()[0])
                        left.discard(k)
                        print(f"[bench] cpu run {r}: timeslice {k} solved in {res[k]['seconds']:.1f}s "
                              f"({time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)
                time.sleep(0.2)
            wall = time.perf_counter() - t0
            assert all(x["nconv"] >= NEV for x in res), res
            rates.append(len(paths) / wall)
            per_ts += [x["seconds"] for x in res]
            print(f"[bench] cpu run {r}: {len(paths)} timeslices in {wall:.1f}s", file=sys.stderr, flush=True)
    finally:
        import shutil
        shutil.rmtree(tmp, ignore_errors=True)
    return {"value": float(np.median(rates)), "unit": "timeslices/s", "cores": ranks * threads, "kind": "port",
            "sample": f"{ranks} timeslices of the 32^3x64 field at once, {threads} OpenMP thread(s) each (the "
                      f"oracle's Krylov-Schur to {TOL}, lambda_c {lam_c:.4f}), median rate of {runs} runs (process "
                      "start and HYP included in each run's wall time)",
            "rates": rates, "seconds_per_timeslice": spread(per_ts), **host_info()}


if __name__ == "__main__":
    main()
