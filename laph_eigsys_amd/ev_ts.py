"""Per-configuration eigensystem program: the host side of ev_ts.cpp
(src/ev_ts.cpp:29-330) on one MI355X.

    python -m laph_eigsys_amd.ev_ts --conf 1000 [--phase_reset 1] [--params parameters.txt]
        [--solver batched|plain] [--tstart T0 --tend T1] [--group G] [--mem-gb M]
        [--tune-file PATH]

Reads <config_path>/conf.%04d (ILDG LIME, io.cpp:22-146), HYP-smears every
timeslice (alpha_1, alpha_2, iter), computes the NEV eigenpairs of the
covariant Laplacian of each timeslice nearest zero and writes, per timeslice t,
<result_path>/eigenvectors.%04d.%03d (column-major complex128, checked with
tr(V^H V) before writing), eigenvalues.%04d.%03d and phases.%04d.%03d
(read_write.cpp:176-201, 237-250), after fixing the phases
(eigensystem.cpp:4-30) or, with --phase_reset, resetting them to a stored
convention (eigensystem.cpp:32-57).

Solvers:
  plain    the drop-in, ev_ts.cpp:176-303 step for step: per timeslice a
           device context (map_timeslice_to_eigen, smearing_hyp), the shell
           operator T_DEG(B) with LAM_L / LAM_C from the parameter file,
           Krylov-Schur (HEP, largest magnitude, NEV), recover_spectrum.
  batched  the mixed-precision batched solver over groups of timeslices (the
           damped interval adapts per timeslice, LAM_C is not used); the
           group size fits the work buffers into --mem-gb of HBM, so 48^3 x 96
           streams through in groups of 24.  The first solve of a process
           times the filter and Gram layouts and the warm-up degree (several
           seconds and up to five solves at 32^3 x 64); the choice is kept in
           a small text file (LAPH_TUNE_FILE; by default laph_layouts.txt in
           the result directory, --tune-file to move it, 'none' to turn it
           off) so that the next configurations of a production run read it.

The reference divides LT over MPI ranks (ev_ts.cpp:96-126); one process
here takes [tstart, tend] (default: all of LT), the share one rank would get.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import lapheig as lp


class IO:
    """The parameter file of par_io.cpp:8-133: `key = value` lines, `#`
    comments; LY = LZ = LX; V3, MAT_ENTRIES, V_TS derived."""
    PATHS = ("config_path", "result_path", "in_path", "phase_path", "ev_path_1", "ev_path_2")
    INTS = ("LT", "LX", "NEV", "iter", "DEG", "OMP_THRDS")
    FLOATS = ("alpha_1", "alpha_1_i", "alpha_1_f", "alpha_2", "alpha_2_i", "alpha_2_f", "LAM_L", "LAM_C")

    def __init__(self):
        self.names = {k: "" for k in self.PATHS}
        self.ints = {k: 0 for k in self.INTS}
        self.floats = {k: 0.0 for k in self.FLOATS}

    def set_values(self, infile):
        with open(infile) as f:
            for line in f:
                line = line.rstrip("\n")
                if not line or line[0] == "#" or "=" not in line:
                    continue
                key, value = (s.strip() for s in line.split("=", 1))
                if key in self.names:
                    self.names[key] = value
                elif key in self.ints:
                    self.ints[key] = int(value.split()[0])
                elif key in self.floats:
                    self.floats[key] = float(value.split()[0])
        return self

    def get_int(self, spec):
        LX = self.ints["LX"]
        derived = {"LY": LX, "LZ": LX, "NDIR": 4, "NCOL": 3, "V3": LX ** 3, "MAT_ENTRIES": 3 * LX ** 3,
                   "V_TS": 72 * LX ** 3, "V4_LIME": 72 * LX ** 3 * self.ints["LT"]}
        return derived[spec] if spec in derived else self.ints[spec]

    def get_float(self, spec):
        return self.floats[spec]

    def get_path(self, spec):
        # par_io.cpp get_path: "conf" / "res" / "in" / "phase_path" ...
        alias = {"conf": "config_path", "res": "result_path", "in": "in_path"}
        return self.names[alias.get(spec, spec)]

    def print_summary(self, out=sys.stdout):
        g = self.get_int
        print("Configuration Summary for the eigensystem calculation:", file=out)
        print(f"Lattice size is: {g('LT')} x {g('LX')} x {g('LY')} x {g('LZ')}", file=out)
        print(f"Number of eigenvectors is: {g('NEV')}", file=out)
        print(f"Hyp-smearing parameters (alpha_1, alpha_2, iter): {self.floats['alpha_1']}, "
              f"{self.floats['alpha_2']}, {g('iter')}", file=out)
        print(f"Using Chebyshev parameters: Lambda_l = {self.floats['LAM_L']} Lambda_c = {self.floats['LAM_C']}",
              file=out)
        print(f"Degree of Chebyshev polynomial: {g('DEG')}", file=out)
        print(f"Configurations are read in from: {self.names['config_path']}", file=out)
        print(f"Results are stored in: {self.names['result_path']}", file=out)
        print(f"Entries in Matrix: {g('MAT_ENTRIES')}", file=out)


def batch_bytes_per_timeslice(L, nev, nsub=0):
    """Device bytes of one timeslice in a batch: three FP64 blocks, four FP32
    filter buffers, twelve K x K FP64 matrices (Gram, Rayleigh quotient, the
    inverse Cholesky factor, scratch, eight partial Grams), FP64 links and two
    FP32 copies (as read, and scaled for the filter's window)."""
    n = 3 * L ** 3
    K = nsub if nsub > 0 else nev + max(16, nev // 4)
    K = (K + 63) // 64 * 64
    return n * K * (3 * 16 + 4 * 8) + 12 * K * K * 16 + L ** 3 * 27 * 36


def plan_groups(todo, L, nev, mem_gb, group=0, nsub=0):
    """Contiguous groups of timeslices of balanced size whose work buffers
    fit into mem_gb (or of the given size)."""
    if todo <= 0:
        return []
    if group <= 0:
        gmax = max(1, int(mem_gb * 1e9 // batch_bytes_per_timeslice(L, nev, nsub)))
        ngroups = math.ceil(todo / gmax)
        group = math.ceil(todo / ngroups)
    out, t = [], 0
    while t < todo:
        out.append((t, min(group, todo - t)))
        t += group
    return out


def finish_timeslice(pars, config, t, ev, V, phase_reset, log, batch=None, k=None):
    """ev_ts.cpp:296-325 for one timeslice: phase fix (or reset), the three
    files, the trace prints.  V: (nev, n) with row i = eigenvector i, or None
    with `batch`/`k`: the solved batch writes its timeslice k straight from
    device memory (phase fix and the V^H V of the guard on the device).

    With phase_reset the files hold the eigenvectors in the old phase
    convention, arg V(0, n) = old phase n, and the phases file the old
    phases.  In the reference the fix_phase call that follows reset_phase
    (ev_ts.cpp:302-309) overwrites the reset, so its files are phase-fixed
    either way; that defect is not reproduced."""
    nev = pars.get_int("NEV")
    res = pars.get_path("res")
    if phase_reset:
        old = lp.read_eigenvalues_bin(pars.get_path("phase_path"), "phases", config, t, nev)
        if V is None:
            ev, V, _, _ = batch.get(k)
        Vr = lp.reset_phase(V, old)
        lp.write_eig_sys_bin(res, "eigenvectors", config, t, Vr)
        lp.write_eigenvalues_bin(res, "eigenvalues", config, t, ev)
        lp.write_eigenvalues_bin(res, "phases", config, t, old)
        log(f"{t}: phases reset to {pars.get_path('phase_path')}/phases.{config:04d}.{t:03d}")
        return
    if V is None:
        trc, sm = batch.write(k, res, config, t)
        log(f"{t}: tr(V.adj() * V) = {trc}  sum(V.adj() * V) = {sm}")
        return
    Vf, phase = lp.fix_phase(V)
    lp.write_eig_sys_bin(res, "eigenvectors", config, t, Vf)
    lp.write_eigenvalues_bin(res, "eigenvalues", config, t, ev)
    lp.write_eigenvalues_bin(res, "phases", config, t, phase)
    ok, trc, sm = lp.check_trace(V)
    log(f"{t}: tr(V.adj() * V) = {trc}  sum(V.adj() * V) = {sm}  sum(V) = {complex(V.sum())}")


def run(pars, config, solver="batched", tstart=0, tend=None, group=0, mem_gb=200.0, nsub=0, tol=1e-10,
        phase_reset=False, device=0, log=print):
    """All timeslices [tstart, tend] of configuration `config`; returns a
    dict of timings (seconds) and per-timeslice cycle counts."""
    LT, L, nev, deg = (pars.get_int(k) for k in ("LT", "LX", "NEV", "DEG"))
    a1, a2, it = pars.get_float("alpha_1"), pars.get_float("alpha_2"), pars.get_int("iter")
    lam_l, lam_c = pars.get_float("LAM_L"), pars.get_float("LAM_C")
    tend = LT - 1 if tend is None else tend
    if not (0 <= tstart <= tend < LT):
        raise ValueError(f"timeslice range [{tstart}, {tend}] outside [0, {LT - 1}]")
    todo = tend - tstart + 1
    vts = pars.get_int("V_TS")
    os.makedirs(pars.get_path("res"), exist_ok=True)
    conf_name = f"{pars.get_path('conf')}/conf.{config:04d}"
    t0 = time.perf_counter()
    configuration = lp.read_lime_gauge_field_doubleprec_timeslices(conf_name, LT, L, tstart, tend)
    out = {"read_seconds": time.perf_counter() - t0, "solve_seconds": 0.0, "write_seconds": 0.0,
           "timeslices": todo, "solver": solver, "groups": []}
    log(f"init time {out['read_seconds']:.3f} s")
    if solver == "plain":
        for ts in range(todo):
            t = tstart + ts
            tp = time.perf_counter()
            with lp.Tslice(L, deg, lam_l, lam_c, device=device) as sl:
                sl.map_timeslice_to_eigen(configuration[ts * vts:(ts + 1) * vts])
                sl.smearing_hyp(a1, a2, it)
                theta, V, st = sl.eps_solve(nev, tol=tol)
            dt = time.perf_counter() - tp
            out["solve_seconds"] += dt
            log(f"{t}: solving time {dt:.3f} s, restarts {st.restarts}, converged {st.nconv}")
            ev = lp.recover_spectrum(theta, deg, lam_l, lam_c)
            tw = time.perf_counter()
            finish_timeslice(pars, config, t, ev, V, phase_reset, log)
            out["write_seconds"] += time.perf_counter() - tw
        return out
    if solver != "batched":
        raise ValueError(f"unknown solver {solver!r}")
    b, bsize = None, 0
    try:
        for g0, gn in plan_groups(todo, L, nev, mem_gb, group, nsub):
            if gn != bsize:
                if b is not None:
                    b.close()
                b = lp.Batch(L, deg, lam_l, 0.0, gn, nev, nsub=nsub, tol=tol, device=device)
                bsize = gn
            tp = time.perf_counter()
            for k in range(gn):
                ts = g0 + k
                b.set_timeslice(k, configuration[ts * vts:(ts + 1) * vts], a1, a2, it)
            st = b.solve()
            dt = time.perf_counter() - tp
            out["solve_seconds"] += dt
            out["groups"].append({"tstart": tstart + g0, "timeslices": gn, "seconds": dt, "cycles": st.iterations,
                                  "max_residual": st.max_residual, "layout_from_cache": st.layout_from_cache,
                                  "warmup_degree": st.warmup_degree})
            log(f"group t={tstart + g0}..{tstart + g0 + gn - 1}: {dt:.2f} s, cycles {st.iterations}, "
                f"max residual {st.max_residual:.2e}")
            tw = time.perf_counter()
            for k in range(gn):
                finish_timeslice(pars, config, tstart + g0 + k, None, None, phase_reset, log, batch=b, k=k)
            out["write_seconds"] += time.perf_counter() - tw
    finally:
        if b is not None:
            b.close()
    return out


def tune_file_path(arg, pars, env):
    """Layout cache of a run: --tune-file when given ('none' turns the cache
    off), else LAPH_TUNE_FILE when set, else laph_layouts.txt in the result
    directory, so that the configurations of one production run (one process
    each, like the reference's) time the layouts and the warm-up degree once."""
    if arg == "none":
        return None
    if arg:
        return arg
    if env.get("LAPH_TUNE_FILE"):
        return env["LAPH_TUNE_FILE"]
    return os.path.join(pars.get_path("res"), "laph_layouts.txt")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--conf", type=int, required=True, help="trajectory number of the gauge configuration")
    ap.add_argument("--phase_reset", type=int, default=0,
                    help="write the eigenvectors with the phases of <phase_path>/phases.%%04d.%%03d")
    ap.add_argument("--params", default="parameters.txt")
    ap.add_argument("--solver", choices=("batched", "plain"), default="batched")
    ap.add_argument("--tstart", type=int, default=0)
    ap.add_argument("--tend", type=int, default=None)
    ap.add_argument("--group", type=int, default=0, help="timeslices per batched solve (0: fit --mem-gb)")
    ap.add_argument("--mem-gb", type=float, default=200.0)
    ap.add_argument("--nsub", type=int, default=0, help="search-space width of the batched solver")
    ap.add_argument("--tol", type=float, default=1e-10, help="residual |D v - lambda v| (batched), EPS tol (plain)")
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--json", default=None, help="write the timing summary here")
    ap.add_argument("--tune-file", default="",
                    help="layout cache of the batched solver (read, and appended after timing a new shape); "
                         "default <result_path>/laph_layouts.txt, 'none': time the layouts in every run")
    args = ap.parse_args(argv)
    pars = IO().set_values(args.params)
    tune = tune_file_path(args.tune_file, pars, os.environ)
    if tune:
        os.environ["LAPH_TUNE_FILE"] = tune
    else:
        os.environ.pop("LAPH_TUNE_FILE", None)
    pars.print_summary()
    print(f"calculating config {args.conf}")
    out = run(pars, args.conf, args.solver, args.tstart, args.tend, args.group, args.mem_gb, args.nsub, args.tol,
              bool(args.phase_reset), args.device, log=lambda m: print(m, flush=True))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f)
    print(json.dumps({k: v for k, v in out.items() if k != "groups"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
