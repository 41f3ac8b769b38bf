"""HYP parameter scan: the host program of tune_hyp.cpp:23-271 on one MI355X.

    python -m laph_eigsys_amd.tune_hyp --conf 1000 [--params parameters.txt] [--ts 10]
        [--solver batched|plain] [--group G] [--mem-gb M]

For every (alpha_1, alpha_2) of the grid alpha_1_i .. alpha_1_f x alpha_2_i ..
alpha_2_f in steps of 0.05 (the bounds from the parameter file; the grid
accumulates the step in double precision and keeps points <= the upper bound,
as the loops of tune_hyp.cpp:116-117 do), timeslice `ts` of configuration
`conf` is HYP-smeared with (alpha_1, alpha_2, iter), its NEV eigenpairs are
computed and phase-fixed, and

    <result_path>/eigenvalues_hyp-<iter>_a1-<a1>_a2-<a2>.%04d.%03d
    <result_path>/phases_hyp-<iter>_a1-<a1>_a2-<a2>.%04d.%03d

are written (tune_hyp.cpp:219-226); the table "#alpha1 alpha2 lowest_ev" of
the lowest eigenvalue per point closes the output (tune_hyp.cpp:239-243).

With the batched solver every grid point is one timeslice of a batch (the
same links smeared differently), so the whole scan is one batched solve (or a
few, when the grid outgrows --mem-gb).  The plain solver runs the drop-in
per point with LAM_C from the parameter file, as the reference does.
"""
import argparse
import sys
import time

from . import lapheig as lp
from .ev_ts import IO, plan_groups


def alpha_grid(lo, hi, step=0.05):
    """for (double a = lo; a <= hi; a += step) -- the values the reference's
    loop visits, rounding of the accumulated step included."""
    out, a = [], lo
    while a <= hi:
        out.append(a)
        a += step
    return out


def tune_hyp(pars, config, ts=10, solver="batched", group=0, mem_gb=200.0, tol=1e-10, device=0, log=print):
    """Returns [(a1, a2, lowest eigenvalue)] in the order of the reference's loops."""
    LT, L, nev, deg, it = (pars.get_int(k) for k in ("LT", "LX", "NEV", "DEG", "iter"))
    lam_l, lam_c = pars.get_float("LAM_L"), pars.get_float("LAM_C")
    pts = [(a1, a2) for a1 in alpha_grid(pars.get_float("alpha_1_i"), pars.get_float("alpha_1_f"))
           for a2 in alpha_grid(pars.get_float("alpha_2_i"), pars.get_float("alpha_2_f"))]
    if not pts:
        raise ValueError("empty alpha grid")
    if not 0 <= ts < LT:
        raise ValueError(f"timeslice {ts} outside [0, {LT - 1}]")
    cfg = lp.read_lime_gauge_field_doubleprec_timeslices(f"{pars.get_path('conf')}/conf.{config:04d}", LT, L, ts, ts)
    res = pars.get_path("res")
    rows = []

    def store(a1, a2, ev, phase):
        lp.write_eigenvalues_bin(res, f"eigenvalues_hyp-{it}_a1-{a1:.2f}_a2-{a2:.2f}", config, ts, ev)
        lp.write_eigenvalues_bin(res, f"phases_hyp-{it}_a1-{a1:.2f}_a2-{a2:.2f}", config, ts, phase)
        rows.append((a1, a2, float(ev[0])))

    t0 = time.perf_counter()
    if solver == "plain":
        for a1, a2 in pts:
            with lp.Tslice(L, deg, lam_l, lam_c, device=device, ts_config=cfg) as sl:
                sl.smearing_hyp(a1, a2, it)
                theta, V, st = sl.eps_solve(nev, tol=tol)
            _, phase = lp.fix_phase(V)
            store(a1, a2, lp.recover_spectrum(theta, deg, lam_l, lam_c), phase)
            log(f"a1 {a1:.2f} a2 {a2:.2f}: restarts {st.restarts}, lowest {rows[-1][2]:.10f}")
    elif solver == "batched":
        b, bsize = None, 0
        try:
            for g0, gn in plan_groups(len(pts), L, nev, mem_gb, group):
                if gn != bsize:
                    if b is not None:
                        b.close()
                    b = lp.Batch(L, deg, lam_l, 0.0, gn, nev, tol=tol, device=device)
                    bsize = gn
                for k in range(gn):
                    a1, a2 = pts[g0 + k]
                    b.set_timeslice(k, cfg, a1, a2, it)
                st = b.solve()
                log(f"points {g0}..{g0 + gn - 1}: cycles {st.iterations}, max residual {st.max_residual:.2e}")
                for k in range(gn):
                    ev, _, phase, _ = b.get(k, want_vectors=False)
                    store(*pts[g0 + k], ev, phase)
        finally:
            if b is not None:
                b.close()
    else:
        raise ValueError(f"unknown solver {solver!r}")
    log(f"{len(pts)} grid points in {time.perf_counter() - t0:.2f} s")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--conf", type=int, required=True)
    ap.add_argument("--params", default="parameters.txt")
    ap.add_argument("--ts", type=int, default=10, help="timeslice (tune_hyp.cpp:107 uses 10)")
    ap.add_argument("--solver", choices=("batched", "plain"), default="batched")
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--mem-gb", type=float, default=200.0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    pars = IO().set_values(args.params)
    pars.print_summary()
    rows = tune_hyp(pars, args.conf, args.ts, args.solver, args.group, args.mem_gb, args.tol, args.device,
                    log=lambda m: print(m, flush=True))
    print("#alpha1\talpha2\tlowest_ev")
    for a1, a2, lo in rows:
        print(f"{a1:g}\t{a2:g}\t{lo:.10g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
