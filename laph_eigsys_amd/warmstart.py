"""Learned warm start for the batched solver: training on the solver's own
eigenvectors and evaluation against the random start.

Configuration C5 of the per-timeslice path: a 4-layer gauge-covariant
parallel-transport network (16 complex channels per group, BF16 activations,
FP32 arithmetic; laph_eigsys_amd/csrc/net.hip) maps Gaussian colour vectors
to the starting block.  Training timeslices come from seeded fields with
eps in [0.25, 0.45]; the evaluation uses held-out seeds and compares, at the
same residual tolerance, the cycles, the stencil applications and the wall
time of the solve from a random start and from the network's proposal, and
checks that both give the same eigenvalues.

    python -m laph_eigsys_amd.warmstart --L 24 --nev 120 --deg 12 \\
        --train-seeds 20 --eval-seeds 5 --ts-per-seed 4 --steps 150 --out result.json
"""
import argparse
import json
import time

import numpy as np

from . import lapheig as lp


def field_timeslices(seed, eps, L, nts):
    """The first nts timeslices of the seeded field (the generator's counter
    includes t, so they are the same timeslices as those of a longer field)."""
    return lp.synth_config(seed, eps, L, nts)


def make_batch(L, deg, nev, cfgs, nts_each, tol, nsub, device, warm):
    nts = len(cfgs) * nts_each
    b = lp.Batch(L, deg, 12.0, 0.0, nts, nev, nsub=nsub, tol=tol, seed=17, device=device, warm_start=warm)
    vol = 72 * L ** 3
    k = 0
    for cfg in cfgs:
        for t in range(nts_each):
            b.set_timeslice(k, cfg[t * vol:(t + 1) * vol], 0.62, 0.62, 3)
            k += 1
    return b


def eps_of(seed, lo=0.25, hi=0.45, n=9):
    return lo + (hi - lo) * (seed % n) / (n - 1)


def train(net, L, nev, deg, seeds, ts_per_seed, steps, lr, tol, nsub, device=0, log=None):
    cfgs = [field_timeslices(s, eps_of(s), L, ts_per_seed) for s in seeds]
    b = make_batch(L, deg, nev, cfgs, ts_per_seed, tol, nsub, device, warm=0)
    t0 = time.time()
    st = b.solve()
    t_solve = time.time() - t0
    losses = []
    t0 = time.time()
    for i in range(steps):
        lr_i = lr * (0.1 ** (i / max(1, steps)))  # exponential decay to lr/10
        losses.append(net.train_step(b, seed=1000 + i, lr=lr_i))
        if log and (i % 25 == 0 or i == steps - 1):
            log(f"[train] step {i} loss {losses[-1]:.5f}")
    t_train = time.time() - t0
    b.close()
    return {"timeslices": len(seeds) * ts_per_seed, "solve_seconds": t_solve, "cycles": st.iterations,
            "train_seconds": t_train, "steps": steps, "loss_first": losses[0], "loss_last": losses[-1]}


MODES = (("random", 1), ("learned", 1), ("learned_no_warmup", 2))


def evaluate(net, L, nev, deg, seeds, ts_per_seed, tol, nsub, device=0, reps=1, modes=MODES):
    """Cycles, stencil applications and wall time of the solve (proposal
    included) per start mode, at the same residual bound: "random" = the
    seeded Gaussian block; "learned" = the network's proposal, Rayleigh-Ritz on
    it, then the usual warm-up cycle; "learned_no_warmup" = the proposal's Ritz
    values set the first window and the warm-up cycle is skipped."""
    cfgs = [field_timeslices(s, eps_of(s), L, ts_per_seed) for s in seeds]
    rows = {}
    for mode, warm in modes:
        b = make_batch(L, deg, nev, cfgs, ts_per_seed, tol, nsub, device, warm=warm)
        b.solve()  # warm-up (graph capture, library initialisation)
        acc = {"seconds": 0.0, "cycles": 0, "stencil32": 0, "stencil64": 0, "max_residual": 0.0}
        for r in range(reps):
            t0 = time.perf_counter()
            if mode != "random":
                net.propose(b, seed=4242 + r)
            st = b.solve()
            acc["seconds"] += time.perf_counter() - t0
            acc["cycles"] += st.iterations
            acc["stencil32"] += st.stencil_vecs
            acc["stencil64"] += st.stencil64_vecs
            acc["max_residual"] = max(acc["max_residual"], st.max_residual)
            assert st.converged == b.nts
        for k in ("seconds", "cycles", "stencil32", "stencil64"):
            acc[k] /= reps
        acc["evals"] = [b.get(t, want_vectors=False)[0] for t in range(b.nts)]
        rows[mode] = acc
        b.close()
    out = {m: {k: v for k, v in rows[m].items() if k != "evals"} for m in rows}
    out["timeslices"] = len(seeds) * ts_per_seed
    r = rows["random"]
    for m in rows:
        if m == "random":
            continue
        w = rows[m]
        out[m]["max_rel_eigenvalue_difference"] = max(
            float(np.max(np.abs(a - c) / np.maximum(np.abs(c), 1e-300))) for a, c in zip(w["evals"], r["evals"]))
        out[m]["matvec_reduction"] = 1.0 - (w["stencil32"] + w["stencil64"]) / (r["stencil32"] + r["stencil64"])
        out[m]["time_reduction"] = 1.0 - w["seconds"] / r["seconds"]
        out[m]["lambda_0_last"] = [float(w["evals"][0][0]), float(w["evals"][0][-1])]
    out["random"]["lambda_0_last"] = [float(r["evals"][0][0]), float(r["evals"][0][-1])]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--L", type=int, default=24)
    ap.add_argument("--nev", type=int, default=120)
    ap.add_argument("--deg", type=int, default=12)
    ap.add_argument("--nsub", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--train-seeds", type=int, default=20)
    ap.add_argument("--eval-seeds", type=int, default=5)
    ap.add_argument("--ts-per-seed", type=int, default=4)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--save", default=None, help="write the trained parameters (.npy)")
    ap.add_argument("--load", default=None, help="start from saved parameters (.npy)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--net-target", choices=("nev", "all"), default="nev",
                    help="train towards the span of the nev eigenvectors or of all K Ritz vectors")
    args = ap.parse_args()
    if args.net_target == "all":
        import os
        os.environ["LAPH_NET_TARGET"] = "1"  # read when the training batch is created
    log = lambda m: print(m, flush=True)  # noqa: E731
    net = lp.Net(layers=args.layers, seed=7)
    if args.load:
        net.set_params(np.load(args.load))
    train_seeds = list(range(3000, 3000 + args.train_seeds))
    eval_seeds = list(range(5000, 5000 + args.eval_seeds))
    res = {"config": vars(args)}
    res["untrained"] = evaluate(net, args.L, args.nev, args.deg, eval_seeds, args.ts_per_seed, args.tol, args.nsub,
                                reps=args.reps)
    if args.steps > 0:
        res["train"] = train(net, args.L, args.nev, args.deg, train_seeds, args.ts_per_seed, args.steps, args.lr,
                             args.tol, args.nsub, log=log)
    if args.save:
        np.save(args.save, net.get_params())
    res["eval"] = evaluate(net, args.L, args.nev, args.deg, eval_seeds, args.ts_per_seed, args.tol, args.nsub,
                           reps=args.reps)
    s = json.dumps(res)
    print(s)
    if args.out:
        with open(args.out, "w") as f:
            f.write(s + "\n")


if __name__ == "__main__":
    main()
