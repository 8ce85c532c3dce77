"""Train the warm-start network on two-galaxy collisions and measure what it gains.

    python -m galaxy_collision_amd.train_warmstart [--grid 256] [--pool 4] [--n-per 5000000]
        [--runs 5] [--steps 80] [--epochs 40] [--eval-steps 40] [--out weights.npz]
        [--json report.json] [--profile] [--plain-tol 1e-3 ...]

1. Pairs: `runs` collisions of the same recipe as the benchmark, each with its own seed,
   disk inclinations and mass ratio (1:1 to 1:4), are stepped `steps` times by the PM chain
   with the extrapolated mixed-precision multigrid; every solve writes its training pair
   (network input and the pooled error of the extrapolated guess, see pm_mg_solve_warm).
2. Training: Adam on the network's own forward/backward kernels; every 10th pair is held
   out for validation.
3. A/B: the benchmark collision (seed 0, inclinations 30/-45, 1:1) is stepped with each
   warm start (previous phi as in the reference, time extrapolation, extrapolation plus the
   network) and with the float32 reference multigrid; V-cycles per step, wall time per step,
   the kept fraction of network corrections and the phi error of the last step against the
   exact discrete solution are reported. All multigrid variants stop on the same estimated
   error, so they deliver the same accuracy.

No reference counterpart (the learned warm start is an addition).
"""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.accuracy import phi_rel_error_discrete
from galaxy_collision_amd.chain import PMChain
from galaxy_collision_amd.data import ParticleState
from galaxy_collision_amd.ic import sort_by_cell, two_galaxy_ic
from galaxy_collision_amd.solver import make_solver
from galaxy_collision_amd.warmstart import PairSet, WarmStartNet, relative_error, train

# (seed, inclinations, stellar masses): varied seeds, orientations and mass ratios 1:1..1:4
TRAIN_RUNS = [
    (1, (30.0, -45.0), (1.0, 1.0)),
    (2, (10.0, 60.0), (1.0, 0.5)),
    (3, (-20.0, 35.0), (1.0, 0.25)),
    (4, (45.0, -15.0), (1.0, 0.75)),
    (5, (0.0, 80.0), (1.0, 0.33)),
]
BENCH_RUN = (0, (30.0, -45.0), (1.0, 1.0))


def log(msg: str) -> None:
    print(f"[warmstart] {msg}", file=sys.stderr, flush=True)


def collision(n_per: int, grid: int, box: float, run) -> tuple[ParticleState, float]:
    seed, inc, masses = run
    dx = box / grid
    c = (grid - 1) * dx / 2.0
    icr = two_galaxy_ic(n_per, box_centre=(c, c, c), separation=20.0, inclinations=inc,
                        masses=masses, seed=seed)
    icr = sort_by_cell(icr, grid, dx)
    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    return parts, dx


def collect_pairs(args, pairs: PairSet) -> None:
    for run in TRAIN_RUNS[: args.runs]:
        t0 = time.perf_counter()
        parts, dx = collision(args.n_per, args.grid, args.box, run)
        solver = make_solver("multigrid_mp", args.grid, dx=dx, grav_constant=1.0,
                             warm="extrapolate", pool=args.pool)
        chain = PMChain(parts, args.grid, dx, solver)
        chain.prime()
        solver.pairs = pairs
        for _ in range(args.steps):
            chain.step(args.dt)
        solver.pairs = None
        torch.cuda.synchronize()
        log(f"run {run}: {len(pairs)} pairs so far ({time.perf_counter() - t0:.1f} s)")
        del chain, solver, parts
        torch.cuda.empty_cache()


def ab_test(args, net: WarmStartNet | None) -> dict:
    variants = [("multigrid_mp", "previous"), ("multigrid_mp", "extrapolate")]
    if net is not None:
        variants.append(("multigrid_mp", "learned"))
    for tol in args.plain_tol:
        variants.append(("multigrid", tol))
    out = {}
    for name, warm in variants:
        parts, dx = collision(args.n_per, args.grid, args.box, BENCH_RUN)
        kw = {"dx": dx, "grav_constant": 1.0}
        if name == "multigrid_mp":
            kw.update(warm=warm, pool=args.pool, net=net if warm == "learned" else None)
        else:
            kw.update(tol=warm, n_cycles=args.plain_cycles)
        solver = make_solver(name, args.grid, **kw)
        chain = PMChain(parts, args.grid, dx, solver)
        chain.prime()
        for _ in range(3):
            chain.step(args.dt)
        torch.cuda.synchronize()
        chain.last_cycles.clear()
        used, ratios = [], []
        t0 = time.perf_counter()
        for _ in range(args.eval_steps):
            chain.step(args.dt)
            w = getattr(solver, "last_warm", None)
            if w and warm == "learned" and w["extrapolated"]:
                used.append(w["used_net"])
                ratios.append(w["resid_after"] / w["resid_before"])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernels = None
        if args.profile:
            nat.timing_reset()
            nat.timing_enable(True)
            for _ in range(args.eval_steps):
                chain.step(args.dt)
            nat.timing_enable(False)
            rep = nat.timing_report()
            kernels = {k: round(v["total_ms"] / args.eval_steps * 1e3, 1)
                       for k, v in sorted(rep.items(), key=lambda kv: -kv[1]["total_ms"])[:24]}
        err = phi_rel_error_discrete(chain.grid.rho.to_numpy(), chain.grid.phi.to_numpy(), dx, 1.0)
        key = f"{name}:{warm}" if name == "multigrid_mp" else f"{name}:tol{warm:g}"
        out[key] = {
            "vcycles_per_step": float(np.mean(chain.last_cycles[: args.eval_steps])),
            "ms_per_step": wall / args.eval_steps * 1e3,
            "phi_rel_l2_vs_exact_discrete": err,
        }
        if kernels:
            out[key]["us_per_step_by_kernel"] = kernels
        if used:
            out[key]["net_kept_fraction"] = float(np.mean(used))
            out[key]["pooled_residual_ratio_mean"] = float(np.mean(ratios))
        log(f"A/B {key}: {out[key]}")
        del chain, solver, parts
        torch.cuda.empty_cache()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--box", type=float, default=40.0)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--pool", type=int, default=4)
    ap.add_argument("--n-per", type=int, default=5_000_000)
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--eval-steps", type=int, default=40)
    ap.add_argument("--plain-tol", type=float, nargs="*", default=[1e-3],
                    help="tolerances of the float32 reference multigrid to compare")
    ap.add_argument("--profile", action="store_true",
                    help="repeat each variant with per-kernel timing")
    ap.add_argument("--plain-cycles", type=int, default=20)
    ap.add_argument("--weights", default=None, help="load instead of training")
    ap.add_argument("--out", default=None, help="save the trained weights here (.npz)")
    ap.add_argument("--json", default=None, help="write the report here")
    args = ap.parse_args(argv)
    nat.require_device()
    torch.cuda.set_device(0)
    report = {"config": vars(args)}
    res = args.grid // args.pool
    if args.weights:
        net = WarmStartNet.load(args.weights)
    else:
        pairs = PairSet(res)
        collect_pairs(args, pairs)
        val = list(range(9, len(pairs), 10))
        net = WarmStartNet(res, seed=0)
        t0 = time.perf_counter()
        hist = train(net, pairs, epochs=args.epochs, batch=args.batch, lr=args.lr, val_idx=val,
                     log=lambda r: log(f"epoch {r}"))
        torch.cuda.synchronize()
        train_s = time.perf_counter() - t0
        report["train"] = {"pairs": len(pairs), "val_pairs": len(val), "seconds": train_s,
                           "history": hist,
                           "train_rel_err": relative_error(net, pairs, [i for i in range(len(pairs))
                                                                        if i not in set(val)][:20]),
                           "val_rel_err": relative_error(net, pairs, val)}
        log(f"trained on {len(pairs)} pairs in {train_s:.1f} s: {report['train']['val_rel_err']:.3f} "
            "validation relative error")
        if args.out:
            net.save(args.out)
        del pairs
        torch.cuda.empty_cache()
    report["ab"] = ab_test(args, net)
    text = json.dumps(report)
    if args.json:
        with open(args.json, "w") as f:
            f.write(text)
    print(text, flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
