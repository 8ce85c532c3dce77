"""Train the warm-start network on two-galaxy collisions and measure what it gains.

    python -m galaxy_collision_amd.train_warmstart [--grid 256] [--pool 4] [--n-per 5000000]
        [--runs 5] [--steps 80] [--epochs 40] [--eval-steps 40] [--out weights.npz]
        [--json report.json] [--profile] [--plain-tol 1e-3 ...]

1. Pairs: `runs` collisions of the same recipe as the benchmark (drawn on the device, disks
   launched on the grid's rotation curve plus the recipe's 10% dispersion), each with its
   own seed, disk inclinations and mass ratio (1:1 to 1:4), are stepped `steps` times by the
   PM chain with the extrapolated mixed-precision multigrid; every solve writes its
   training pair (network input and the pooled error of the extrapolated guess, see
   pm_mg_solve_warm). The defaults (25 runs x 80 steps) give 2,000 pairs.
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
from galaxy_collision_amd.chain import PMChain
from galaxy_collision_amd.data import ParticleState
from galaxy_collision_amd.ic import sort_by_cell, two_galaxy_ic, two_galaxy_ic_tensors
from galaxy_collision_amd.solver import make_solver
from galaxy_collision_amd.warmstart import PairSet, WarmStartNet, relative_error, train

BENCH_RUN = (0, (30.0, -45.0), (1.0, 1.0))


def train_runs(count: int) -> list:
    """(seed, inclinations, stellar masses) of the training collisions: seeds 1.., disk
    inclinations drawn in [-80, 80] degrees, mass ratios spread over 1:1 .. 1:4."""
    rng = np.random.default_rng(12345)
    ratios = np.linspace(1.0, 0.25, count) if count > 1 else np.array([1.0])
    rng.shuffle(ratios)
    return [(k + 1, tuple(float(a) for a in np.round(rng.uniform(-80.0, 80.0, 2), 1)),
             (1.0, float(round(r, 3)))) for k, r in enumerate(ratios)]


def log(msg: str) -> None:
    print(f"[warmstart] {msg}", file=sys.stderr, flush=True)


def _equilibrated(icr, grid: int, dx: float, seed: int, parts=None) -> None:
    from galaxy_collision_amd.data import GridState
    from galaxy_collision_amd.ic import equilibrate_disk_velocities

    g = GridState(grid)
    solver = make_solver("multigrid_mp", grid, dx=dx, grav_constant=1.0)
    equilibrate_disk_velocities(icr, dx=dx, solver=solver, rho=g.rho, phi=g.phi, ax=g.ax,
                                ay=g.ay, az=g.az, parts=parts, dispersion=0.1, seed=seed)


def collision(n_per: int, grid: int, box: float, run, device_ic: bool = False
              ) -> tuple[ParticleState, float]:
    """The collision of `run` on the device; the benchmark's input (host generator) unless
    device_ic (the same recipe drawn with torch, for the many training runs)."""
    seed, inc, masses = run
    dx = box / grid
    c = (grid - 1) * dx / 2.0
    if device_ic:
        icr = two_galaxy_ic_tensors(n_per, box_centre=(c, c, c), separation=20.0,
                                    inclinations=inc, masses=masses, seed=seed,
                                    device=torch.device("cuda"), sort_cells=(grid, dx))
        parts = ParticleState(icr.n)
        parts.load_tensors(icr.pos, icr.vel, icr.mass, icr.gid)
    else:
        icr = sort_by_cell(two_galaxy_ic(n_per, box_centre=(c, c, c), separation=20.0,
                                         inclinations=inc, masses=masses, seed=seed), grid, dx)
        parts = ParticleState(icr.n)
        parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    _equilibrated(icr, grid, dx, seed, parts)
    return parts, dx


def collect_pairs(args, pairs: PairSet) -> None:
    for run in train_runs(args.runs):
        t0 = time.perf_counter()
        parts, dx = collision(args.n_per, args.grid, args.box, run, device_ic=True)
        solver = make_solver("multigrid_mp", args.grid, dx=dx, grav_constant=1.0,
                             warm="extrapolate", pool=args.pool)
        chain = PMChain(parts, args.grid, dx, solver, reorder_every=4)
        chain.prime()
        solver.pairs = pairs
        chain.run(args.steps, args.dt)
        solver.pairs = None
        torch.cuda.synchronize()
        log(f"run {run}: {len(pairs)} pairs so far ({time.perf_counter() - t0:.1f} s)")
        del chain, solver, parts
        torch.cuda.empty_cache()


def ab_test(args, net: WarmStartNet | None) -> dict:
    """The benchmark collision from one initial state per variant, stepped with the fused
    production path (PMChain.run) at the benchmark's solver settings (--tol, --sweeps);
    the kept fraction of network corrections from a further stepped pass."""
    from galaxy_collision_amd.accuracy import phi_rel_error_discrete_device

    variants = [("multigrid_mp", "previous"), ("multigrid_mp", "extrapolate")]
    if net is not None:
        variants.append(("multigrid_mp", "learned"))
    for tol in args.plain_tol:
        variants.append(("multigrid", tol))
    if args.variants:
        variants = [v for v in variants if v[1] in args.variants or
                    (v[0] == "multigrid" and "plain" in args.variants)]
    out = {}
    for name, warm in variants:
        parts, dx = collision(args.n_per, args.grid, args.box, BENCH_RUN)
        kw = {"dx": dx, "grav_constant": 1.0}
        if name == "multigrid_mp":
            kw.update(warm=warm, pool=args.pool, net=net if warm == "learned" else None,
                      tol=args.tol, pre_sweeps=args.sweeps[0], post_sweeps=args.sweeps[1])
        else:
            kw.update(tol=warm, n_cycles=args.plain_cycles)
        solver = make_solver(name, args.grid, **kw)
        chain = PMChain(parts, args.grid, dx, solver, reorder_every=4)
        chain.prime()
        chain.run(3, args.dt)
        torch.cuda.synchronize()
        chain.last_cycles.clear()
        t0 = time.perf_counter()
        chain.run(args.eval_steps, args.dt)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        cycles = float(np.mean(chain.last_cycles))
        err = phi_rel_error_discrete_device(chain.grid.rho, chain.grid.phi, dx, 1.0)
        used, ratios = [], []
        if warm == "learned":
            for _ in range(10):
                chain.step(args.dt)
                w = solver.last_warm
                if w and w["extrapolated"]:
                    used.append(w["used_net"])
                    ratios.append(w["resid_after"] / w["resid_before"])
        kernels = None
        if args.profile:
            nat.timing_reset()
            nat.timing_enable(True)
            chain.run(args.eval_steps, args.dt)
            torch.cuda.synchronize()
            nat.timing_enable(False)
            rep = nat.timing_report()
            kernels = {k: round(v["total_ms"] / args.eval_steps * 1e3, 1)
                       for k, v in sorted(rep.items(), key=lambda kv: -kv[1]["total_ms"])[:32]}
        key = f"{name}:{warm}" if name == "multigrid_mp" else f"{name}:tol{warm:g}"
        out[key] = {"vcycles_per_step": cycles, "ms_per_step": wall / args.eval_steps * 1e3,
                    "phi_rel_l2_vs_exact_discrete": err}
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
    ap.add_argument("--runs", type=int, default=25)
    ap.add_argument("--steps", type=int, default=80)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=5e-4)
    ap.add_argument("--eval-steps", type=int, default=40)
    ap.add_argument("--tol", type=float, default=2e-6,
                    help="mixed-precision solver tolerance of the A/B (the benchmark's)")
    ap.add_argument("--sweeps", type=int, nargs=2, default=[1, 1],
                    help="pre/post sweeps of the A/B's mixed-precision solves (the benchmark's)")
    ap.add_argument("--plain-tol", type=float, nargs="*", default=[1e-3],
                    help="tolerances of the float32 reference multigrid to compare")
    ap.add_argument("--profile", action="store_true",
                    help="repeat each variant with per-kernel timing")
    ap.add_argument("--plain-cycles", type=int, default=20)
    ap.add_argument("--variants", nargs="*", default=None,
                    help="A/B variants to run: previous, extrapolate, learned, plain (default all)")
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
