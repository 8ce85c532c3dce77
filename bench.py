"""PM force-chain benchmark on one MI355X: KDK steps/s of the full particle-mesh step.

Workloads (seeded synthetic inputs, built on the host, resident in HBM before timing):
  collision10m  two-galaxy collision, 10M particles, 256^3 grid, box 40, dt 0.01,
                mixed-precision multigrid, V(1,1) cycles, to an estimated phi error of 2e-6
                (measured error against the exact discrete solution ~1e-7, the float32
                floor, reported with every run), warm-started by linear extrapolation of
                phi in time (default)
  collision2m   the same recipe with 2 x 1,048,576 particles on a 128^3 grid, box 40
  collision100m the full-scale case: 2 x 50M particles (2:1 disk mass ratio), 256^3
  plummer       131,072-particle Plummer sphere, 64^3 grid over a box of 20a, dt 0.01

A step is one fused KDK step (kick+drift, deposit, Poisson solve, -grad, gather+kick); the
timed steps run back to back, so one step's closing kick shares a particle pass with the next
step's opening kick and drift (bitwise the same results as separate steps).
The collision inputs are launched with their disks on the rotation curve the grid produces
plus the recipe's 10% dispersion (ic.equilibrate_disk_velocities, as reference
bench.py:119-123 does). Pass 1 times --steps steps with no instrumentation (the reported
value). Pass 2 repeats them with per-kernel HIP events on the launch stream, giving
per-kernel times, the roofline figures of the dominant kernel and of every finest-level
multigrid kernel. Pass 3 is the reference's per-stage protocol (bench.py:66-72, 176-186):
the unfused stages, each bracketed by device synchronizations. Pass 4 continues the run to
step --steady-after and times --steps more steps (the steady state after the particles have
mixed). The accuracy series restarts from the initial state and runs --accuracy-steps steps,
comparing phi and the particle forces every --accuracy-cadence steps with the exact float64
solution of the discrete problem and with the zero-padded 2n^3 float64 FFT (the 10M / 256^3
configuration's check). The A/B pass (collisions) restarts from the initial state once per
multigrid warm start (previous phi, extrapolation, extrapolation plus the trained network)
and for the float32 reference multigrid: the learned-vs-plain comparison. The CPU baseline
runs the reference algorithm (OpenMP C port: atomic CIC scatter, plain float32 multigrid tol
1e-3) on the same input with the same warmup / throughput / per-stage protocol.

Prints one JSON line. ``--gpus`` above 1 is refused: this build runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "galaxy-collision-with-cuda-fft_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published
REFERENCE_STEPS_PER_S = 24.8  # reference docs/performance.md:17 (RTX 3070, 256^3, 10M)
# reference docs/performance.md:28-32, 10M particles on 256^3, RTX 3070, ms per step
REFERENCE_STAGE_MS = {"deposit": 2.54, "solve": 27.8, "grad": 1.17, "gather": 5.88,
                      "integrate": 3.18}

WORKLOADS = {
    "collision10m": dict(kind="two_galaxy", n_per=5_000_000, grid=256, box=40.0, dt=0.01,
                         solver="multigrid_mp", warm="extrapolate", tol=2e-6, sweeps=(1, 1),
                         deposit="bricked", reorder_every=8),
    "collision2m": dict(kind="two_galaxy", n_per=1_048_576, grid=128, box=40.0, dt=0.01,
                        solver="multigrid"),
    "collision100m": dict(kind="two_galaxy", n_per=50_000_000, grid=256, box=40.0, dt=0.01,
                          solver="multigrid_mp", warm="extrapolate", masses=(1.0, 0.5),
                          tol=2e-6, sweeps=(1, 1), deposit="bricked", reorder_every=8),
    "plummer": dict(kind="plummer", n=131_072, grid=64, box=20.0, dt=0.01, solver="multigrid"),
}


def equilibrate(icr, w: dict, dx: float, grav: float, seed: int) -> None:
    """Disk velocities from the grid's rotation curve plus the recipe's 10% dispersion,
    measured with a cold solve of the workload's solver settings (before any timing)."""
    from galaxy_collision_amd.data import GridState
    from galaxy_collision_amd.ic import equilibrate_disk_velocities
    from galaxy_collision_amd.solver import make_solver

    grid = GridState(w["grid"])
    solver = make_solver("multigrid_mp", w["grid"], dx=dx, grav_constant=grav)
    equilibrate_disk_velocities(icr, dx=dx, solver=solver, rho=grid.rho, phi=grid.phi,
                                ax=grid.ax, ay=grid.ay, az=grid.az, dispersion=0.1, seed=seed)
    del grid, solver


def build_ic(w: dict, seed: int):
    from galaxy_collision_amd import ic

    n, box = w["grid"], w["box"]
    dx = box / n
    if w["kind"] == "plummer":
        c = (n // 2) * dx
        icr = ic.plummer_ic(w["n"], total_mass=1.0, scale_a=1.0, centre=(c, c, c), seed=seed)
    else:
        c = (n - 1) * dx / 2.0
        icr = ic.two_galaxy_ic(w["n_per"], box_centre=(c, c, c), separation=20.0, seed=seed,
                               masses=w.get("masses", (1.0, 1.0)))
    return ic.sort_by_cell(icr, n, dx), dx


# rocprofv3 kernel symbol of a timed kernel name (the finest-grid dispatch is the timed one)
_PROFILE_KERNELS = {"mg.rb_pass": "rb_pass_kernel", "gather_kick": "gather_kick4_kernel",
                    "gather_kick_drift": "gather_kick_drift4_kernel",
                    "mg.rbs": "rbs4_kernel<false>", "mg.residual": "residual_kernel",
                    "mg.prolong": "prolong_add_kernel", "deposit.piece": "piece_kernel",
                    "deposit.key": "key_kernel", "deposit.radix_hist": "radix_hist_kernel",
                    "deposit.radix_scatter": "radix_scatter_kernel",
                    "deposit.unit_sums": "unit_deposit_kernel<true, false>",
                    "deposit.combine": "brick_combine_kernel",
                    "deposit.brick_reduce": "brick_reduce_kernel"}


def measured_traffic(workload: str, grid: int, timer_name: str):
    """HBM bytes per dispatch of the timed kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (profiles/), or None."""
    path = ROOT / "profiles" / f"pm_step_{workload}_{grid}_hbm_traffic.json"
    sym = _PROFILE_KERNELS.get(timer_name.split("[")[0])
    if sym is None or not path.exists():
        return None
    table = json.loads(path.read_text()).get("hbm_bytes_per_dispatch", {})
    hits = [(int(k.rsplit("@", 1)[1]), v) for k, v in table.items() if k.rsplit("@", 1)[0] == sym]
    return max(hits)[1] if hits else None


def phi_accuracy(chain, grav: float) -> dict:
    """phi of the last step against the exact discrete solution and the continuum FFT."""
    from galaxy_collision_amd.accuracy import phi_accuracy as acc

    return acc(chain.grid.rho, chain.grid.phi, chain.n, chain.dx, grav)


def accuracy_series(icr, w: dict, dx: float, grav: float, make, steps: int,
                    cadence: int) -> dict:
    """A run of `steps` steps from the initial state with the benchmark's solver, checked
    every `cadence` steps (and at step 0) against the exact discrete solution and the
    zero-padded float64 FFT: phi and particle-force relative L2 errors."""
    import torch

    from galaxy_collision_amd.accuracy import Checkpoint
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState

    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    chain = PMChain(parts, w["grid"], dx, make(), reorder_every=w.get("reorder_every", 0),
                    deposit_mode=w.get("deposit", "sorted"))
    check = Checkpoint(w["grid"], dx, grav)
    chain.prime()
    points = [{"step": 0, **check.measure(parts, chain.grid.rho, chain.grid.phi)}]
    done = 0
    while done < steps:
        k = min(cadence, steps - done)
        chain.run(k, w["dt"])
        done += k
        points.append({"step": done, **check.measure(parts, chain.grid.rho, chain.grid.phi)})
        print(f"[bench] accuracy step {done}: {points[-1]}", file=sys.stderr, flush=True)
    keys = [k for k in points[0] if k != "step"]
    out = {"steps": steps, "cadence": cadence,
           "max": {k: max(p[k] for p in points) for k in keys},
           "last": points[-1], "points": points,
           "vcycles_per_step": float(sum(chain.last_cycles[1:]) / max(len(chain.last_cycles) - 1, 1))}
    del chain, check, parts
    torch.cuda.empty_cache()
    return out


def warm_start_ab(icr, w: dict, dx: float, grav: float, steps: int) -> dict:
    """The same collision from the same initial state, stepped with each multigrid warm start
    (previous phi as in the reference, time extrapolation, extrapolation plus the trained
    warm-start network) at the benchmark's tolerance and V-cycle shape, and with the float32
    reference multigrid at its default tolerance: V-cycles and wall time per step, phi error
    of the last step."""
    import numpy as np
    import torch

    from galaxy_collision_amd.accuracy import phi_rel_error_discrete_device
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.solver import make_solver
    from galaxy_collision_amd.warmstart import WEIGHTS_DIR, WarmStartNet

    pool = 4
    variants = [("multigrid_mp", "previous"), ("multigrid_mp", "extrapolate")]
    weights = WEIGHTS_DIR / f"wsnet_r{w['grid'] // pool}.npz"
    net = WarmStartNet.load(weights) if weights.exists() else None
    if net is not None:
        variants.append(("multigrid_mp", "learned"))
    variants.append(("multigrid", "previous"))
    out = {}
    for name, warm in variants:
        parts = ParticleState(icr.n)
        parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
        kw = {"dx": dx, "grav_constant": grav}
        if name == "multigrid_mp":
            # the benchmark's own solver settings (tolerance, V-cycle shape)
            kw.update(warm=warm, pool=pool, net=net if warm == "learned" else None)
            if "tol" in w:
                kw["tol"] = w["tol"]
            if "sweeps" in w:
                kw["pre_sweeps"], kw["post_sweeps"] = w["sweeps"]
        solver = make_solver(name, w["grid"], **kw)
        chain = PMChain(parts, w["grid"], dx, solver, deposit_mode=w.get("deposit", "sorted"))
        chain.prime()
        for _ in range(3):
            chain.step(w["dt"])
        torch.cuda.synchronize()
        chain.last_cycles.clear()
        kept = []
        t0 = time.perf_counter()
        for _ in range(steps):
            chain.step(w["dt"])
            lw = getattr(solver, "last_warm", None)
            if warm == "learned" and lw and lw["extrapolated"]:
                kept.append(lw["used_net"])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        key = f"{name}:{warm}" if name == "multigrid_mp" else f"{name}:tol{solver.tol:g}"
        out[key] = {
            "vcycles_per_step": float(np.mean(chain.last_cycles)),
            "ms_per_step": wall / steps * 1e3,
            "phi_rel_l2_vs_exact_discrete": phi_rel_error_discrete_device(
                chain.grid.rho, chain.grid.phi, dx, grav),
        }
        if kept:
            out[key]["net_kept_fraction"] = float(np.mean(kept))
        print(f"[bench] A/B {key}: {out[key]}", file=sys.stderr, flush=True)
        del chain, solver, parts
        torch.cuda.empty_cache()
    return out


def cpu_baseline(icr, w: dict, dx: float, budget_s: float, threads: int | None) -> dict:
    """The reference algorithm on host cores with the reference's protocol: prime, one
    untimed warmup step, a throughput pass, then a per-stage pass (each stage timed)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pm_oracle

    chain = pm_oracle.CpuChain(icr.pos, icr.vel, icr.mass, w["grid"], dx, 1.0, threads=threads)
    t0 = time.perf_counter()
    chain.accel()
    prime_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    chain.step(w["dt"])  # warmup
    step_s = time.perf_counter() - t0
    # the two passes share the budget; at most 10 steps each
    k = max(1, min(10, int(0.5 * budget_s / max(step_s, 1e-9))))
    t0 = time.perf_counter()
    for _ in range(k):
        chain.step(w["dt"])
    elapsed = time.perf_counter() - t0
    stages: dict = {}
    for _ in range(k):
        for name, sec in chain.stage_times(w["dt"]).items():
            stages[name] = stages.get(name, 0.0) + sec
    chain.close()
    return {"value": k / elapsed, "unit": "steps/s", "cores": chain.threads, "kind": "port",
            "per_stage_ms": {s: stages[s] / k * 1e3 for s in ("integrate", "deposit", "solve",
                                                              "grad", "gather")},
            "sample": (f"{icr.n:,} particles on {w['grid']}^3, reference algorithm (OpenMP C "
                       f"port, atomic CIC, plain float32 multigrid tol 1e-3): prime "
                       f"{prime_s:.1f} s, 1 warmup step, {k} steps timed back to back, then "
                       f"{k} steps with every stage timed")}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="collision10m", choices=sorted(WORKLOADS))
    ap.add_argument("--solver", default=None, help="override the workload's solver")
    ap.add_argument("--tol", type=float, default=None, help="solver tolerance override")
    ap.add_argument("--warm", default=None, choices=["previous", "extrapolate"],
                    help="multigrid warm start across steps (default: the workload's)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reorder-every", type=int, default=None,
                    help="steps between moves of the particles into cell order (0: never; "
                         "default: the workload's)")
    ap.add_argument("--deposit-mode", default=None, choices=["sorted", "fixed", "bricked"],
                    help="the chain's deposit mode (default: the workload's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of timed CPU steps")
    ap.add_argument("--cpu-threads", type=int, default=None)
    ap.add_argument("--no-accuracy", action="store_true", help="skip the phi accuracy check")
    ap.add_argument("--no-kernel-pass", action="store_true", help="skip the per-kernel pass")
    ap.add_argument("--no-ab", action="store_true",
                    help="skip the warm-start A/B pass (learned vs plain multigrid)")
    ap.add_argument("--ab-steps", type=int, default=20)
    ap.add_argument("--no-equilibrate", action="store_true",
                    help="keep the analytic disk velocities (no grid rotation curve)")
    ap.add_argument("--no-stage-pass", action="store_true", help="skip the per-stage pass")
    ap.add_argument("--steady-after", type=int, default=300,
                    help="time --steps more steps after the run reaches this step (0: skip)")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the accuracy series from the initial state (0: skip)")
    ap.add_argument("--accuracy-cadence", type=int, default=50)
    ap.add_argument("--copy-probe", action="store_true",
                    help="first copy a 256 MiB buffer 3 times (a known-size kernel for scaling "
                         "profiler byte counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on exactly one GPU; --gpus {args.gpus} is not supported in this "
              "build (the multi-GPU slab decomposition is a design note, not code).",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState, estimate_memory
    from galaxy_collision_amd.ic import box_fit_fraction
    from galaxy_collision_amd.solver import make_solver

    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        for _ in range(3):
            nat.call("pm_copy_f32", nat.ptr(dst), nat.ptr(src), src.numel(), nat.stream())
        torch.cuda.synchronize()
        del src, dst
    w = dict(WORKLOADS[args.workload])
    if args.deposit_mode:
        w["deposit"] = args.deposit_mode
    solver_name = args.solver or w["solver"]
    grav = 1.0
    t_ic = time.perf_counter()
    icr, dx = build_ic(w, args.seed)
    equilibrated = w["kind"] == "two_galaxy" and not args.no_equilibrate
    if equilibrated:
        equilibrate(icr, w, dx, grav, args.seed)
    ic_s = time.perf_counter() - t_ic
    outside = box_fit_fraction(icr, w["grid"], dx)
    print(f"[bench] workload={args.workload} N={icr.n:,} grid={w['grid']}^3 dx={dx:.5f} "
          f"solver={solver_name} ic={ic_s:.1f}s outside={outside:.2%}", file=sys.stderr, flush=True)

    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    kw = {"dx": dx, "grav_constant": grav}
    if args.tol is not None:
        kw["tol"] = args.tol
    elif "tol" in w and solver_name == w["solver"]:
        kw["tol"] = w["tol"]
    if "sweeps" in w and solver_name == w["solver"]:
        kw["pre_sweeps"], kw["post_sweeps"] = w["sweeps"]
    warm = args.warm or (w.get("warm", "previous") if solver_name == w["solver"] else "previous")
    if solver_name in ("multigrid", "multigrid_mp") and warm != "previous":
        kw["warm"] = warm
    make = lambda: make_solver(solver_name, w["grid"], **kw)  # noqa: E731
    solver = make()
    reorder = args.reorder_every if args.reorder_every is not None else w.get("reorder_every", 0)
    chain = PMChain(parts, w["grid"], dx, solver, reorder_every=reorder,
                    deposit_mode=w.get("deposit", "sorted"))
    chain.prime()
    chain.run(args.warmup, w["dt"])
    torch.cuda.synchronize()

    # pass 1: throughput. The dominant kernel's own dispatches record HIP event pairs on its
    # stream (timing mode 2 adds no packets, so the steps run as untimed ones)
    chain.last_cycles.clear()
    nat.timing_reset()
    nat.timing_enable(2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    chain.run(args.steps, w["dt"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nat.timing_enable(False)
    live = nat.timing_report()
    nat.timing_reset()
    cycles = list(chain.last_cycles)
    ms_per_step = wall / args.steps * 1e3
    value = args.steps / wall
    print(f"[bench] pass 1: {ms_per_step:.3f} ms/step, V-cycles/step {np.mean(cycles):.2f}",
          file=sys.stderr, flush=True)

    # pass 2: per-kernel HIP events (same steps, instrumented)
    kernels, roofline, pass2_ms, name = {}, None, None, None
    if not args.no_kernel_pass:
        nat.timing_reset()
        nat.timing_enable(True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        chain.run(args.steps, w["dt"])
        torch.cuda.synchronize()
        pass2_ms = (time.perf_counter() - t1) / args.steps * 1e3
        nat.timing_enable(False)
        rep = nat.timing_report()
        for k, r in rep.items():
            avg_ms = r["total_ms"] / r["launches"]
            kernels[k] = {
                "ms_per_step": r["total_ms"] / args.steps,
                "launches_per_step": r["launches"] / args.steps,
                "avg_launch_us": avg_ms * 1e3,
                "GBps": (r["bytes"] / r["launches"]) / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0,
            }
        name = max(rep.items(), key=lambda kv: kv[1]["total_ms"])[0]
        print(f"[bench] pass 2: {pass2_ms:.3f} ms/step; dominant {name}", file=sys.stderr,
              flush=True)
    elif live:
        name = max(live.items(), key=lambda kv: kv[1]["total_ms"])[0]
    if name is not None:
        # the dominant kernel's average launch: from the timed steps of pass 1 when its
        # dispatches recorded their own events there, else from pass 2's event pairs
        if name in live:
            r = live[name]
            timing = "events recorded by the kernel's own dispatches over the timed steps"
        else:
            r = rep[name]
            timing = "event pairs around each launch (instrumented pass)"
        avg_s = r["total_ms"] / r["launches"] * 1e-3
        per_launch = r["bytes"] / r["launches"]
        achieved = per_launch / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                    "traffic": measured_traffic(args.workload, w["grid"], name),
                    "algorithmic_bytes_per_launch": per_launch, "avg_launch_us": avg_s * 1e6,
                    "timing": timing}
        print(f"[bench] roofline: {name} {avg_s * 1e6:.2f} us/launch, {achieved:.0f} GB/s",
              file=sys.stderr, flush=True)

    mg_fractions = {}
    for k, r in kernels.items():
        finest = k.endswith(f"[{w['grid']}]") or "[" not in k
        if k.startswith("mg.") and finest and r["GBps"] > 0:
            mg_fractions[k] = {"GBps": r["GBps"], "frac": r["GBps"] / HBM_PEAK_GBS,
                               "avg_launch_us": r["avg_launch_us"],
                               "launches_per_step": r["launches_per_step"]}

    # pass 3: the reference's per-stage protocol (unfused stages, each synchronized)
    per_stage = None
    if not args.no_stage_pass:
        from galaxy_collision_amd.benchmark import UnfusedStep

        per_stage = UnfusedStep(chain, w["dt"]).stage_pass(args.steps)
        print(f"[bench] pass 3 per stage (ms): {per_stage}", file=sys.stderr, flush=True)

    # pass 4: the steady state after the particles have mixed
    steady = None
    if args.steady_after > 0:
        ahead = args.steady_after - chain.steps_done
        if ahead > 0:
            chain.run(ahead, w["dt"])
        torch.cuda.synchronize()
        chain.last_cycles.clear()
        t2 = time.perf_counter()
        chain.run(args.steps, w["dt"])
        torch.cuda.synchronize()
        dt_s = time.perf_counter() - t2
        steady = {"first_step": chain.steps_done - args.steps + 1, "steps": args.steps,
                  "ms_per_step": dt_s / args.steps * 1e3, "steps_per_s": args.steps / dt_s,
                  "vcycles_per_step": float(np.mean(chain.last_cycles))}
        print(f"[bench] pass 4 steady state: {steady}", file=sys.stderr, flush=True)

    acc = {}
    if not args.no_accuracy:
        acc = phi_accuracy(chain, grav)
        print(f"[bench] accuracy {acc}", file=sys.stderr, flush=True)
    stats = getattr(solver, "last_stats", None) or {}
    series = None
    if args.accuracy_steps > 0 and not args.no_accuracy:
        del chain
        torch.cuda.empty_cache()
        series = accuracy_series(icr, w, dx, grav, make, args.accuracy_steps,
                                 args.accuracy_cadence)

    ab, learned_vs_plain = None, None
    if not args.no_ab and w["kind"] == "two_galaxy":
        ab = warm_start_ab(icr, w, dx, grav, args.ab_steps)
        plain, learned = ab.get("multigrid_mp:previous"), ab.get("multigrid_mp:learned")
        if plain and learned:
            learned_vs_plain = {
                "plain": "mixed-precision multigrid warm-started from the previous phi (the "
                         "reference's warm start), same tolerance",
                "learned": "the same solver warm-started by extrapolation plus the trained "
                           "warm-start network",
                "vcycles_per_step": [learned["vcycles_per_step"], plain["vcycles_per_step"]],
                "ms_per_step": [learned["ms_per_step"], plain["ms_per_step"]],
                "vcycles_ratio": learned["vcycles_per_step"] / plain["vcycles_per_step"],
                "wall_ratio": learned["ms_per_step"] / plain["ms_per_step"],
            }

    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(icr, w, dx, args.cpu_budget, args.cpu_threads)
        print(f"[bench] CPU baseline {cpu['value']:.4f} steps/s on {cpu['cores']} threads",
              file=sys.stderr, flush=True)

    out = {
        "metric": "PM steps/s @256^3,10M particles" if args.workload == "collision10m"
        else f"PM steps/s ({args.workload})",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REFERENCE_STEPS_PER_S if args.workload == "collision10m" else None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "particles": icr.n,
            "grid": w["grid"],
            "box": w["box"],
            "dt": w["dt"],
            "solver": solver_name,
            "solver_tol": getattr(solver, "tol", None),
            "vcycle": (f"V({solver.nu1},{solver.nu2})" if hasattr(solver, "nu1") else None),
            "warm_start": getattr(solver, "warm", None),
            "precision": "f32 particles/V-cycles; f64 residual, coarsest solve, reductions"
            if solver_name == "multigrid_mp" else "f32",
            "parallelism": "single GPU",
            "particle_order": "sorted by initial cell" + (
                f"; moved into cell order every {reorder} steps" if reorder else ""),
            "disk_velocities": ("rotation curve measured on the grid + 10% dispersion"
                                if equilibrated else "analytic enclosed-mass curve + 10% "
                                "dispersion" if w["kind"] == "two_galaxy" else None),
        },
        "particle_steps_per_s": value * icr.n,
        "mean_vcycles_per_step": float(np.mean(cycles)) if cycles else None,
        "solver_last_stats": stats,
        "ms_per_step_instrumented": pass2_ms,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": dict(sorted(kernels.items(), key=lambda kv: -kv[1]["ms_per_step"])),
        "mg_finest_roofline": mg_fractions or None,
        "per_stage_ms": per_stage,
        "per_stage_protocol": "unfused stages, each bracketed by device synchronizations "
                              "(reference bench.py:66-72, 176-186); serializes the device",
        "per_stage_reference_ms_rtx3070": REFERENCE_STAGE_MS,
        "steady_state": steady,
        "mem_estimate_gb": estimate_memory(icr.n, w["grid"]).total_gb,
        **acc,
        "accuracy_series": ({k: v for k, v in series.items() if k != "points"}
                            if series else None),
        "learned_vs_plain_mg": learned_vs_plain,
        "warm_start_ab": ab,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
