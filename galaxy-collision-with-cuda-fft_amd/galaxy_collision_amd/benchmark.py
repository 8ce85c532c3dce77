"""Benchmark harness of the force chain (reference bench.py:66-252, ``galaxy-bench``).

``run_benchmark(config, warmup=3, measure=10, deposit_fn=None)`` follows the reference's
protocol: build the IC, launch the disks on the grid's rotation curve (two-galaxy presets,
reference bench.py:119-123), prime, ``warmup`` untimed steps, a throughput pass of
``measure`` steps with one synchronization at the end (steps/s and the solver's V-cycles per
step), then a per-stage pass of ``measure`` steps in which every stage (integrate = both
half kicks and the drift, deposit, solve, grad, gather) is bracketed by device
synchronizations (reference bench.py:66-72, 176-186). It returns the reference's report keys.

The throughput pass runs the production path, ``PMChain.run`` (fused particle passes, bitwise
the unfused step). With ``deposit_fn`` (called as ``deposit_fn(parts, rho, dx)`` like
``deposit_density``, the reference's deposit plug) the throughput pass runs the unfused step
with that deposit instead, so deposit variants can be compared. The per-stage pass always
runs the unfused stages (the fused step has no stage boundaries); it serializes the device,
so its sum exceeds the throughput step time.
"""

from __future__ import annotations

import sys
import time
from typing import Any, Callable

import numpy as np
import torch

STAGES = ("integrate", "deposit", "solve", "grad", "gather")


def _timed(fn: Callable[[], None]) -> float:
    """Wall time of fn() bracketed by device synchronizations (reference bench.py:66-72)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


class UnfusedStep:
    """The reference step kick -> drift -> deposit -> solve -> grad -> gather -> kick as
    separate device calls on a chain's particles, grid and acceleration buffers."""

    def __init__(self, chain, dt: float, deposit_fn: Callable | None = None):
        self.chain = chain
        self.dt = float(dt)
        self.deposit_fn = deposit_fn

    def kick(self) -> None:
        from galaxy_collision_amd.integrator import kick

        p = self.chain.parts
        kick(p.vel_x, p.vel_y, p.vel_z, *self.chain.acc, p.n, 0.5 * self.dt)

    def drift(self) -> None:
        from galaxy_collision_amd.integrator import drift

        p = self.chain.parts
        drift(p.pos_x, p.pos_y, p.pos_z, p.vel_x, p.vel_y, p.vel_z, p.n, self.dt)

    def deposit(self) -> None:
        if self.deposit_fn is None:
            self.chain.deposit()  # the chain's own plan (sorted deposit)
        else:
            self.deposit_fn(self.chain.parts, self.chain.grid.rho, self.chain.dx)

    def solve(self) -> None:
        self.chain.solve()

    def grad(self) -> None:
        from galaxy_collision_amd.deposit import potential_to_accel

        g = self.chain.grid
        potential_to_accel(g.phi, g.ax, g.ay, g.az, self.chain.dx)

    def gather(self) -> None:
        from galaxy_collision_amd.deposit import gather_acceleration

        g = self.chain.grid
        gather_acceleration(self.chain.parts, g.ax, g.ay, g.az, *self.chain.acc, self.chain.dx)

    def __call__(self) -> None:
        self.kick()
        self.drift()
        self.deposit()
        self.solve()
        self.grad()
        self.gather()
        self.kick()

    def stage_pass(self, steps: int) -> dict[str, float]:
        """ms per step of each stage, every stage synchronized on both sides."""
        total = {s: 0.0 for s in STAGES}
        for _ in range(steps):
            total["integrate"] += _timed(self.kick)
            total["integrate"] += _timed(self.drift)
            total["deposit"] += _timed(self.deposit)
            total["solve"] += _timed(self.solve)
            total["grad"] += _timed(self.grad)
            total["gather"] += _timed(self.gather)
            total["integrate"] += _timed(self.kick)
        return {s: total[s] / steps * 1e3 for s in STAGES}


def run_benchmark(config, *, warmup: int = 3, measure: int = 10,
                  deposit_fn: Callable[..., None] | None = None, icr=None,
                  equilibrate: bool | None = None, solver_kwargs: dict | None = None
                  ) -> dict[str, Any]:
    """Throughput and per-stage breakdown of the force chain for ``config`` (a SimConfig);
    see the module docstring. ``icr`` overrides the configured initial conditions."""
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import GridState, ParticleState, estimate_memory
    from galaxy_collision_amd.sim import build_ic, check_box_fit
    from galaxy_collision_amd.solver import make_solver

    nat.require_device()
    dx, gs = float(config.dx), int(config.grid_size)
    if icr is None:
        icr = build_ic(config)
    check_box_fit(icr, gs, dx)
    n = icr.n
    parts = ParticleState(n)
    parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    grid = GridState(gs)
    solver = make_solver(config.solver, gs, dx=dx, grav_constant=config.grav,
                         **(solver_kwargs or {}))
    if equilibrate is None:
        equilibrate = config.ic_preset == "two_galaxy"
    if equilibrate:
        from galaxy_collision_amd.ic import equilibrate_disk_velocities

        equilibrate_disk_velocities(icr, dx=dx, solver=solver, rho=grid.rho, phi=grid.phi,
                                    ax=grid.ax, ay=grid.ay, az=grid.az, parts=parts,
                                    dispersion=0.1, seed=config.seed)
    chain = PMChain(parts, gs, dx, solver, grid=grid,
                    reorder_every=config.reorder_every if deposit_fn is None else 0)
    unfused = UnfusedStep(chain, config.dt, deposit_fn)

    chain.prime()
    if deposit_fn is None:
        chain.run(warmup, config.dt)
    else:
        for _ in range(warmup):
            unfused()
    torch.cuda.synchronize()
    free, total_mem = torch.cuda.mem_get_info()

    chain.last_cycles.clear()
    t0 = time.perf_counter()
    if deposit_fn is None:
        chain.run(measure, config.dt)
    else:
        for _ in range(measure):
            unfused()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cycles = list(chain.last_cycles)

    per_stage_ms = unfused.stage_pass(measure)
    return {
        "device": torch.cuda.get_device_name(0),
        "backend_resolved": "hip",
        "n_particles": n,
        "grid_size": gs,
        "solver": config.solver,
        "path": "fused PMChain.run" if deposit_fn is None else "unfused step with deposit_fn",
        "steps_per_sec": measure / wall,
        "ms_per_step": wall / measure * 1e3,
        "per_stage_ms": per_stage_ms,
        "mem_used_mib": (total_mem - free) / 2 ** 20,
        "mem_label": "device memory used",
        "est_total_gb": estimate_memory(n, gs).total_gb,
        "mean_solver_cycles": float(np.mean(cycles)) if cycles else None,
    }


def format_report(r: dict[str, Any]) -> str:
    lines = [
        f"device      : {r['device']}  (arch={r['backend_resolved']})",
        f"N particles : {r['n_particles']:,}",
        f"grid        : {r['grid_size']}^3   solver={r['solver']}   path={r['path']}",
        f"throughput  : {r['steps_per_sec']:.2f} steps/s   ({r['ms_per_step']:.3f} ms/step)",
        f"est. fields : {r['est_total_gb']:.2f} GB (analytical: particles + grid)",
        f"{r['mem_label']:<12s}: {r['mem_used_mib']:.0f} MiB",
    ]
    if r["mean_solver_cycles"] is not None:
        lines.append(f"solver      : {r['mean_solver_cycles']:.2f} V-cycles/step (mean, adaptive)")
    lines.append("per-stage (ms/step, individually synced):")
    total = sum(r["per_stage_ms"].values())
    for s in STAGES:
        ms = r["per_stage_ms"][s]
        lines.append(f"    {s:<10s} {ms:8.3f} ms   {100.0 * ms / total if total else 0.0:5.1f}%")
    lines.append(f"    {'(sum)':<10s} {total:8.3f} ms")
    return "\n".join(lines)


def main(argv=None) -> int:
    """``python -m galaxy_collision_amd.benchmark --n 10000000 --grid 256``."""
    import argparse

    from galaxy_collision_amd.sim import IC_PRESETS, SimConfig

    ap = argparse.ArgumentParser(prog="galaxy-bench", description="Benchmark the PM force chain.")
    ap.add_argument("--n", type=int, default=1_000_000, help="particle count")
    ap.add_argument("--grid", type=int, default=256, help="grid size per side")
    ap.add_argument("--box", type=float, default=40.0, help="box side (dx = box / grid)")
    ap.add_argument("--solver", default="multigrid_mp")
    ap.add_argument("--preset", default="two_galaxy", choices=IC_PRESETS)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--measure", type=int, default=10)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    cfg = SimConfig(name="bench", n_particles=a.n, dt=a.dt, steps=0, ic_preset=a.preset,
                    solver=a.solver, grid_size=a.grid, dx=a.box / a.grid, seed=a.seed)
    print(format_report(run_benchmark(cfg, warmup=a.warmup, measure=a.measure)),
          file=sys.stdout)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())


__all__ = ["STAGES", "UnfusedStep", "run_benchmark", "format_report", "main"]
