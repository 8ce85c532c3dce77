"""The simulation driver on the device force chain (reference sim.py:173-450).

``run_simulation(config, ...)`` builds (or takes) the initial conditions, refuses ones that
do not fit the box, primes the force chain, runs ``config.steps`` fused KDK steps, samples
the conservation diagnostics at a cadence (the reference's ``measure`` contract, sim.py:
341-371: energy, kinetic, potential, momentum, angular momentum, half-mass radius and
virial ratio, reduced in float64 on the device) and optionally writes snapshots and
device-projected frames. It returns the reference's summary keys (energy drift at the
endpoint and over the run, final virial ratio, history, snapshot paths, tracer path,
frames).
"""

from __future__ import annotations

import sys
from dataclasses import asdict, dataclass
from pathlib import Path
from typing import Any

import numpy as np
import torch

IC_PRESETS = ("plummer", "two_galaxy")


@dataclass
class SimConfig:
    """One run (the reference's SimConfig fields that drive the hot path).

    Units follow ``grav``: G = 1 with box units for the synthetic benchmarks, or
    ``units.G`` with kpc / Myr / M_sun as in the reference.
    """

    name: str = "run"
    n_particles: int = 10_000
    dt: float = 0.01
    steps: int = 1
    ic_preset: str = "plummer"
    solver: str = "multigrid_mp"
    grid_size: int = 64
    dx: float = 1.0
    grav: float = 1.0
    total_mass: float = 1.0     # plummer: sphere mass; two_galaxy: stellar mass per galaxy
    scale_a: float = 1.0        # plummer scale radius
    separation: float = 20.0    # two_galaxy centre distance
    output_cadence: int = 0     # steps between snapshots (0: none)
    output_dir: str = "outputs"
    seed: int = 0
    reorder_every: int = 4      # steps between moves of the particles into cell order (0: never)

    def __post_init__(self) -> None:
        self.validate()

    def validate(self) -> None:
        from galaxy_collision_amd.solver import SOLVER_NAMES

        if self.solver not in SOLVER_NAMES:
            raise ValueError(f"solver must be one of {SOLVER_NAMES}, got {self.solver!r}")
        if self.ic_preset not in IC_PRESETS:
            raise ValueError(f"ic_preset must be one of {IC_PRESETS}, got {self.ic_preset!r}")
        if self.n_particles <= 0:
            raise ValueError(f"n_particles must be positive, got {self.n_particles}")
        if self.dt <= 0:
            raise ValueError(f"dt must be positive, got {self.dt}")
        if self.steps < 0:
            raise ValueError(f"steps must be non-negative, got {self.steps}")
        if self.grid_size < 3 or self.dx <= 0:
            raise ValueError("grid_size must be >= 3 and dx positive")
        if self.reorder_every < 0:
            raise ValueError(f"reorder_every must be non-negative, got {self.reorder_every}")
        if self.output_cadence < 0:
            raise ValueError(f"output_cadence must be non-negative, got {self.output_cadence}")

    def to_dict(self) -> dict:
        return asdict(self)


def build_ic(config: SimConfig):
    """The configured preset, centred in the box."""
    from galaxy_collision_amd import ic

    c = (config.grid_size - 1) * config.dx / 2.0
    if config.ic_preset == "plummer":
        return ic.plummer_ic(config.n_particles, total_mass=config.total_mass,
                             scale_a=config.scale_a, grav=config.grav, centre=(c, c, c),
                             seed=config.seed)
    n_per = max(1, (config.n_particles - 2) // 2)
    return ic.two_galaxy_ic(n_per, box_centre=(c, c, c), separation=config.separation,
                            masses=(config.total_mass, config.total_mass), grav=config.grav,
                            seed=config.seed)


def check_box_fit(icr, grid_size: int, dx: float, tol: float = 0.10) -> float:
    """Mass fraction outside the node span; raises ValueError above ``tol`` (the CIC
    deposit drops out-of-box weight, so a misplaced IC would silently lose mass;
    reference sim.py:173-191)."""
    from galaxy_collision_amd.ic import box_fit_fraction

    frac = box_fit_fraction(icr, grid_size, dx)
    if frac > tol:
        raise ValueError(f"{frac:.1%} of the IC mass lies outside the {grid_size}^3 box: the "
                         "initial conditions do not fit (check grid_size vs placement)")
    return frac


def run_simulation(config: SimConfig, *, history_cadence: int | None = None,
                   write_snapshots: bool = True, solver_kwargs: dict | None = None,
                   direct_pe_softening: float | None = None, icr=None, tracer_indices=None,
                   frame_cadence: int | None = None, frame_bins: int = 512,
                   frame_axes: tuple[int, int] = (0, 1), accuracy_cadence: int | None = None,
                   equilibrate: bool | None = None,
                   equilibrate_kwargs: dict | None = None) -> dict[str, Any]:
    """Run the PM pipeline and return its summary and diagnostics history.

    history_cadence: steps between diagnostics samples (default ~20 over the run, plus the
    first and last); direct_pe_softening: also record the O(N^2) softened pair-sum energy
    (small N only); tracer_indices: particle indices whose positions are recorded at each
    sample; frame_cadence: capture a device-projected surface-density frame every that many
    steps (plus first and last); accuracy_cadence: every that many steps (plus the first and
    last) compare phi and the particle forces with the exact float64 solution of the same
    discrete problem and with the zero-padded float64 continuum FFT (``accuracy.Checkpoint``;
    results under "accuracy"); equilibrate: launch the disks of a two-galaxy IC on the
    rotation curve the grid produces plus the recipe's 10% dispersion
    (``ic.equilibrate_disk_velocities``, options in ``equilibrate_kwargs``; default: on for
    two-galaxy presets as in reference sim.py:310-314).
    """
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import GridState, ParticleState
    from galaxy_collision_amd.diagnostics import potential_energy_direct, virial_ratio
    from galaxy_collision_amd.diagnostics_device import DeviceDiagnostics
    from galaxy_collision_amd.io import snapshot_from_states, write_snapshot
    from galaxy_collision_amd.project import project_density
    from galaxy_collision_amd.solver import make_solver

    nat.require_device()
    dx, gsize = float(config.dx), int(config.grid_size)
    if icr is None:
        icr = build_ic(config)
    outside = check_box_fit(icr, gsize, dx)
    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
    grid = GridState(gsize)
    solver = make_solver(config.solver, gsize, dx=dx, grav_constant=config.grav,
                         **(solver_kwargs or {}))
    if equilibrate is None:
        equilibrate = config.ic_preset == "two_galaxy"
    curves = None
    if equilibrate:
        from galaxy_collision_amd.ic import equilibrate_disk_velocities

        # the recipe's 10% dispersion is kept on top of the measured curve
        ekw = {"dispersion": 0.1, "seed": config.seed, **(equilibrate_kwargs or {})}
        curves = equilibrate_disk_velocities(icr, dx=dx, solver=solver, rho=grid.rho,
                                             phi=grid.phi, ax=grid.ax, ay=grid.ay, az=grid.az,
                                             parts=parts, **ekw)
    if accuracy_cadence is not None and int(accuracy_cadence) <= 0:
        raise ValueError(f"accuracy_cadence must be positive, got {accuracy_cadence}")
    checker = None
    accuracy: list[dict] = []
    if accuracy_cadence:
        from galaxy_collision_amd.accuracy import Checkpoint

        checker = Checkpoint(gsize, dx, config.grav)
    chain = PMChain(parts, gsize, dx, solver, grid=grid, reorder_every=config.reorder_every)
    diag = DeviceDiagnostics(gsize, dx=dx)
    tracer_idx = None if tracer_indices is None else [int(i) for i in np.atleast_1d(tracer_indices)]
    frames, frame_times = [], []
    extent = (0.0, gsize * dx, 0.0, gsize * dx)
    frame_img = (torch.zeros((frame_bins, frame_bins), device=torch.device("cuda"))
                 if frame_cadence else None)

    def measure(step: int) -> dict[str, Any]:
        d = diag.sample(parts, grid, dx)
        ke, pe = d["kinetic"], d["potential"]
        rec = {"step": step, "time": step * config.dt, "energy": d["energy"], "kinetic": ke,
               "potential": pe, "momentum": d["momentum"], "ang_momentum": d["ang_momentum"],
               "half_mass_radius": d["half_mass_radius"], "virial": virial_ratio(ke, pe)}
        if direct_pe_softening is not None:
            pos = parts.positions().astype(np.float64)
            mass_np = parts.mass.to_numpy().astype(np.float64)  # same (current) slot order
            pe_d = potential_energy_direct(mass_np, pos, direct_pe_softening, config.grav)
            rec["potential_direct"] = pe_d
            rec["energy_direct"] = ke + pe_d
        if tracer_idx is not None:
            # tracers are load indices; the chain may have moved the particles
            slots = parts.slots_of(tracer_idx)
            rec["tracer_pos"] = torch.stack(
                [f.tensor[slots] for f in (parts.pos_x, parts.pos_y, parts.pos_z)],
                1).cpu().numpy().astype(np.float64)
        return rec

    diag_keys = ("energy", "kinetic", "potential", "momentum", "ang_momentum",
                 "half_mass_radius", "virial")
    snapshots: list[str] = []
    out_dir = Path(config.output_dir)

    def snapshot_due(step: int) -> bool:
        if not (write_snapshots and config.output_cadence > 0):
            return False
        return step % config.output_cadence == 0 or step == config.steps

    def write_snap(step: int, m: dict) -> None:
        snap = snapshot_from_states(step, step * config.dt, parts, grid=grid,
                                    diagnostics={k: np.asarray(m[k]) for k in diag_keys}, dx=dx)
        snapshots.append(str(write_snapshot(snap, out_dir / f"snapshot_{step:05d}.npz")))

    def capture_frame(step: int) -> None:
        frames.append(project_density(parts, frame_img, extent, frame_axes))
        frame_times.append(step * config.dt)

    chain.prime()
    cadence = history_cadence or max(1, config.steps // 20)
    m0 = measure(0)
    history = [m0]
    if snapshot_due(0):
        write_snap(0, m0)
    if frame_cadence:
        capture_frame(0)
    def frame_due(step: int) -> bool:
        return bool(frame_cadence) and (step % frame_cadence == 0 or step == config.steps)

    def accuracy_due(step: int) -> bool:
        return bool(accuracy_cadence) and (step % accuracy_cadence == 0 or step == config.steps)

    def check_accuracy(step: int) -> None:
        rec = {"step": step, "time": step * config.dt}
        rec.update(checker.measure(parts, grid.rho, grid.phi))
        accuracy.append(rec)

    if checker is not None:
        check_accuracy(0)
    step = 0
    while step < config.steps:
        # run up to the next step that is sampled, written, captured or checked, in one
        # fused run
        nxt = step + 1
        while not (nxt % cadence == 0 or nxt == config.steps or snapshot_due(nxt)
                   or frame_due(nxt) or accuracy_due(nxt)):
            nxt += 1
        chain.run(nxt - step, config.dt)
        step = nxt
        if accuracy_due(step):
            check_accuracy(step)
        want_hist = step % cadence == 0 or step == config.steps
        want_snap = snapshot_due(step)
        if want_hist or want_snap:
            m = measure(step)
            if want_hist:
                history.append(m)
            if want_snap:
                write_snap(step, m)
        if frame_due(step):
            capture_frame(step)
    torch.cuda.synchronize()

    e0, ef = history[0]["energy"], history[-1]["energy"]
    drift = abs((ef - e0) / e0) if e0 != 0.0 else float("nan")
    drift_max = (max(abs((h["energy"] - e0) / e0) for h in history) if e0 != 0.0
                 else float("nan"))
    return {
        "name": config.name, "device": torch.cuda.get_device_name(0), "preset": config.ic_preset,
        "solver": config.solver, "n_particles": icr.n, "grid_size": gsize, "steps": config.steps,
        "dt": config.dt, "mass_outside_box": outside,
        "energy_initial": e0, "energy_final": ef, "energy_drift": drift,
        "energy_drift_max": drift_max, "virial_final": history[-1]["virial"],
        "vcycles_per_step": (float(np.mean(chain.last_cycles[1:])) if len(chain.last_cycles) > 1
                             else None),
        "tracer_indices": tracer_idx,
        "tracer_path": (np.stack([h["tracer_pos"] for h in history]) if tracer_idx is not None
                        else None),
        "tracer_times": (np.array([h["time"] for h in history]) if tracer_idx is not None
                         else None),
        "frames": np.stack(frames) if frames else None,
        "frame_times": np.array(frame_times) if frames else None,
        "frame_extent": extent if frame_cadence else None,
        "history": history, "snapshots": snapshots, "status": "ok",
        "accuracy": accuracy if checker is not None else None,
        "accuracy_max": ({k: max(a[k] for a in accuracy) for k in accuracy[0]
                          if k not in ("step", "time")} if accuracy else None),
        "rotation_curves": curves,
    }


def main(argv=None) -> int:
    """``python -m galaxy_collision_amd.sim --preset plummer --n 131072 --grid 64 ...``"""
    import argparse
    import json

    ap = argparse.ArgumentParser(description="run a PM simulation on the device")
    ap.add_argument("--preset", default="plummer", choices=IC_PRESETS)
    ap.add_argument("--n", type=int, default=131_072)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--dx", type=float, default=20.0 / 64)
    ap.add_argument("--dt", type=float, default=0.01)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--solver", default="multigrid_mp")
    ap.add_argument("--output-cadence", type=int, default=0)
    ap.add_argument("--output-dir", default="outputs")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--accuracy-cadence", type=int, default=None,
                    help="check phi and forces against the exact and FFT solutions every K steps")
    ap.add_argument("--no-equilibrate", action="store_true",
                    help="keep the analytic disk velocities of a two-galaxy IC")
    a = ap.parse_args(argv)
    cfg = SimConfig(name=a.preset, n_particles=a.n, dt=a.dt, steps=a.steps, ic_preset=a.preset,
                    solver=a.solver, grid_size=a.grid, dx=a.dx, output_cadence=a.output_cadence,
                    output_dir=a.output_dir, seed=a.seed)
    res = run_simulation(cfg, accuracy_cadence=a.accuracy_cadence,
                         equilibrate=False if a.no_equilibrate else None)
    keep = {k: v for k, v in res.items()
            if k not in ("history", "frames", "tracer_path", "rotation_curves")}
    print(json.dumps(keep, default=float), file=sys.stdout)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
