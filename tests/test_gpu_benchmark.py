"""``run_benchmark`` (reference bench.py:79-204): the reference's report keys, a per-stage
breakdown with every stage, the deposit plug, and the fused throughput path giving the same
state as the unfused reference step."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import announce

pytestmark = pytest.mark.gpu


def _config(**kw):
    from galaxy_collision_amd.sim import SimConfig

    n = 64
    args = dict(name="b", n_particles=40_002, dt=0.01, steps=0, ic_preset="two_galaxy",
                solver="multigrid_mp", grid_size=n, dx=40.0 / n, seed=2)
    args.update(kw)
    return SimConfig(**args)


def test_run_benchmark_report(device):
    announce("run_benchmark report, fused path")
    from galaxy_collision_amd.benchmark import STAGES, format_report, run_benchmark

    r = run_benchmark(_config(), warmup=2, measure=3)
    for key in ("device", "backend_resolved", "n_particles", "grid_size", "solver",
                "steps_per_sec", "ms_per_step", "per_stage_ms", "mem_used_mib", "mem_label",
                "est_total_gb", "mean_solver_cycles"):
        assert key in r, key
    assert r["n_particles"] == 40_002 and r["grid_size"] == 64
    assert set(r["per_stage_ms"]) == set(STAGES)
    assert all(v > 0 for v in r["per_stage_ms"].values())
    assert r["steps_per_sec"] > 0 and r["mean_solver_cycles"] >= 1
    assert "per-stage" in format_report(r)


def test_run_benchmark_deposit_plug(device):
    """deposit_fn replaces the deposit (reference bench.py:84, 102): it is called on every
    step of the unfused throughput and per-stage passes."""
    announce("run_benchmark with a deposit_fn")
    from galaxy_collision_amd.benchmark import run_benchmark
    from galaxy_collision_amd.deposit import deposit_density

    calls = []

    def fixed_point(parts, rho, dx):
        calls.append(1)
        deposit_density(parts, rho, dx, mode="fixed")

    r = run_benchmark(_config(), warmup=1, measure=2, deposit_fn=fixed_point)
    assert r["path"].startswith("unfused")
    assert len(calls) == 1 + 2 + 2  # warmup, throughput, per-stage (the prime uses the chain)
    assert np.isfinite(r["ms_per_step"])


def test_unfused_step_equals_fused_run(device):
    """The per-stage pass's unfused step is the reference step: after the same number of
    steps from the same state it leaves the particles bitwise where PMChain.run does."""
    announce("unfused step vs fused run")
    from galaxy_collision_amd.benchmark import UnfusedStep
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.solver import make_solver

    icr = plummer_ic(8000, centre=(8.0, 8.0, 8.0), seed=3)
    out = []
    for fused in (True, False):
        parts = ParticleState(icr.n)
        parts.load(icr.pos, icr.vel, icr.mass)
        chain = PMChain(parts, 32, 0.5, make_solver("multigrid", 32, dx=0.5, grav_constant=1.0))
        chain.prime()
        if fused:
            chain.run(4, 0.02)
        else:
            step = UnfusedStep(chain, 0.02)
            for _ in range(4):
                step()
        out.append((parts.positions(), parts.velocities()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
