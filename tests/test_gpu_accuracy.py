"""Accuracy checkpoints on the device: the device DST-I against the oracle, and
``run_simulation(accuracy_cadence=K)`` recording phi and force errors against the exact
discrete solution and the zero-padded float64 FFT at the cadence (the check the 10M / 256^3
configuration runs every 50 steps), each value equal to the oracle's recomputation.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import pm_oracle as O
from conftest import announce

pytestmark = pytest.mark.gpu
F32 = np.float32


def test_device_dst_matches_oracle(device):
    announce("device DST-I vs scipy")
    from galaxy_collision_amd.accuracy import dst_exact_device

    rng = np.random.default_rng(1)
    for n, dx in ((9, 0.5), (64, 0.3125)):
        f = rng.normal(size=(n, n, n)).astype(F32)
        b = rng.normal(size=(n, n, n)).astype(F32)
        ref = O.dirichlet_solve_f64(f, b, dx)
        got = dst_exact_device(torch.from_numpy(f).to(device), torch.from_numpy(b).to(device),
                               dx).cpu().numpy()
        assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_run_simulation_accuracy_cadence(device):
    announce("run_simulation accuracy cadence, 64^3 collision")
    from galaxy_collision_amd.sim import SimConfig, run_simulation

    n, dx = 64, 40.0 / 64
    cfg = SimConfig(name="acc", n_particles=60_002, dt=0.01, steps=10, ic_preset="two_galaxy",
                    solver="multigrid_mp", grid_size=n, dx=dx, output_cadence=0, seed=1)
    res = run_simulation(cfg, accuracy_cadence=4, write_snapshots=False, equilibrate=False)
    acc = res["accuracy"]
    assert [a["step"] for a in acc] == [0, 4, 8, 10]
    for a in acc:
        assert a["phi_rel_l2_vs_exact_discrete"] <= 1e-5, a
        assert a["force_rel_l2_vs_exact_discrete"] <= 1e-4, a
        # discretization: the 7-point Laplacian vs the continuum kernel. phi differs at the
        # percent level; the forces differ most at the cell scale, and this 64^3 grid puts a
        # galaxy's scale length on 1.6 cells around a point-mass black hole
        assert 1e-4 < a["phi_rel_l2_vs_fft_continuum"] < 5e-2, a
        assert 1e-4 < a["force_rel_l2_vs_fft_continuum"] < 1.0, a
    assert res["accuracy_max"]["phi_rel_l2_vs_exact_discrete"] == max(
        a["phi_rel_l2_vs_exact_discrete"] for a in acc)


def test_checkpoint_values_equal_oracle_recomputation(device):
    """The recorded phi and force errors of a chain state equal the oracle's float64 DST-I,
    gradient and gather applied to the same arrays."""
    announce("checkpoint vs oracle recomputation")
    from galaxy_collision_amd.accuracy import Checkpoint
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.solver import make_solver

    n, dx = 48, 0.5
    icr = plummer_ic(20_000, centre=(12.0, 11.5, 12.2), seed=4)
    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass)
    chain = PMChain(parts, n, dx, make_solver("multigrid", n, dx=dx, grav_constant=1.0))
    chain.prime()
    chain.run(3, 0.01)
    rec = Checkpoint(n, dx, 1.0).measure(parts, chain.grid.rho, chain.grid.phi)
    rho, phi = chain.grid.rho.to_numpy(), chain.grid.phi.to_numpy()
    exact = O.dirichlet_solve_f64((F32(4.0 * math.pi) * rho).astype(F32), phi, dx)
    e_phi = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    assert rec["phi_rel_l2_vs_exact_discrete"] == pytest.approx(e_phi, rel=1e-6)
    pos = parts.positions()
    a = np.stack(O.gather(pos, *O.grad(phi, dx), dx), 1).astype(np.float64)
    a_ex = np.stack(O.gather(pos, *O.grad(exact.astype(F32), dx), dx), 1).astype(np.float64)
    e_f = np.linalg.norm(a - a_ex) / np.linalg.norm(a_ex)
    assert rec["force_rel_l2_vs_exact_discrete"] == pytest.approx(e_f, rel=1e-5)
    fft = O.fft_solve_f64(rho, dx, 1.0)
    e_fft = np.linalg.norm(phi - fft) / np.linalg.norm(fft)
    assert rec["phi_rel_l2_vs_fft_continuum"] == pytest.approx(e_fft, rel=1e-8)
