"""The simulation driver and the device projection (reference tests/test_sim.py:20-175,
tests/test_project.py:14-47).

A dense Plummer sphere (4e11 M_sun, a = 5 kpc, 48^3 grid of 1 kpc, G in kpc/Myr/M_sun)
run through the device force chain must stay in equilibrium over > 5 dynamical times:
half-mass radius within [0.8, 1.2] of its start, virial ratio in (0.45, 0.55), momentum drift
< 1e-3 of M*sqrt(GM/a), and the direct softened pair-sum energy conserved to 2% (the
reference's bands and thresholds). Snapshots, the box-fit guard, a two-galaxy smoke run and
the projection against a host float64 histogram complete the driver's contract.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from conftest import announce

pytestmark = pytest.mark.gpu

PLUMMER_MASS, PLUMMER_A = 4.0e11, 5.0
DIRECT_PE_SOFTENING = 1.0
R_HALF_LO, R_HALF_HI = 0.80, 1.20
VIRIAL_LO, VIRIAL_HI = 0.45, 0.55
PDRIFT_TOL, EDRIFT_DIRECT = 1.0e-3, 0.02


def _plummer(n, grid, seed):
    from galaxy_collision_amd import units
    from galaxy_collision_amd.ic import plummer_ic

    c = grid / 2.0
    return plummer_ic(n, total_mass=PLUMMER_MASS, scale_a=PLUMMER_A, grav=units.G,
                      centre=(c, c, c), seed=seed)


def _config(solver, **kw):
    from galaxy_collision_amd import units
    from galaxy_collision_amd.sim import SimConfig

    base = dict(name="plummer-gate", ic_preset="plummer", solver=solver, grid_size=48, dx=1.0,
                grav=units.G, n_particles=2000, dt=0.5, steps=120, seed=7,
                total_mass=PLUMMER_MASS, scale_a=PLUMMER_A)
    base.update(kw)
    return SimConfig(**base)


def _assert_stable(history, cfg):
    from galaxy_collision_amd import units

    rh0 = history[0]["half_mass_radius"]
    rh = np.array([h["half_mass_radius"] / rh0 for h in history])
    ed0 = history[0]["energy_direct"]
    edrift = np.array([abs((h["energy_direct"] - ed0) / ed0) for h in history])
    virial = np.array([h["kinetic"] / abs(h["potential"]) for h in history])
    p0 = history[0]["momentum"]
    pscale = PLUMMER_MASS * np.sqrt(units.G * PLUMMER_MASS / PLUMMER_A)
    pdrift = np.array([np.linalg.norm(h["momentum"] - p0) for h in history]) / pscale
    n_tdyn = cfg.steps * cfg.dt / np.sqrt(PLUMMER_A ** 3 / (units.G * PLUMMER_MASS))
    print(f"\n[sim] {cfg.solver}: {n_tdyn:.1f} t_dyn, r_half [{rh.min():.3f}, {rh.max():.3f}], "
          f"virial [{virial.min():.3f}, {virial.max():.3f}], edrift {edrift.max():.2e}, "
          f"pdrift {pdrift.max():.2e}", flush=True)
    assert n_tdyn >= 5.0
    assert rh.min() > R_HALF_LO and rh.max() < R_HALF_HI, (rh.min(), rh.max())
    assert VIRIAL_LO < virial.min() and virial.max() < VIRIAL_HI, (virial.min(), virial.max())
    assert pdrift.max() < PDRIFT_TOL, pdrift.max()
    assert edrift.max() < EDRIFT_DIRECT, edrift.max()


@pytest.mark.parametrize("solver,n,dt,steps,kw", [
    ("fft", 2000, 0.5, 120, {}),
    ("multigrid", 1500, 1.0, 45, {"n_cycles": 10}),
    ("multigrid_mp", 1500, 1.0, 45, {"warm": "extrapolate"}),
])
def test_plummer_stays_stable(device, solver, n, dt, steps, kw):
    announce(f"sim plummer stability {solver}")
    from galaxy_collision_amd.sim import run_simulation

    cfg = _config(solver, n_particles=n, dt=dt, steps=steps)
    res = run_simulation(cfg, write_snapshots=False, history_cadence=8, solver_kwargs=kw,
                         direct_pe_softening=DIRECT_PE_SOFTENING, icr=_plummer(n, 48, 7))
    assert res["status"] == "ok" and res["n_particles"] == n
    _assert_stable(res["history"], cfg)


def test_snapshots_written_and_readable(device, tmp_path):
    announce("sim snapshots")
    from galaxy_collision_amd.io import read_snapshot
    from galaxy_collision_amd.sim import run_simulation

    # output_cadence 2 with 5 steps: snapshots at 0, 2, 4 and the forced final step 5
    cfg = _config("fft", grid_size=32, n_particles=500, steps=5, output_cadence=2,
                  output_dir=str(tmp_path), seed=3)
    from galaxy_collision_amd import units
    from galaxy_collision_amd.ic import plummer_ic

    icr = plummer_ic(500, total_mass=1.0e10, scale_a=2.0, grav=units.G, centre=(16.0,) * 3, seed=3)
    res = run_simulation(cfg, history_cadence=2, icr=icr, tracer_indices=[0, 7],
                         frame_cadence=2, frame_bins=64)
    assert len(res["snapshots"]) == 4
    snap = read_snapshot(res["snapshots"][-1])
    assert snap.step == 5 and snap.n == 500
    assert snap.rho is not None and snap.phi is not None and "energy" in snap.diagnostics
    assert res["tracer_path"].shape == (len(res["history"]), 2, 3)
    assert res["frames"].shape == (4, 64, 64)  # steps 0, 2, 4, 5
    # the frame (surface density over 64 bins of 0.5 kpc) holds all the mass of the compact
    # sphere, which lies inside the extent
    binw = 32 * 1.0 / 64
    assert res["frames"][-1].sum() * binw * binw == pytest.approx(
        float(icr.mass.astype(np.float64).sum()), rel=1e-3)


def test_run_rejects_ic_that_does_not_fit_box(device):
    from galaxy_collision_amd.sim import SimConfig, run_simulation

    cfg = SimConfig(ic_preset="two_galaxy", grid_size=16, dx=1.0, n_particles=2000, steps=1,
                    separation=20.0)
    with pytest.raises(ValueError, match="outside the .* box"):
        run_simulation(cfg, write_snapshots=False)


def test_two_galaxy_collision_smoke(device):
    announce("sim two-galaxy smoke")
    from galaxy_collision_amd.sim import SimConfig, run_simulation

    cfg = SimConfig(name="collision-smoke", ic_preset="two_galaxy", solver="multigrid_mp",
                    grid_size=64, dx=40.0 / 64, n_particles=20_002, dt=0.01, steps=3, seed=0)
    res = run_simulation(cfg, write_snapshots=False, history_cadence=1)
    assert res["status"] == "ok" and res["n_particles"] == 20_002
    assert np.isfinite(res["energy_drift_max"]) and len(res["history"]) == 4


def test_project_density_matches_hist2d(device):
    """Reference tests/test_project.py:14-47: the device projection equals a float64
    mass-weighted histogram2d divided by the bin area, transposed; out-of-extent particles
    are dropped. Deterministic: two projections are bitwise equal."""
    announce("projection vs histogram2d")
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.project import project_density

    n, bins = 50_000, 128
    extent = (0.0, 256.0, 0.0, 256.0)
    rng = np.random.default_rng(0)
    pos = rng.uniform(10.0, 246.0, (n, 3)).astype(np.float32)
    pos[:200] = rng.uniform(300.0, 400.0, (200, 3))  # outside: dropped
    mass = rng.uniform(500.0, 1500.0, n).astype(np.float32)
    parts = ParticleState(n)
    parts.load(pos, None, mass)
    img = torch.zeros((bins, bins), device=device)
    binw = (extent[1] - extent[0]) / bins
    for axes in [(0, 1), (0, 2), (1, 2)]:
        dev = project_density(parts, img, extent, axes=axes)
        a, b = axes
        h, _, _ = np.histogram2d(pos[:, a].astype(np.float64), pos[:, b].astype(np.float64),
                                 bins=bins, range=[extent[:2], extent[2:]],
                                 weights=mass.astype(np.float64))
        ref = h.T / (binw * binw)
        np.testing.assert_allclose(dev, ref, rtol=1e-6, atol=ref.max() * 1e-7)
        again = project_density(parts, img, extent, axes=axes)
        assert np.array_equal(dev, again)


def test_reordered_run_keeps_particle_identity(device, tmp_path):
    """With the chain moving particles into cell order every 2 steps, snapshots still list
    the particles in load order (mass and gid exactly as loaded) and tracers follow their
    particles: both agree with an unordered run to rounding level."""
    announce("sim with reordering")
    from galaxy_collision_amd import units
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.io import read_snapshot
    from galaxy_collision_amd.sim import run_simulation

    icr = plummer_ic(3000, total_mass=1.0e10, scale_a=2.0, grav=units.G, centre=(16.0,) * 3,
                     seed=4)
    res = {}
    for k in (0, 2):
        cfg = _config("multigrid", grid_size=32, n_particles=3000, steps=6, dt=0.5,
                      output_cadence=3, output_dir=str(tmp_path / f"k{k}"), seed=4,
                      reorder_every=k)
        res[k] = run_simulation(cfg, history_cadence=3, icr=icr, tracer_indices=[0, 11, 2999])
    snaps = {k: read_snapshot(res[k]["snapshots"][-1]) for k in res}
    assert np.array_equal(snaps[2].mass, icr.mass.astype(np.float32))
    assert np.array_equal(snaps[2].gid, snaps[0].gid)
    np.testing.assert_allclose(snaps[2].pos, snaps[0].pos, rtol=0, atol=1e-3)
    np.testing.assert_allclose(res[2]["tracer_path"], res[0]["tracer_path"], rtol=0, atol=1e-3)
