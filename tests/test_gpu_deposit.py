"""GPU parity of the sort-based CIC deposit (replaces reference deposit.py:47-82).

Cell keys and the stable permutation are integer work: bit-exact against np.floor and
np.argsort(kind="stable"). rho is bit-exact against the oracle's restatement of the same
summation order (oracle/pm_oracle.c), and the reference's closed forms hold.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import pm_oracle as O
from conftest import announce, assert_bitwise

pytestmark = pytest.mark.gpu
F32 = np.float32


def _run(device, pos, mass, n, dx, offset=0):
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import deposit_keys

    parts = ParticleState(max(len(mass), 1))
    if len(mass):
        parts.load(pos, None, mass)
    parts.n = len(mass)
    # deposit must overwrite everything; offset > 0 makes rho a view that is not 16-byte
    # aligned (the node assembly then stores scalars instead of float4)
    buf = torch.full((n * n * n + offset,), 7.0, device=device)
    rho = buf[offset:].view(n, n, n)
    keys, perm = deposit_keys(parts, rho, dx)
    assert bool((buf[:offset] == 7.0).all())
    return rho.cpu().numpy(), keys.cpu().numpy().view(np.uint32), perm.cpu().numpy()


def _check(device, pos, mass, n, dx, offset=0):
    rho, keys, perm = _run(device, pos, mass, n, dx, offset)
    ref_rho, ref_keys, ref_perm = O.deposit_cic(pos, mass, n, dx)
    assert_bitwise(keys, O.cell_keys(pos, n, dx))
    assert_bitwise(keys, ref_keys)
    assert_bitwise(perm, O.stable_perm(keys))
    assert_bitwise(rho, ref_rho)
    return rho


def test_closed_forms(device):
    announce("deposit closed forms")
    rho, _, _ = _run(device, np.array([[8.0, 8, 8]], F32), np.array([3.0], F32), 16, 1.0)
    assert rho[8, 8, 8] == 3.0 and rho.sum() == 3.0
    rho, _, _ = _run(device, np.array([[8.5, 8.5, 8.5]], F32), np.array([8.0], F32), 16, 1.0)
    assert np.all(rho[8:10, 8:10, 8:10] == 1.0) and rho.sum() == 8.0
    rho, _, _ = _run(device, np.array([[-3.0, 8, 8], [40.0, 8, 8], [15.5, 8, 8]], F32),
                     np.ones(3, F32), 16, 1.0)
    assert rho[15, 8, 8] == 0.5 and rho[0].sum() == 0.0 and rho.sum() == 0.5
    dx = 2.5
    rho, _, _ = _run(device, np.array([[2 * dx, 3 * dx, 4 * dx]], F32), np.array([5.0], F32), 8, dx)
    assert rho[2, 3, 4] == pytest.approx(5.0 / dx ** 3, rel=1e-6)


@pytest.mark.parametrize("seed,n,dx", [(0, 64, 1.0), (3, 32, 2.5)])
def test_mass_conservation(device, seed, n, dx):
    announce(f"deposit mass conservation seed={seed}")
    rng = np.random.default_rng(seed)
    pos = rng.uniform(2.0 * dx, (n - 3) * dx, size=(5000, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, size=5000).astype(F32)
    rho = _check(device, pos, mass, n, dx)
    assert rho.astype(np.float64).sum() * dx ** 3 == pytest.approx(mass.astype(np.float64).sum(), rel=1e-4)


def test_edge_cases(device):
    announce("deposit edge cases")
    # empty
    rho, keys, perm = _run(device, np.zeros((0, 3), F32), np.zeros(0, F32), 8, 1.0)
    assert not rho.any() and keys.size == 0
    # everything outside the box
    pos = np.array([[-5.0, 1, 1], [100.0, 2, 2], [np.nan, 3, 3], [3, np.inf, 3]], F32)
    _check(device, pos, np.ones(4, F32), 8, 1.0)
    # exactly on the -1 and n boundaries of the key range
    pos = np.array([[-1.0, 0, 0], [7.999, 7.999, 7.999], [8.0, 1, 1], [-1.0001, 2, 2]], F32)
    _check(device, pos, np.ones(4, F32), 8, 1.0)


@pytest.mark.parametrize("npart,n,dx,dense,offset", [
    (1, 4, 1.0, 0, 0), (17, 4, 1.0, 17, 0), (4097, 8, 1.0, 0, 0), (50_000, 32, 0.3125, 20_000, 0),
    (300_000, 64, 1.0, 150_000, 0), (1_000_003, 128, 0.5, 0, 0), (20_000, 37, 0.7, 3000, 0),
    (40_000, 30, 1.0, 0, 1), (30_000, 32, 1.0, 500, 3), (2_000_000, 256, 0.15625, 40_000, 0)])
def test_random_bitwise(device, npart, n, dx, dense, offset):
    """Includes a cell holding `dense` particles: thousands of pieces, the multi-batch tree.
    Grid sizes that are not multiples of 4 and unaligned rho views cover the node assembly's
    scalar-store path; 256^3 (25-bit keys) the 9-bit radix digit."""
    announce(f"deposit random npart={npart} n={n} dense={dense} offset={offset}")
    rng = np.random.default_rng(npart)
    pos = rng.normal(n * dx / 2, n * dx / 5, size=(npart, 3)).astype(F32)
    if dense:
        pos[:dense] = rng.uniform(0.3 * n * dx, 0.3 * n * dx + 0.9 * dx, size=(dense, 3))
        pos = pos[rng.permutation(npart)]
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    _check(device, pos, mass, n, dx, offset)


@pytest.mark.parametrize("dense", [129, 1000, 16_383, 16_384 + 7, 32_768 + 1000])
def test_populous_cell_trees(device, dense):
    """One cell with `dense` particles after a few sparse ones: its pieces span 9..2111
    chunks with arbitrary alignment, covering the in-thread, one-batch and multi-batch
    tree paths of the segmented reduction."""
    announce(f"deposit populous cell dense={dense}")
    rng = np.random.default_rng(dense)
    n = 16
    sparse = rng.uniform(0.0, 5.0, size=(37, 3)).astype(F32)
    cell = rng.uniform(9.0, 9.999, size=(dense, 3)).astype(F32)
    pos = np.concatenate([sparse, cell, rng.uniform(10.0, 15.0, size=(11, 3)).astype(F32)])
    pos = pos[rng.permutation(len(pos))]
    mass = rng.uniform(0.5, 2.0, len(pos)).astype(F32)
    _check(device, pos, mass, n, 1.0)


def test_deterministic_repeat(device):
    announce("deposit determinism")
    rng = np.random.default_rng(11)
    pos = rng.normal(32, 4, size=(400_000, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, 400_000).astype(F32)
    a, _, _ = _run(device, pos, mass, 64, 1.0)
    b, _, _ = _run(device, pos, mass, 64, 1.0)
    assert_bitwise(a, b)


def test_dispatch_recorded_timing(device):
    """Timing mode 2: only the radix scatter's dispatch-recorded pairs are kept (one per sort
    pass, with the pass's key/index bytes), and the deposit is bitwise the untimed one."""
    announce("deposit dispatch-recorded timing")
    from galaxy_collision_amd import _native as nat

    rng = np.random.default_rng(12)
    npart = 300_000
    pos = rng.normal(32, 4, size=(npart, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    ref, _, _ = _run(device, pos, mass, 64, 1.0)
    nat.timing_reset()
    nat.timing_enable(2)
    try:
        got, _, _ = _run(device, pos, mass, 64, 1.0)
        torch.cuda.synchronize()
    finally:
        nat.timing_enable(False)
    rep = nat.timing_report()
    nat.timing_reset()
    assert_bitwise(got, ref)
    assert set(rep) == {"deposit.radix_scatter"}, sorted(rep)
    r = rep["deposit.radix_scatter"]
    passes = r["launches"]
    assert passes >= 1 and r["total_ms"] > 0.0
    # 12 bytes per particle in the first pass (key read, key and index written), 16 after
    assert r["bytes"] == pytest.approx((12.0 + 16.0 * (passes - 1)) * npart)


# ---------------------------------------------------------------- fixed-point mode

def _run_fixed(device, pos, mass, n, dx):
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import deposit_plan

    parts = ParticleState(max(len(mass), 1))
    if len(mass):
        parts.load(pos, None, mass)
    parts.n = len(mass)
    rho = torch.full((n, n, n), 7.0, device=device)
    keys = torch.empty(max(parts.n, 1), dtype=torch.int32, device=device)
    plan = deposit_plan(n, parts.n)
    plan.set_mode("fixed")
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho, keys)
    return rho.cpu().numpy(), keys.cpu().numpy().view(np.uint32)[: parts.n]


def _check_fixed(device, pos, mass, n, dx):
    rho, keys = _run_fixed(device, pos, mass, n, dx)
    ref_rho, ref_keys = O.deposit_cic_fixed(pos, mass, n, dx)
    assert_bitwise(keys, ref_keys)
    assert_bitwise(rho, ref_rho)
    return rho


def test_fixed_closed_forms_and_edges(device):
    announce("fixed deposit closed forms / edges")
    rho, _ = _run_fixed(device, np.array([[8.5, 8.5, 8.5]], F32), np.array([8.0], F32), 16, 1.0)
    assert np.all(rho[8:10, 8:10, 8:10] == 1.0) and rho.sum() == 8.0
    rho, _ = _run_fixed(device, np.array([[-3.0, 8, 8], [40.0, 8, 8], [15.5, 8, 8]], F32),
                        np.ones(3, F32), 16, 1.0)
    assert rho[15, 8, 8] == 0.5 and rho.sum() == 0.5
    rho, _ = _run_fixed(device, np.zeros((0, 3), F32), np.zeros(0, F32), 8, 1.0)
    assert not rho.any()
    pos = np.array([[-5.0, 1, 1], [100.0, 2, 2], [np.nan, 3, 3], [3, np.inf, 3], [-1.0, 0, 0],
                    [7.999, 7.999, 7.999], [8.0, 1, 1], [2.5, 2.5, 2.5]], F32)
    _check_fixed(device, pos, np.ones(8, F32), 8, 1.0)


@pytest.mark.parametrize("npart,n,dx,dense,sort", [
    (1, 4, 1.0, 0, False), (4097, 8, 1.0, 0, False), (50_000, 32, 0.3125, 20_000, False),
    (300_000, 64, 1.0, 150_000, True), (1_000_003, 128, 0.5, 0, True),
    (1_000_003, 128, 0.5, 0, False)])
def test_fixed_random_bitwise(device, npart, n, dx, dense, sort):
    """Bitwise equal to the C oracle in random order and in cell order (long wave runs),
    with a dense cell (one atomic target shared by many waves)."""
    announce(f"fixed deposit npart={npart} n={n} dense={dense} sorted={sort}")
    rng = np.random.default_rng(npart + 1)
    pos = rng.normal(n * dx / 2, n * dx / 5, size=(npart, 3)).astype(F32)
    if dense:
        pos[:dense] = rng.uniform(0.3 * n * dx, 0.3 * n * dx + 0.9 * dx, size=(dense, 3))
    pos = pos[rng.permutation(npart)]
    if sort:
        pos = pos[np.argsort(O.cell_keys(pos, n, dx), kind="stable")]
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    rho = _check_fixed(device, pos, mass, n, dx)
    # the sorted mode gives the same density to float32 rounding of a different order
    ref_sorted, _, _ = O.deposit_cic(pos, mass, n, dx)
    assert np.abs(rho - ref_sorted).max() <= 4e-6 * np.abs(ref_sorted).max()


def test_fixed_order_independent(device):
    announce("fixed deposit order independence")
    rng = np.random.default_rng(5)
    pos = rng.normal(32, 4, size=(400_000, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, 400_000).astype(F32)
    a, _ = _run_fixed(device, pos, mass, 64, 1.0)
    order = rng.permutation(400_000)
    b, _ = _run_fixed(device, pos[order], mass[order], 64, 1.0)
    assert_bitwise(a, b)


def test_plan_reuse_across_disjoint_particle_sets(device):
    """One plan deposits particle sets in different regions in turn (with a fixed-mode
    deposit in between): the map is only reset around the previous deposit's occupied
    bricks, so a stale entry would show up as a wrong density. Each result must equal the
    oracle bit for bit."""
    announce("deposit plan reuse")
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import DepositPlan

    n, dx = 48, 0.5
    rng = np.random.default_rng(21)
    plan = DepositPlan(n, 60_000)
    sets = [rng.normal(6.0, 1.0, size=(50_000, 3)), rng.normal(17.0, 1.5, size=(30_000, 3)),
            rng.uniform(-1.0, n * dx + 1.0, size=(60_000, 3)), np.zeros((0, 3))]
    for q, pos in enumerate(sets):
        pos = pos.astype(F32)
        mass = rng.uniform(0.5, 2.0, len(pos)).astype(F32)
        parts = ParticleState(max(len(pos), 1))
        if len(pos):
            parts.load(pos, None, mass)
        parts.n = len(pos)
        rho = torch.full((n, n, n), 7.0, device=device)
        if q == 2:  # a fixed-mode deposit leaves the sorted map alone
            plan.set_mode("fixed")
            plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho)
            plan.set_mode("sorted")
        plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho)
        ref, _, _ = O.deposit_cic(pos, mass, n, dx)
        assert_bitwise(rho.cpu().numpy(), ref, f"set {q}")
    plan.close()


def test_drift_keys_and_keyed_deposit(device):
    """The fused drift pass writes each moved particle's cell key exactly as the deposit's
    key pass computes it (out-of-box particles included), and the deposit that takes those
    keys gives the same density and particle order bit for bit."""
    announce("drift keys + keyed deposit")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import deposit_plan

    rng = np.random.default_rng(21)
    n, dx, npart = 48, 0.75, 200_003
    box = n * dx
    pos = rng.normal(box / 2, box / 6, size=(npart, 3)).astype(F32)
    pos[:1000] = rng.uniform(-2 * dx, box + 2 * dx, size=(1000, 3))  # edges and outside
    pos[1000:1100] = np.float32(box - 0.5 * dx)  # one crowded cell at the far edge
    vel = rng.normal(0, 1, size=(npart, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    parts = ParticleState(npart)
    parts.load(pos, vel, mass)
    g4 = torch.randn((n, n, n, 4), device=device)
    keys = torch.empty(npart, dtype=torch.int32, device=device)
    p = lambda f: nat.ptr(f.tensor)  # noqa: E731
    nat.call("pm_gather_kick_drift4", p(parts.pos_x), p(parts.pos_y), p(parts.pos_z),
             nat.ptr(g4), p(parts.vel_x), p(parts.vel_y), p(parts.vel_z), npart, n, dx,
             0.01, 0.01, 0.05, nat.ptr(keys, torch.int32), nat.stream())
    plan = deposit_plan(n, npart)
    plan.set_mode("sorted")
    rho_a = torch.full((n, n, n), 7.0, device=device)
    rho_b = torch.full((n, n, n), 7.0, device=device)
    keys_ref = torch.empty(npart, dtype=torch.int32, device=device)
    perm_a = torch.empty(npart, dtype=torch.int32, device=device)
    perm_b = torch.empty(npart, dtype=torch.int32, device=device)
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, npart, dx, rho_a, keys_ref,
                 perm_a)
    keys_before = keys.clone()
    plan.deposit_keyed(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, npart, dx, rho_b,
                       keys, perm_b)
    assert_bitwise(_np_u32(keys), _np_u32(keys_ref))
    assert_bitwise(_np_u32(keys), _np_u32(keys_before))  # only read
    assert_bitwise(rho_b.cpu().numpy(), rho_a.cpu().numpy())
    assert_bitwise(perm_b.cpu().numpy(), perm_a.cpu().numpy())
    sentinel = (n + 1) ** 3
    k = _np_u32(keys)
    assert (k == sentinel).any() and (k < sentinel).any()
    # and against the oracle on the moved positions
    ref_rho, ref_keys = O.deposit_cic(parts.positions(), mass, n, dx)[:2]
    assert_bitwise(k, ref_keys)
    assert_bitwise(rho_b.cpu().numpy(), ref_rho)


def _np_u32(t):
    return t.cpu().numpy().view(np.uint32)
