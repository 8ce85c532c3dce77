"""GPU parity of the bricked fixed-point deposit (PM_DEPOSIT_BRICKED; replaces reference
deposit.py:47-82 like the other modes).

Its density is the fixed-point one, so every case is bit-exact against the oracle's
oracle_deposit_fixed (itself pinned by closed forms, order independence and a float64 sum
of the same terms in tests/test_oracle.py):
- binned on every call (pm_deposit_cic) for uniform, clustered (units split inside one
  brick, whole waves in one cell), out-of-box and non-multiple-of-8 grids;
- reusing one binning while the particles move (pm_deposit_cic_step): within a cell of their
  brick, beyond it (the global-atomic path), out of and into the box;
- after the particles are moved into the binning's order (pm_deposit_plan_mark_ordered);
- the binning order is a stable argsort of the (brick, cell) keys, and the node-brick list
  holds every CIC corner node of the particles (all node bricks when one fell outside);
- PMChain runs in "bricked" mode, with reordering, equal "fixed" mode runs bit for bit.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import pm_oracle as O
from conftest import announce, assert_bitwise

pytestmark = pytest.mark.gpu
F32 = np.float32


def _parts(pos, mass):
    from galaxy_collision_amd.data import ParticleState

    parts = ParticleState(max(len(mass), 1))
    if len(mass):
        parts.load(pos, None, mass)
    parts.n = len(mass)
    return parts


def _plan(n, cap):
    from galaxy_collision_amd.deposit import DepositPlan

    plan = DepositPlan(n, max(cap, 1))
    plan.set_mode("bricked")
    return plan


def _deposit(device, plan, parts, n, dx, step=None):
    rho = torch.full((n, n, n), 7.0, device=device)
    if step is None:
        plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho)
    else:
        plan.deposit_step(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho,
                          bin=step)
    return rho.cpu().numpy()


def _check(device, pos, mass, n, dx):
    plan = _plan(n, len(mass))
    rho = _deposit(device, plan, _parts(pos, mass), n, dx)
    ref, _ = O.deposit_cic_fixed(pos, mass, n, dx)
    assert_bitwise(rho, ref)
    plan.close()
    return rho


def _bin_keys(pos, n, dx):
    """Binning key of each particle: (cell brick << 9) | cell within the brick, over the
    extended cells (ceil((n+1)/8) bricks per axis), or the sentinel (bricks << 9) when no
    corner lies in the box."""
    m = n + 1
    nbc = (m + 7) // 8
    key = O.cell_keys(pos, n, dx).astype(np.int64)
    out = key == m ** 3
    ex, ey, ez = key // (m * m), (key // m) % m, key % m
    b = ((ex >> 3) * nbc + (ey >> 3)) * nbc + (ez >> 3)
    cell = ((ex & 7) << 6) | ((ey & 7) << 3) | (ez & 7)
    return np.where(out, nbc ** 3 << 9, (b << 9) | cell)


def test_bricked_closed_forms_and_edges(device):
    announce("bricked deposit closed forms / edges")
    plan = _plan(16, 8)
    rho = _deposit(device, plan, _parts(np.array([[8.5, 8.5, 8.5]], F32), np.array([8.0], F32)),
                   16, 1.0)
    assert np.all(rho[8:10, 8:10, 8:10] == 1.0) and rho.sum() == 8.0
    rho = _deposit(device, plan, _parts(np.array([[-3.0, 8, 8], [40.0, 8, 8], [15.5, 8, 8]], F32),
                                        np.ones(3, F32)), 16, 1.0)
    assert rho[15, 8, 8] == 0.5 and rho.sum() == 0.5
    rho = _deposit(device, plan, _parts(np.zeros((0, 3), F32), np.zeros(0, F32)), 16, 1.0)
    assert not rho.any()
    plan.close()
    pos = np.array([[-5.0, 1, 1], [100.0, 2, 2], [np.nan, 3, 3], [3, np.inf, 3], [-1.0, 0, 0],
                    [7.999, 7.999, 7.999], [8.0, 1, 1], [2.5, 2.5, 2.5]], F32)
    _check(device, pos, np.ones(8, F32), 8, 1.0)


@pytest.mark.parametrize("npart,n,dx,dense,sort", [
    (1, 4, 1.0, 0, False), (4097, 8, 1.0, 0, False), (30_000, 9, 0.7, 0, False),
    (50_000, 32, 0.3125, 20_000, False), (120_000, 37, 0.5, 40_000, True),
    (300_000, 64, 1.0, 150_000, True), (1_000_003, 128, 0.5, 0, False),
    # 2^21 particles and more: the float64 LDS sums
    (2_200_000, 128, 0.5, 400_000, True), (2_100_003, 96, 0.4, 0, False)])
def test_bricked_random_bitwise(device, npart, n, dx, dense, sort):
    """Bitwise equal to the C oracle in random and in cell order, with a dense cell (one
    brick split over tens of units, whole waves in one cell), for both kinds of LDS sums
    (64-bit integers below 2^21 particles, exact float64 from there)."""
    announce(f"bricked deposit npart={npart} n={n} dense={dense} sorted={sort}")
    rng = np.random.default_rng(npart + 3)
    pos = rng.normal(n * dx / 2, n * dx / 5, size=(npart, 3)).astype(F32)
    if dense:
        pos[:dense] = rng.uniform(0.3 * n * dx, 0.3 * n * dx + 0.9 * dx, size=(dense, 3))
    pos = pos[rng.permutation(npart)]
    if sort:
        pos = pos[np.argsort(O.cell_keys(pos, n, dx), kind="stable")]
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    _check(device, pos, mass, n, dx)


@pytest.mark.parametrize("case", ["tiny_mass", "mass_range", "tiny_dx", "huge_dx",
                                  "tiny_pos", "tiny_mass_f64"])
def test_bricked_scales_and_quotient_ranges(device, case):
    """The unit pass's shortcuts stay bitwise the oracle's arithmetic at the edges of their
    ranges: fixed-point scales above 2^64 (terms scaled one by one) and far below (m' * 2^s
    folded once, with terms that underflow), masses over 33 decades, dx outside [2^-30, 2^30]
    and coordinates below 2^-60 (the IEEE division itself), zeros and signed zeros."""
    announce(f"bricked deposit scale / quotient range case {case}")
    rng = np.random.default_rng(7)
    n, dx, npart = 32, 0.5, 20_000
    if case == "tiny_mass_f64":
        npart = 2_100_000  # float64 LDS sums
    if case == "tiny_dx":
        dx = 2.0 ** -34
    elif case == "huge_dx":
        dx = 2.0 ** 31
    pos = rng.normal(n * dx / 2, n * dx / 5, size=(npart, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    if case in ("tiny_mass", "tiny_mass_f64"):
        mass *= F32(1e-9)  # m' near 2^-27: scale 2^s with s > 64
    elif case == "mass_range":
        mass = (10.0 ** rng.uniform(-30, 3, npart)).astype(F32)
    elif case == "tiny_pos":
        k = npart // 4
        pos[:k, 0] = rng.choice(np.array([0.0, -0.0, 1e-25, -1e-25, 3e-39, -3e-39, 1e-45],
                                         F32), k)
        pos[k:2 * k, 1] = rng.uniform(-1e-19, 1e-19, k).astype(F32)
        pos[2 * k:2 * k + 50, 2] = np.array([1e30, -1e30, np.inf, -np.inf, 2.0 ** 60] * 10, F32)
    _check(device, pos, mass, n, dx)


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_bricked_unaligned_arrays(device, shift):
    """Particle arrays that start 4, 8 or 12 bytes past a 16-byte boundary (read one value
    at a time instead of four) give the same bitwise result, binned anew and reused."""
    announce(f"bricked deposit, arrays offset by {shift} floats")
    rng = np.random.default_rng(11 + shift)
    n, dx, npart = 40, 0.5, 50_001
    pos = rng.normal(n * dx / 2, n * dx / 6, size=(npart, 3)).astype(F32)
    pos = pos[np.argsort(O.cell_keys(pos, n, dx), kind="stable")]
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    ref, _ = O.deposit_cic_fixed(pos, mass, n, dx)

    def shifted(a):
        t = torch.empty(npart + shift, dtype=torch.float32, device=device)
        t[shift:] = torch.from_numpy(np.ascontiguousarray(a)).to(device)
        return t[shift:]

    cols = [shifted(pos[:, a]) for a in range(3)] + [shifted(mass)]
    assert all(c.data_ptr() % 16 == 4 * shift for c in cols)
    plan = _plan(n, npart)
    for step in (None, True, False):
        rho = torch.full((n, n, n), 7.0, device=device)
        if step is None:
            plan.deposit(*cols, npart, dx, rho)
        else:
            plan.deposit_step(*cols, npart, dx, rho, bin=step)
        assert_bitwise(rho.cpu().numpy(), ref)
    plan.close()


def test_bricked_step_reuses_binning_exactly(device):
    """One binning, then the particles move without binning again: by less than a cell, by
    several cells (outside their unit's region), out of the box and into it. Every deposit
    equals the oracle, also with rho kept between the deposits (the node bricks far from the
    binned particles are then stored only after a deposit with particles outside: the one
    after it must clear what that one put there); the node-brick count covers every brick
    when a particle fell outside."""
    announce("bricked step deposits on one binning")
    n, dx = 96, 0.5
    box = n * dx
    rng = np.random.default_rng(8)
    npart = 200_000
    pos = rng.normal(box / 2, 2.5, size=(npart, 3)).astype(F32)
    pos[:5000] = rng.uniform(-2.0, -0.6, size=(5000, 3))  # no corner in the box yet
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    parts = _parts(pos, mass)
    plan = _plan(n, npart)
    nnb = ((n + 7) // 8) ** 3

    rho_t = torch.full((n, n, n), 7.0, device=device)  # one array, kept between deposits

    def check(p, step, what):
        parts.pos_x.tensor.copy_(torch.from_numpy(np.ascontiguousarray(p[:, 0])))
        parts.pos_y.tensor.copy_(torch.from_numpy(np.ascontiguousarray(p[:, 1])))
        parts.pos_z.tensor.copy_(torch.from_numpy(np.ascontiguousarray(p[:, 2])))
        plan.deposit_step(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx,
                          rho_t, bin=step, rho_kept=not step)
        rho = rho_t.cpu().numpy()
        ref, _ = O.deposit_cic_fixed(p, mass, n, dx)
        assert_bitwise(rho, ref, what)
        _, cnt, _ = plan.brick_list()
        return _read_u32(cnt)

    base = check(pos, True, "binned")
    assert 0 < base < nnb
    small = pos + rng.uniform(-0.4 * dx, 0.4 * dx, size=pos.shape).astype(F32)
    small[:5000] = pos[:5000]
    assert check(small, False, "moved < 1 cell") == base
    far = small.copy()
    far[5000:5200] += rng.uniform(-6 * dx, 6 * dx, size=(200, 3)).astype(F32)
    far[5200:5300] = rng.uniform(box + 1.0, box + 3.0, size=(100, 3))  # leave the box
    far[:2000] = rng.uniform(0.0, box, size=(2000, 3))  # enter it
    assert check(far, False, "moved several cells / through the box faces") == nnb
    assert check(small, False, "back within a cell") == base
    plan.close()


def _read_u32(addr: int) -> int:
    """A device uint32 at ``addr`` (a device count the plan owns), read through torch."""
    out = torch.empty(1, dtype=torch.int32, device="cuda")
    from galaxy_collision_amd import _native as nat

    nat.call("pm_copy_f32", nat.ptr(out, torch.int32), ctypes.c_void_p(addr), 1, nat.stream())
    torch.cuda.synchronize()
    return int(out.cpu().numpy().view(np.uint32)[0])


def test_bricked_order_and_in_place_reads(device):
    """The binning order is the stable argsort of the (brick, cell) keys; once the particles
    are moved into it (mark_ordered) later step deposits read them in place, still exact."""
    announce("bricked binning order / in-place reads")
    n, dx = 48, 0.5
    rng = np.random.default_rng(12)
    npart = 150_000
    pos = rng.normal(12.0, 3.0, size=(npart, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    parts = _parts(pos, mass)
    plan = _plan(n, npart)
    rho = _deposit(device, plan, parts, n, dx, step=True)
    assert_bitwise(rho, O.deposit_cic_fixed(pos, mass, n, dx)[0])
    addr = plan.order()
    assert addr
    order = torch.empty(npart, dtype=torch.int32, device=device)
    from galaxy_collision_amd import _native as nat

    nat.call("pm_copy_f32", nat.ptr(order, torch.int32), ctypes.c_void_p(addr), npart,
             nat.stream())
    torch.cuda.synchronize()
    perm = order.cpu().numpy()
    assert_bitwise(perm, np.argsort(_bin_keys(pos, n, dx), kind="stable").astype(np.int32))
    # move the particles into the binning's order, as the reordering drift does
    pos2, mass2 = pos[perm], mass[perm]
    parts2 = _parts(pos2, mass2)
    plan.deposit_step(parts2.pos_x, parts2.pos_y, parts2.pos_z, parts2.mass, npart, dx,
                      torch.empty((n, n, n), device=device), bin=False)
    assert plan.order() is None  # no binning in that call
    plan.mark_ordered()
    moved = pos2 + rng.uniform(-0.7 * dx, 0.7 * dx, size=pos2.shape).astype(F32)
    for k, p in enumerate((pos2, moved)):
        for a, t in zip((parts2.pos_x, parts2.pos_y, parts2.pos_z), p.T):
            a.tensor.copy_(torch.from_numpy(np.ascontiguousarray(t)))
        rho = _deposit(device, plan, parts2, n, dx, step=False)
        assert_bitwise(rho, O.deposit_cic_fixed(p, mass2, n, dx)[0], f"in place {k}")
    plan.close()


def test_bricked_node_list_covers_corners(device):
    announce("bricked node-brick list")
    n, dx = 56, 0.5
    rng = np.random.default_rng(4)
    pos = np.concatenate([rng.normal(7.0, 2.0, size=(40_000, 3)),
                          rng.normal(20.0, 1.0, size=(20_000, 3))]).astype(F32)
    mass = np.ones(len(pos), F32)
    plan = _plan(n, len(pos))
    _deposit(device, plan, _parts(pos, mass), n, dx)
    lst, cnt, mx = plan.brick_list()
    count = _read_u32(cnt)
    nb = (n + 7) // 8
    assert mx == nb ** 3
    listed = torch.empty(mx, dtype=torch.int32, device=device)
    from galaxy_collision_amd import _native as nat

    nat.call("pm_copy_f32", nat.ptr(listed, torch.int32), ctypes.c_void_p(lst), mx, nat.stream())
    torch.cuda.synchronize()
    listed = listed.cpu().numpy()
    assert sorted(listed.tolist()) == list(range(mx))  # every brick once: listed ones first
    g = np.floor(pos.astype(F32) / F32(dx)).astype(np.int64)
    need = set()
    for c in range(8):
        j = g + np.array([c >> 2, (c >> 1) & 1, c & 1])
        ok = np.all((j >= 0) & (j < n), axis=1)
        jb = j[ok] >> 3
        need.update(((jb[:, 0] * nb + jb[:, 1]) * nb + jb[:, 2]).tolist())
    assert need <= set(listed[:count].tolist())
    plan.close()


def test_chain_bricked_equals_fixed(device):
    """PMChain in "bricked" mode (binning reused between reorders, particles reordered every
    3 steps) gives the "fixed" mode's run bit for bit, particle by particle."""
    announce("PMChain bricked vs fixed, 12 steps")
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.solver import make_solver

    n, dx = 48, 0.5
    icr = plummer_ic(60_000, centre=(12.0, 11.7, 12.3), seed=6)
    out = {}
    for mode, reorder in (("fixed", 0), ("bricked", 3), ("bricked", 0)):
        parts = ParticleState(icr.n)
        parts.load(icr.pos, icr.vel, icr.mass)
        chain = PMChain(parts, n, dx, make_solver("multigrid", n, dx=dx, grav_constant=1.0),
                        deposit_mode=mode, reorder_every=reorder)
        chain.prime()
        chain.run(7, 0.02)
        chain.run(5, 0.02)
        torch.cuda.synchronize()
        pid = parts.pid.tensor.cpu().numpy() if reorder else np.arange(icr.n)
        pos, vel = np.empty((icr.n, 3), F32), np.empty((icr.n, 3), F32)
        pos[pid], vel[pid] = parts.positions(), parts.velocities()
        out[(mode, reorder)] = (pos, vel, chain.grid.rho.to_numpy())
    ref = out[("fixed", 0)]
    for key in (("bricked", 3), ("bricked", 0)):
        for a, b, what in zip(out[key], ref, ("pos", "vel", "rho")):
            assert_bitwise(a, b, f"{key} {what}")
