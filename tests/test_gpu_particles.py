"""GPU parity: gradient, gather, kick/drift and their fused forms, direct force, KDK.

Every kernel keeps the reference's float32 operation order, so results are compared
bit for bit with the numpy restatement in the oracle (reference deposit.py:85-174,
integrator.py:27-131).
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


def _t(a, device):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=F32)).to(device)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("n,dx", [(2, 1.0), (17, 0.3125), (64, 1.0)])
def test_grad_bitwise(device, n, dx):
    announce(f"grad n={n} dx={dx}")
    from galaxy_collision_amd.deposit import potential_to_accel

    phi = np.random.default_rng(n).normal(size=(n, n, n)).astype(F32)
    g = [torch.empty((n, n, n), device=device) for _ in range(3)]
    potential_to_accel(_t(phi, device), *g, dx=dx)
    ref = O.grad(phi, dx)
    for a in range(3):
        assert_bitwise(_np(g[a]), ref[a])
    # interleaved variant of the fused step: g4[node] = (gx, gy, gz, 0)
    from galaxy_collision_amd import _native as nat

    phi_d = _t(phi, device)
    g4 = torch.full((n, n, n, 4), 7.0, device=device)
    nat.call("pm_grad4", nat.ptr(phi_d), nat.ptr(g4), n, dx, nat.stream())
    g4h = _np(g4)
    for a in range(3):
        assert_bitwise(g4h[..., a], ref[a])
    assert not np.any(g4h[..., 3])


def _particles(rng, npart, n, dx, spread=1.5):
    pos = rng.uniform(-spread * dx, (n + spread) * dx, size=(npart, 3)).astype(F32)
    pos[: npart // 8] = rng.normal(n * dx / 2, dx, size=(npart // 8, 3))
    return pos


@pytest.mark.parametrize("npart,n,dx", [(1, 8, 1.0), (1000, 16, 1.0), (200_003, 64, 0.3125)])
def test_gather_and_fused_gather_kick_bitwise(device, npart, n, dx):
    announce(f"gather npart={npart} n={n}")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import gather_acceleration

    rng = np.random.default_rng(npart)
    pos = _particles(rng, npart, n, dx)
    grids = [rng.normal(size=(n, n, n)).astype(F32) for _ in range(3)]
    parts = ParticleState(npart)
    parts.load(pos, rng.normal(size=(npart, 3)).astype(F32), np.ones(npart, F32))
    acc = [torch.empty(npart, device=device) for _ in range(3)]
    gt = [_t(g, device) for g in grids]
    gather_acceleration(parts, *gt, *acc, dx=dx)
    ref = O.gather(pos, *grids, dx)
    for a in range(3):
        assert_bitwise(_np(acc[a]), ref[a])
    # fused gather + kick == gather then kick
    v0 = parts.velocities()
    acc2 = [torch.empty(npart, device=device) for _ in range(3)]
    h = 0.005
    nat.call("pm_gather_kick", *(nat.ptr(f.tensor) for f in (parts.pos_x, parts.pos_y, parts.pos_z)),
             *(nat.ptr(g) for g in gt), *(nat.ptr(a) for a in acc2),
             *(nat.ptr(f.tensor) for f in (parts.vel_x, parts.vel_y, parts.vel_z)),
             npart, n, dx, h, nat.stream())
    v1 = parts.velocities()
    for a in range(3):
        assert_bitwise(_np(acc2[a]), ref[a])
        assert_bitwise(v1[:, a], O.kick(v0[:, a], ref[a], h))
    # the interleaved-field variant used by the fused step
    g4 = torch.stack([*gt, torch.zeros_like(gt[0])], -1).contiguous()
    acc3 = [torch.empty(npart, device=device) for _ in range(3)]
    nat.call("pm_gather_kick4", *(nat.ptr(f.tensor) for f in (parts.pos_x, parts.pos_y, parts.pos_z)),
             nat.ptr(g4), *(nat.ptr(a) for a in acc3),
             *(nat.ptr(f.tensor) for f in (parts.vel_x, parts.vel_y, parts.vel_z)),
             npart, n, dx, h, nat.stream())
    v2 = parts.velocities()
    for a in range(3):
        assert_bitwise(_np(acc3[a]), ref[a])
        assert_bitwise(v2[:, a], O.kick(v1[:, a], ref[a], h))


@pytest.mark.parametrize("dx", [2.0 ** -34, 0.3125, 2.0 ** 31])
def test_gather4_quotient_ranges(device, dx):
    """The interleaved gather forms x/dx from a refined reciprocal where that equals the IEEE
    quotient (dx in [2^-30, 2^30], every |coordinate| >= 2^-60) and divides otherwise: both
    sides of each bound, zeros, subnormals and non-finite coordinates, bitwise vs the
    oracle's float32 division."""
    announce(f"gather4 quotient ranges dx={dx}")
    from galaxy_collision_amd import _native as nat

    n = 16
    rng = np.random.default_rng(7)
    npart = 4096
    pos = _particles(rng, npart, n, F32(dx))
    special = np.array([0.0, -0.0, 1e-40, -1e-40, 1e-20, 2.0 ** -60, 2.0 ** -61, -(2.0 ** -60),
                        2.0 ** 60, -(2.0 ** 61), np.inf, -np.inf, np.nan, 3.0e38], dtype=F32)
    k = len(special)
    for axis in range(3):
        rows = slice(axis * k, (axis + 1) * k)
        pos[rows, axis] = special
    # a coordinate near 2^-60 on one axis while the others lie in the box
    pos[3 * k:3 * k + 64, 0] = (np.arange(64, dtype=F32) - 32) * F32(2.0 ** -66)
    grids = [rng.normal(size=(n, n, n)).astype(F32) for _ in range(3)]
    with np.errstate(over="ignore", invalid="ignore"):  # 3e38/dx and the non-finite rows
        ref = O.gather(pos, *grids, F32(dx))
    g4 = torch.from_numpy(np.stack([*grids, np.zeros_like(grids[0])], -1)).to(device).contiguous()
    p = [_t(pos[:, a], device) for a in range(3)]
    vel = [torch.zeros(npart, device=device) for _ in range(3)]
    acc = [torch.empty(npart, device=device) for _ in range(3)]
    nat.call("pm_gather_kick4", *(nat.ptr(t) for t in p), nat.ptr(g4), *(nat.ptr(a) for a in acc),
             *(nat.ptr(v) for v in vel), npart, n, float(dx), 0.0, nat.stream())
    for a in range(3):
        assert_bitwise(_np(acc[a]), ref[a])


@pytest.mark.parametrize("mode,how", [("fixed", "step"), ("sorted", "step"), ("sorted", "run"),
                                      ("fixed", "run")])
def test_fused_chain_step_equals_unfused_sequence(device, mode, how):
    """PMChain.step (kick+drift fused, interleaved gradient, gather+kick fused) and
    PMChain.run (one step's closing kick, the next one's opening kick and drift in one pass)
    equal the reference step kick -> drift -> deposit -> solve -> grad -> gather -> kick bit
    for bit (reference integrator.py:100-131 with sim.py:318-325's accel_fn)."""
    announce(f"fused chain {how} vs unfused ({mode} deposit)")
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import (deposit_density, gather_acceleration,
                                              potential_to_accel)
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.integrator import drift, kick
    from galaxy_collision_amd.solver import make_solver

    n, dx, dt = 32, 0.5, 0.02
    icr = plummer_ic(5000, total_mass=1.0, scale_a=1.0, centre=(8.0, 8.0, 8.0), seed=5)

    def state():
        p = ParticleState(icr.n)
        p.load(icr.pos, icr.vel, icr.mass)
        return p

    a = state()
    chain = PMChain(a, n, dx, make_solver("multigrid", n, dx=dx, grav_constant=1.0),
                    deposit_mode=mode)
    chain.prime()
    if how == "run":
        chain.run(3, dt)
    else:
        for _ in range(3):
            chain.step(dt)
    b = state()
    solver = make_solver("multigrid", n, dx=dx, grav_constant=1.0)
    rho, phi = (torch.zeros((n, n, n), device=device) for _ in range(2))
    g = [torch.zeros((n, n, n), device=device) for _ in range(3)]
    acc = [torch.zeros(icr.n, device=device) for _ in range(3)]
    fields = lambda p: (p.pos_x.tensor, p.pos_y.tensor, p.pos_z.tensor)  # noqa: E731
    vels = lambda p: (p.vel_x.tensor, p.vel_y.tensor, p.vel_z.tensor)  # noqa: E731

    def accel(warm):
        deposit_density(b, rho, dx, mode=mode)
        solver.solve(rho, phi, warm_start=warm)
        potential_to_accel(phi, *g, dx=dx)
        gather_acceleration(b, *g, *acc, dx=dx)

    accel(False)
    for _ in range(3):
        kick(*vels(b), *acc, icr.n, 0.5 * dt)
        drift(*fields(b), *vels(b), icr.n, dt)
        accel(True)
        kick(*vels(b), *acc, icr.n, 0.5 * dt)
    assert_bitwise(a.positions(), b.positions(), "positions")
    assert_bitwise(a.velocities(), b.velocities(), "velocities")
    for k in range(3):
        assert_bitwise(_np(chain.acc[k]), _np(acc[k]), f"acc[{k}]")
    assert_bitwise(chain.grid.phi.to_numpy(), _np(phi), "phi")


def test_kick_drift_and_fused_bitwise(device):
    announce("kick/drift")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.integrator import drift, kick

    rng = np.random.default_rng(3)
    npart = 100_001
    x, v, a = (rng.normal(size=(3, npart)).astype(F32) for _ in range(3))
    tx, tv, ta = ([_t(r, device) for r in arr] for arr in (x, v, a))
    kick(*tv, *ta, npart, 0.005)
    drift(*tx, *tv, npart, 0.01)
    for q in range(3):
        vq = O.kick(v[q], a[q], 0.005)
        assert_bitwise(_np(tv[q]), vq)
        assert_bitwise(_np(tx[q]), O.drift(x[q], vq, 0.01))
    fx, fv = [_t(r, device) for r in x], [_t(r, device) for r in v]
    nat.call("pm_kick_drift", *(nat.ptr(t) for t in fx), *(nat.ptr(t) for t in fv),
             *(nat.ptr(t) for t in ta), npart, 0.005, 0.01, nat.stream())
    for q in range(3):
        assert_bitwise(_np(fx[q]), _np(tx[q]))
        assert_bitwise(_np(fv[q]), _np(tv[q]))


def test_direct_accel_matches_float64(device):
    announce("direct_accel")
    from galaxy_collision_amd.integrator import direct_accel

    rng = np.random.default_rng(4)
    npart = 700
    pos = rng.normal(size=(npart, 3)).astype(F32)
    m = rng.uniform(0.5, 1.5, npart).astype(F32)
    acc = [torch.empty(npart, device=device) for _ in range(3)]
    direct_accel(*[_t(pos[:, a], device) for a in range(3)], _t(m, device), *acc, npart, 1.0, 0.01)
    ref = O.direct_accel_f64(pos, m, 1.0, 0.01)
    got = np.stack([_np(a) for a in acc], 1)
    assert np.abs(got - ref).max() < 1e-4 * np.abs(ref).max()


def _two_body(m, r, v_each):
    from galaxy_collision_amd.data import ParticleState

    parts = ParticleState(2)
    parts.load(np.array([[-r / 2, 0, 0], [r / 2, 0, 0]], F32),
               np.array([[0, -v_each, 0], [0, v_each, 0]], F32), np.array([m, m], F32))
    return parts


def test_kdk_step_one_eval_and_exact_inline_equality(device):
    """kdk_step == the naive 2-evaluation KDK loop bit for bit (integrator.py:100-131,
    reference tests/test_integrator.py:189-257); energy drift bounded over one orbit."""
    announce("kdk exact equality")
    from galaxy_collision_amd.integrator import direct_accel, drift, kdk_step, kick

    G = 4.498e-12
    m, r, eps = 1.0e10, 10.0, 0.01
    v_rel = np.sqrt(G * 2 * m / r)
    period = 2 * np.pi * np.sqrt(r ** 3 / (G * 2 * m))
    dt, steps = period / 1000.0, 1000
    parts = _two_body(m, r, 0.5 * v_rel)
    acc = [torch.zeros(2, device=device) for _ in range(3)]
    calls = {"n": 0}

    def accel_fn():
        calls["n"] += 1
        direct_accel(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, *acc, 2, G, eps ** 2)

    pos0, vel0 = parts.positions().astype(np.float64), parts.velocities().astype(np.float64)
    accel_fn()
    for _ in range(steps):
        kdk_step(parts, *acc, accel_fn, dt)
    assert calls["n"] == steps + 1
    pos_k, vel_k = parts.positions(), parts.velocities()

    p2 = _two_body(m, r, 0.5 * v_rel)
    b = [torch.zeros(2, device=device) for _ in range(3)]
    for _ in range(steps):
        direct_accel(p2.pos_x, p2.pos_y, p2.pos_z, p2.mass, *b, 2, G, eps ** 2)
        kick(p2.vel_x, p2.vel_y, p2.vel_z, *b, 2, dt / 2)
        drift(p2.pos_x, p2.pos_y, p2.pos_z, p2.vel_x, p2.vel_y, p2.vel_z, 2, dt)
        direct_accel(p2.pos_x, p2.pos_y, p2.pos_z, p2.mass, *b, 2, G, eps ** 2)
        kick(p2.vel_x, p2.vel_y, p2.vel_z, *b, 2, dt / 2)
    assert_bitwise(pos_k, p2.positions(), 'pos')
    assert_bitwise(vel_k, p2.velocities(), 'vel')

    mm = np.array([m, m])

    def energy(p, v):
        d = np.linalg.norm(p[1] - p[0])
        return 0.5 * (mm * (v ** 2).sum(1)).sum() - G * m * m / np.sqrt(d ** 2 + eps ** 2)

    e0 = energy(pos0, vel0)
    ef = energy(pos_k.astype(np.float64), vel_k.astype(np.float64))
    assert abs((ef - e0) / e0) < 1e-3
    assert np.linalg.norm(pos_k[0] - pos0[0]) / r < 0.02


@pytest.mark.parametrize("n,npart", [(32, 3000), (37, 5000), (64, 40_000)])
def test_grad4_on_deposit_bricks(device, n, npart):
    """pm_grad4_bricks writes pm_grad4's values (bitwise) on every node of the 8^3 bricks the
    last deposit marked, leaves the other nodes alone, and the marked bricks cover every CIC
    corner node of every deposited particle (so the gather reads only written nodes)."""
    announce(f"grad4 on deposit bricks n={n}")
    import ctypes

    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.deposit import DepositPlan

    dx = 0.5
    rng = np.random.default_rng(n)
    # a compact blob, a uniform slab over the lower 40% of x (some particles outside the
    # box), so that some bricks stay unmarked
    slab = rng.uniform(-1.0, n * dx + 1.0, size=(npart - npart // 2, 3))
    slab[:, 0] = rng.uniform(-1.0, 0.4 * n * dx, size=len(slab))
    pos = np.concatenate([rng.normal(n * dx * 0.3, 1.0, size=(npart // 2, 3)), slab])
    pos = pos.astype(F32)
    parts = ParticleState(npart)
    parts.load(pos, None, np.ones(npart, F32))
    rho = torch.zeros((n, n, n), device=device)
    plan = DepositPlan(n, npart)
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, npart, dx, rho)
    ptr = plan.node_bricks()
    assert ptr
    nb = -(-n // 8)
    torch.cuda.synchronize()
    marks = np.zeros(nb ** 3, dtype=np.uint8)
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipMemcpy(ctypes.c_void_p(marks.ctypes.data), ctypes.c_void_p(ptr),
                         ctypes.c_size_t(marks.nbytes), 2) == 0  # device to host
    phi = torch.from_numpy(rng.normal(size=(n, n, n)).astype(F32)).to(device)
    full = torch.empty((n, n, n, 4), device=device)
    part = torch.full((n, n, n, 4), 12345.0, device=device)
    nat.call("pm_grad4", nat.ptr(phi), nat.ptr(full), n, dx, nat.stream())
    nat.call("pm_grad4_bricks", nat.ptr(phi), nat.ptr(part), n, dx, ctypes.c_void_p(ptr),
             nat.stream())
    full_np, part_np = _np(full), _np(part)
    m = marks.reshape(nb, nb, nb).repeat(8, 0).repeat(8, 1).repeat(8, 2)[:n, :n, :n] != 0
    assert m.any() and not m.all()
    assert_bitwise(part_np[m], full_np[m])
    assert np.all(part_np[~m] == 12345.0)
    # the compacted list holds exactly the marked bricks, and the listed gradient writes
    # the same nodes with the same values
    lst, cnt_ptr, mx = plan.brick_list()
    assert mx == nb ** 3
    cnt = np.zeros(1, dtype=np.uint32)
    assert hip.hipMemcpy(ctypes.c_void_p(cnt.ctypes.data), ctypes.c_void_p(cnt_ptr),
                         ctypes.c_size_t(4), 2) == 0
    ids = np.zeros(max(int(cnt[0]), 1), dtype=np.int32)
    assert hip.hipMemcpy(ctypes.c_void_p(ids.ctypes.data), ctypes.c_void_p(lst),
                         ctypes.c_size_t(4 * int(cnt[0])), 2) == 0
    assert np.array_equal(np.sort(ids[: int(cnt[0])]), np.flatnonzero(marks))
    listed = torch.full((n, n, n, 4), 12345.0, device=device)
    nat.call("pm_grad4_listed", nat.ptr(phi), nat.ptr(listed), n, dx, ctypes.c_void_p(lst),
             ctypes.c_void_p(cnt_ptr), mx, nat.stream())
    assert_bitwise(_np(listed), part_np)
    # every in-box CIC corner node of every particle lies in a marked brick
    base = np.floor(pos.astype(np.float64) / dx).astype(np.int64)
    for c in range(8):
        node = base + np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
        nd = node[np.all((node >= 0) & (node < n), axis=1)]
        assert np.all(m[nd[:, 0], nd[:, 1], nd[:, 2]]), c
    plan.set_mode("fixed")
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, npart, dx, rho)
    assert plan.node_bricks() is None and plan.brick_list() is None
    plan.close()


def test_reordering_drift_moves_particles_bitwise(device):
    """pm_gather_kick_drift4_reorder: slot i of the destination holds particle order[i]
    after exactly pm_gather_kick_drift4's arithmetic (position, velocity, key), with its
    mass, gid and pid moved along; an order entry outside [0, n) keeps the slot's own
    particle."""
    announce("reordering drift")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.data import ParticleState

    rng = np.random.default_rng(8)
    n, dx, npart = 40, 0.5, 150_001
    box = n * dx
    pos = rng.uniform(-dx, box + dx, size=(npart, 3)).astype(F32)
    vel = rng.normal(0, 1, size=(npart, 3)).astype(F32)
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    gid = rng.integers(0, 2, npart).astype(np.int32)
    g4 = torch.randn((n, n, n, 4), device=device)
    h, dt = 0.01, 0.02

    a = ParticleState(npart)
    a.load(pos, vel, mass, gid)
    keys_a = torch.empty(npart, dtype=torch.int32, device=device)
    p = lambda f: nat.ptr(f.tensor)  # noqa: E731
    nat.call("pm_gather_kick_drift4", p(a.pos_x), p(a.pos_y), p(a.pos_z), nat.ptr(g4),
             p(a.vel_x), p(a.vel_y), p(a.vel_z), npart, n, dx, h, h, dt,
             nat.ptr(keys_a, torch.int32), nat.stream())

    order = rng.permutation(npart).astype(np.int32)
    order[5] = -3          # not a permutation entry: slot 5 keeps particle 5
    order[6] = npart + 9   # likewise
    b = ParticleState(npart)
    b.load(pos, vel, mass, gid)
    ot = torch.from_numpy(order).to(device)
    keys_b = torch.empty(npart, dtype=torch.int32, device=device)
    nat.call("pm_gather_kick_drift4_reorder", nat.ptr(ot, torch.int32),
             ctypes.byref(b.arrays()), ctypes.byref(b.arrays(shadow=True)), nat.ptr(g4),
             npart, n, dx, h, h, dt, nat.ptr(keys_b, torch.int32), nat.stream())
    b.swap_shadow()
    src = order.astype(np.int64)
    src[5], src[6] = 5, 6
    assert_bitwise(b.positions(), a.positions()[src], "positions")
    assert_bitwise(b.velocities(), a.velocities()[src], "velocities")
    assert_bitwise(b.mass.to_numpy(), mass[src], "mass")
    assert_bitwise(b.gid.to_numpy(), gid[src], "gid")
    assert_bitwise(b.pid.to_numpy(), src.astype(np.int32), "pid")
    assert_bitwise(keys_b.cpu().numpy(), keys_a.cpu().numpy()[src], "keys")


def test_chain_reorder_every(device):
    """PMChain(reorder_every=K): particles end in cell order with pid a permutation; the run
    is deterministic; the last deposit is the oracle's for the final slot order (bitwise);
    and the state, mapped back by pid, follows the unordered run to rounding level."""
    announce("chain with reordering")
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.solver import make_solver

    n, dx, dt = 32, 0.5, 0.02
    icr = plummer_ic(20_000, total_mass=1.0, scale_a=1.5, centre=(8.0, 8.0, 8.0), seed=9)

    def run(k):
        parts = ParticleState(icr.n)
        parts.load(icr.pos, icr.vel, icr.mass, icr.gid)
        chain = PMChain(parts, n, dx, make_solver("multigrid", n, dx=dx, grav_constant=1.0),
                        reorder_every=k)
        chain.prime()
        chain.run(4, dt)
        chain.run(3, dt)
        torch.cuda.synchronize()
        return parts, chain

    a, ca = run(2)
    b, _ = run(2)
    plain, _ = run(0)
    # a fixed-mode deposit gives no particle order: the chain then never reorders
    fx = ParticleState(icr.n)
    fx.load(icr.pos, icr.vel, icr.mass, icr.gid)
    cf = PMChain(fx, n, dx, make_solver("multigrid", n, dx=dx, grav_constant=1.0),
                 deposit_mode="fixed", reorder_every=1)
    cf.prime()
    cf.run(4, dt)
    torch.cuda.synchronize()
    assert np.array_equal(fx.pid.to_numpy(), np.arange(icr.n))
    pid = a.pid.to_numpy()
    assert not np.array_equal(pid, np.arange(icr.n)), "no reorder happened"
    assert np.array_equal(np.sort(pid), np.arange(icr.n))
    assert_bitwise(a.positions(), b.positions(), "deterministic positions")
    assert_bitwise(a.velocities(), b.velocities(), "deterministic velocities")
    rho_ref = O.deposit_cic(a.positions(), a.mass.to_numpy(), n, dx)[0]
    assert_bitwise(ca.grid.rho.to_numpy(), rho_ref, "rho vs oracle in the reordered slots")
    back = np.argsort(pid)
    assert_bitwise(a.mass.to_numpy()[back], icr.mass.astype(F32), "mass by pid")
    assert_bitwise(a.gid.to_numpy()[back], icr.gid.astype(np.int32), "gid by pid")
    np.testing.assert_allclose(a.positions()[back], plain.positions(), rtol=0, atol=1e-4)
    np.testing.assert_allclose(a.velocities()[back], plain.velocities(), rtol=0, atol=1e-3)
