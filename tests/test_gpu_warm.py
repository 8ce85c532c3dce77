"""Warm starts across time steps (pm_mg_solve_warm): extrapolation, the training pairs the
solver writes, and what a trained warm-start network gains over the method without it.

The density sequence is a Plummer sphere (4096 particles, 64^3 grid, box 20, G = 1) stepped
by the PM chain; every solve in every mode is checked against the exact float64 solution
of its discrete problem (DST-I), so a warm start can only change the work, not the answer.
No reference counterpart (the reference warm-starts from the previous phi only,
solver/multigrid.py:465-466).
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import pm_oracle as O
from conftest import announce

pytestmark = pytest.mark.gpu

N_GRID, BOX, DT = 64, 20.0, 0.05
F32 = np.float32


def _sequence(device, seed: int, steps: int):
    """Densities of `steps` consecutive PM steps of a Plummer sphere."""
    from galaxy_collision_amd.chain import PMChain
    from galaxy_collision_amd.data import ParticleState, as_tensor
    from galaxy_collision_amd.ic import plummer_ic
    from galaxy_collision_amd.solver import make_solver

    dx = BOX / N_GRID
    c = (N_GRID // 2) * dx
    icr = plummer_ic(4096, total_mass=1.0, scale_a=1.0, centre=(c + 0.3, c - 0.2, c), seed=seed)
    parts = ParticleState(icr.n)
    parts.load(icr.pos, icr.vel, icr.mass)
    chain = PMChain(parts, N_GRID, dx, make_solver("multigrid_mp", N_GRID, dx=dx, grav_constant=1.0))
    chain.prime()
    rhos = []
    for _ in range(steps):
        chain.step(DT)
        rhos.append(as_tensor(chain.grid.rho).clone())
    torch.cuda.synchronize()
    return rhos, dx


@pytest.fixture(scope="module")
def seq_a(device):
    return _sequence(device, 0, 36)


@pytest.fixture(scope="module")
def seq_b(device):
    return _sequence(device, 1, 24)


def _exact(phi_np, rho_np, dx):
    f = (F32(4.0 * math.pi) * rho_np).astype(F32)
    return O.dirichlet_solve_f64(f, phi_np, dx)


def _run(solver, rhos, device, check_every=0, dx=None):
    phi = torch.zeros((N_GRID,) * 3, device=device)
    cycles, warm, errs = [], [], []
    for t, rho in enumerate(rhos):
        solver.solve(rho, phi, warm_start=t > 0)
        cycles.append(solver.last_cycles)
        warm.append(solver.last_warm)
        if check_every and t % check_every == check_every - 1:
            p = phi.cpu().numpy()
            ex = _exact(p, rho.cpu().numpy(), dx)
            errs.append(float(np.linalg.norm(p - ex) / np.linalg.norm(ex)))
    return cycles, warm, errs, phi


def test_extrapolated_warm_start_same_accuracy_fewer_cycles(device, seq_a):
    announce("warm start: previous vs extrapolate")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver

    rhos, dx = seq_a
    prev = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="previous")
    extr = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="extrapolate")
    c_prev, _, e_prev, _ = _run(prev, rhos, device, check_every=6, dx=dx)
    c_extr, w_extr, e_extr, _ = _run(extr, rhos, device, check_every=6, dx=dx)
    assert max(e_prev) <= 1e-5 and max(e_extr) <= 1e-5, (e_prev, e_extr)
    assert all(w["extrapolated"] for w in w_extr[2:])
    assert not w_extr[1]["extrapolated"]  # no history before the second warm solve
    assert np.mean(c_extr[2:]) <= np.mean(c_prev[2:]), (c_prev, c_extr)


def test_pair_target_is_pooled_error_of_extrapolated_guess(device, seq_a):
    """pair_target * scale == pool4(phi_solved - (2 phi(t-1) - phi(t-2))) on the same faces,
    scale = RMS(pool4(f - A phi_ex)) * (4 dx)^2, all recomputed in float64 with torch."""
    announce("warm start: training pair")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver
    from galaxy_collision_amd.warmstart import PairSet

    rhos, dx = seq_a
    s = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="extrapolate", pool=4)
    pairs = PairSet(N_GRID // 4)
    phi = torch.zeros((N_GRID,) * 3, device=device)
    hist = []
    for t in range(4):
        x, tg = pairs.new_buffers()
        if t == 0:
            s.solve(rhos[t], phi)
        else:
            s.solve_warm(rhos[t], phi, 1, pair_input=x, pair_target=tg)
        hist.append(phi.clone())
    assert s.last_warm["extrapolated"]
    p2, p1, p0 = (h.double() for h in hist[-3:])  # phi(t-2), phi(t-1), phi(t)
    inner = (slice(1, -1),) * 3
    ex = p0.clone()
    ex[inner] = 2 * p1[inner] - p2[inner]
    f = 4 * math.pi * rhos[3].double()
    lap = torch.zeros_like(ex)
    lap[inner] = (ex[2:, 1:-1, 1:-1] + ex[:-2, 1:-1, 1:-1] + ex[1:-1, 2:, 1:-1] + ex[1:-1, :-2, 1:-1]
                  + ex[1:-1, 1:-1, 2:] + ex[1:-1, 1:-1, :-2] - 6 * ex[inner]) / (dx * dx)
    r = torch.zeros_like(ex)
    r[inner] = f[inner] - lap[inner]
    rp = F.avg_pool3d(r[None, None], 4)
    scale = float(rp.pow(2).mean().sqrt()) * (4 * dx) ** 2
    want = F.avg_pool3d((p0 - ex)[None, None], 4)[0, 0].reshape(-1) / scale
    got = tg.double()
    rel = float((got - want).norm() / want.norm())
    assert rel < 1e-3, rel
    # input channel 0 is the pooled residual over its RMS
    ch0 = x.view(-1, 8)[:, 0].double()
    want0 = (rp / rp.pow(2).mean().sqrt()).reshape(-1)
    assert float((ch0 - want0).norm() / want0.norm()) < 1e-2
    assert torch.count_nonzero(x.view(-1, 8)[:, 3:].float()) == 0


def test_learned_warm_start_gain_over_extrapolation(device, seq_a, seq_b):
    """Train on sequence A's pairs, evaluate on sequence B: the network must lower the
    residual of the extrapolated guess, the multigrid keeps its accuracy, and the V-cycle
    counts are reported for both methods."""
    announce("warm start: learned vs extrapolate")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver
    from galaxy_collision_amd.warmstart import PairSet, WarmStartNet, train

    rhos_a, dx = seq_a
    rhos_b, _ = seq_b
    res = N_GRID // 4
    gen = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="extrapolate", pool=4)
    pairs = PairSet(res)
    phi = torch.zeros((N_GRID,) * 3, device=device)
    for t, rho in enumerate(rhos_a):
        if t == 0:
            gen.solve(rho, phi)
            continue
        x, tg = pairs.new_buffers()
        w = gen.solve_warm(rho, phi, 1, pair_input=x, pair_target=tg)
        if w["extrapolated"]:
            pairs.add(x, tg)
    assert len(pairs) == len(rhos_a) - 2
    net = WarmStartNet(res, seed=0)
    hist = train(net, pairs, epochs=60, batch=4, lr=1e-3, seed=0)
    assert hist[-1]["loss"] < 0.5 * hist[0]["loss"], hist[:: 10]

    extr = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="extrapolate")
    lrn = MixedPrecisionMultigridSolver(N_GRID, dx=dx, grav_constant=1.0, warm="learned",
                                        net=net, pool=4)
    c_e, _, err_e, _ = _run(extr, rhos_b, device, check_every=4, dx=dx)
    c_l, w_l, err_l, _ = _run(lrn, rhos_b, device, check_every=4, dx=dx)
    assert max(err_l) <= 1e-5 and max(err_e) <= 1e-5, (err_l, err_e)
    judged = [w for w in w_l[2:] if w and w["extrapolated"]]
    ratio = [w["resid_after"] / w["resid_before"] for w in judged]
    used = np.mean([w["used_net"] for w in judged])
    print(f"\n[learned] cycles extrapolate {np.mean(c_e[2:]):.2f} learned {np.mean(c_l[2:]):.2f}; "
          f"residual ratio mean {np.mean(ratio):.3f}, kept {used:.0%}", flush=True)
    # the judge keeps the correction when it lowers the pooled residual; on this small
    # problem (34 training pairs of a 4096-particle sphere) the gain is modest, so the test
    # asserts that the network helps at all, keeps the accuracy and never costs cycles
    assert used >= 0.5 and np.median(ratio) < 0.99, (ratio, used)
    assert np.mean(c_l[2:]) <= np.mean(c_e[2:]) + 0.25


@pytest.mark.parametrize("warm", ["previous", "extrapolate", "learned"])
def test_graphs_and_split_fine_level_equal_eager_natural(device, seq_b, warm):
    """Every solve replays its V-cycles from HIP graphs (captured once per set of buffers),
    and MIXED solves run the finest level in colour-split layout with fused launches. Each
    combination must give bitwise the phi and the cycle counts of eager natural-layout
    cycles on every solve of a sequence (the norms only differ by summation order). The
    learned mode with an always-accepting judge swaps the defect buffers every solve, so
    its graphs are re-keyed and replayed alternately; the plain (float32) solver covers
    the other graph kind."""
    announce(f"multigrid graphs/split vs eager natural ({warm})")
    from galaxy_collision_amd.solver.multigrid import (MixedPrecisionMultigridSolver,
                                                       MultigridPoissonSolver)
    from galaxy_collision_amd.warmstart import WarmStartNet

    rhos, dx = seq_b
    kw = dict(dx=dx, grav_constant=1.0)
    if warm == "learned":
        net = WarmStartNet(N_GRID // 4, seed=3)
        kw.update(warm="learned", net=net, pool=4, fallback_ratio=1e30)
    elif warm == "extrapolate":
        kw.update(warm="extrapolate")
    combos = [(False, False), (True, False), (False, True), (True, True)]
    groups = [[MixedPrecisionMultigridSolver(N_GRID, graphs=g, split_fine=sp, **kw)
               for g, sp in combos]]
    if warm == "previous":
        groups.append([MultigridPoissonSolver(N_GRID, dx=dx, grav_constant=1.0, graphs=g)
                       for g in (False, True)])
    used = 0
    for t, rho in enumerate(rhos[:10]):
        for group in groups:
            phis = []
            for s in group:
                phi = getattr(s, "_test_phi", None)
                if phi is None:
                    phi = s._test_phi = torch.zeros((N_GRID,) * 3, device=device)
                s.solve(rho, phi, warm_start=t > 0)
                torch.cuda.synchronize()
                phis.append(phi)
            ref = group[0]
            for s, phi in zip(group[1:], phis[1:]):
                assert torch.equal(phi, phis[0]), (warm, t)
                assert s.last_cycles == ref.last_cycles, (warm, t)
                for key, v in ref.last_stats.items():
                    assert s.last_stats[key] == pytest.approx(v, rel=1e-12, abs=0), (warm, t, key)
        if warm == "learned" and groups[0][-1].last_warm:
            used += groups[0][-1].last_warm["used_net"]
    if warm == "learned":
        assert used >= 5, used


@pytest.mark.parametrize("n", [33, 37, 48])
def test_split_fine_level_odd_and_even_grids(device, n):
    """Colour-split finest level on odd grids (pad slot in half of the rows) and a grid whose
    hierarchy reaches the bottom solver at its second level: bitwise the natural-layout
    solve, cold and warm."""
    announce(f"multigrid split fine level n={n}")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver

    rng = np.random.default_rng(n)
    rho = torch.zeros((n,) * 3, device=device)
    inner = rng.uniform(0.0, 1.0, size=(n - 8,) * 3).astype(F32)
    rho[4:-4, 4:-4, 4:-4] = torch.from_numpy(inner).to(device)
    out = []
    for split in (False, True):
        s = MixedPrecisionMultigridSolver(n, dx=0.5, grav_constant=1.0, split_fine=split)
        phi = torch.zeros((n,) * 3, device=device)
        s.solve(rho, phi)
        c0 = s.last_cycles
        s.solve(rho * 1.01, phi, warm_start=True)
        out.append((phi.clone(), c0, s.last_cycles))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


def test_carried_guess_equals_extrapolation_pass(device, seq_b):
    """The extrapolated warm start without its own pass: each MIXED solve's last update
    writes the next guess 2*phi_t - phi_(t-1), and the next solve starts from it. Over a
    sequence of one-cycle solves phi is bitwise that of the explicit extrapolation pass
    (run here by the learned mode with a judge that rejects every correction), in both
    fine-level layouts."""
    announce("carried guess vs extrapolation pass")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver
    from galaxy_collision_amd.warmstart import WarmStartNet

    rhos, dx = seq_b
    net = WarmStartNet(N_GRID // 4, seed=3)
    for split in (True, False):
        kw = dict(dx=dx, grav_constant=1.0, n_cycles=1, split_fine=split)
        carried = MixedPrecisionMultigridSolver(N_GRID, warm="extrapolate", **kw)
        explicit = MixedPrecisionMultigridSolver(N_GRID, warm="learned", net=net, pool=4,
                                                 fallback_ratio=1e-30, **kw)
        pa = torch.zeros((N_GRID,) * 3, device=device)
        pb = torch.zeros((N_GRID,) * 3, device=device)
        for t, rho in enumerate(rhos[:8]):
            carried.solve(rho, pa, warm_start=t > 0)
            explicit.solve(rho, pb, warm_start=t > 0)
            torch.cuda.synchronize()
            assert carried.last_cycles == explicit.last_cycles == 1
            if t >= 2:
                assert carried.last_warm["extrapolated"] and explicit.last_warm["extrapolated"]
                assert not explicit.last_warm["used_net"]
            assert torch.equal(pa, pb), (split, t)
