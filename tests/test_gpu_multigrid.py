"""GPU parity of the open-boundary multigrid (reference solver/multigrid.py).

- single kernels (RB pass, residual, restriction, prolongation, faces) bit for bit against
  the numpy restatement;
- moments against the numpy closed form (float64, rtol 1e-12);
- the full PLAIN solve against the oracle's solve (same cycle count; phi within 1e-6,
  bitwise when the float64 moment sums round to the same float32 face parameters);
- MIXED mode against the exact float64 DST-I solve of the same discrete problem
  (algebraic phi rel-L2 <= 1e-5);
- the reference test suite's convergence / far-field / warm-start / oracle checks.
"""

from __future__ import annotations

import math

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


def _gaussian_rho(n, dx, sigma, total_mass, centre=None):
    axis = np.arange(n) * dx
    c = np.full(3, (n - 1) / 2.0 * dx) if centre is None else np.asarray(centre)
    xx, yy, zz = np.meshgrid(axis - c[0], axis - c[1], axis - c[2], indexing="ij")
    g = np.exp(-(xx ** 2 + yy ** 2 + zz ** 2) / (2.0 * sigma ** 2))
    return (g / g.sum() * total_mass / dx ** 3).astype(F32)


@pytest.mark.parametrize("n", [3, 8, 33, 64])
def test_rb_pass_and_residual_bitwise(device, n):
    announce(f"rb/residual n={n}")
    from galaxy_collision_amd import _native as nat

    rng = np.random.default_rng(n)
    phi = rng.normal(size=(n, n, n)).astype(F32)
    rhs = rng.normal(size=(n, n, n)).astype(F32)
    tp, tr = _t(phi, device), _t(rhs, device)
    for color in (0, 1, 1, 0):
        nat.call("pm_rb_pass", nat.ptr(tp), nat.ptr(tr), n, 0.25, color, nat.stream())
        O.rb_pass(phi, rhs, 0.25, color)
        assert_bitwise(_np(tp), phi)
    res = torch.empty_like(tp)
    nat.call("pm_residual", nat.ptr(tp), nat.ptr(tr), nat.ptr(res), n, 4.0, nat.stream())
    assert_bitwise(_np(res), O.residual(phi, rhs, 4.0))


@pytest.mark.parametrize("n", [3, 5, 37, 70, 129, 256])
def test_fused_sweep_equals_two_passes(device, n):
    """pm_rb_sweep (streaming red+black in one pass) == pm_rb_pass(0) then pm_rb_pass(1)."""
    announce(f"fused sweep n={n}")
    from galaxy_collision_amd import _native as nat

    rng = np.random.default_rng(100 + n)
    phi = rng.normal(size=(n, n, n)).astype(F32)
    rhs = rng.normal(size=(n, n, n)).astype(F32)
    a, b, tr = _t(phi, device), _t(phi, device), _t(rhs, device)
    tmp = torch.full_like(a, 9.0)
    for _ in range(2):
        nat.call("pm_rb_sweep", nat.ptr(a), nat.ptr(tmp), nat.ptr(tr), n, 0.3, nat.stream())
        a, tmp = tmp, a
        nat.call("pm_rb_pass", nat.ptr(b), nat.ptr(tr), n, 0.3, 0, nat.stream())
        nat.call("pm_rb_pass", nat.ptr(b), nat.ptr(tr), n, 0.3, 1, nat.stream())
    assert_bitwise(_np(a), _np(b), "sweep vs passes")
    if n <= 70:
        for _ in range(2):
            O.rb_pass(phi, rhs, 0.3, 0)
            O.rb_pass(phi, rhs, 0.3, 1)
        assert_bitwise(_np(a), phi, "sweep vs oracle")


@pytest.mark.parametrize("nf,nc", [(6, 3), (16, 8), (64, 32), (34, 18), (33, 17), (65, 33),
                                   (129, 65)])
def test_restrict_prolong_bitwise(device, nf, nc):
    """Halving (PLAIN) and aligned (MIXED, nf = 2nc - 1) hierarchies; workgroups inside one
    i-plane (staged coarse rows) and across planes."""
    announce(f"restrict/prolong nf={nf} nc={nc}")
    from galaxy_collision_amd import _native as nat

    rng = np.random.default_rng(nf)
    if nf == 2 * nc:
        res = rng.normal(size=(nf, nf, nf)).astype(F32)
        out = torch.full((nc, nc, nc), 3.0, device=device)
        nat.call("pm_restrict", nat.ptr(_t(res, device)), nat.ptr(out), nf, nc, nat.stream())
        assert_bitwise(_np(out), O.restrict(res))
    coarse = rng.normal(size=(nc, nc, nc)).astype(F32)
    fine = rng.normal(size=(nf, nf, nf)).astype(F32)
    tf = _t(fine, device)
    nat.call("pm_prolong_add", nat.ptr(_t(coarse, device)), nat.ptr(tf), nf, nc, nat.stream())
    O.prolong_add(coarse, fine)
    assert_bitwise(_np(tf), fine)


def test_faces_bitwise(device):
    announce("faces")
    import ctypes

    from galaxy_collision_amd import _native as nat

    n, dx = 20, 0.3125
    p = np.array([1.0, 0.98, 3.1, 2.9, 3.2, 0.5, 0.4, 0.3, 0.02, -0.01, 0.015], F32)
    phi = np.random.default_rng(1).normal(size=(n, n, n)).astype(F32)
    tp = _t(phi, device)
    nat.call("pm_set_faces", nat.ptr(tp), n, dx, (ctypes.c_float * 11)(*p), nat.stream())
    O.set_faces(phi, dx, p)
    assert_bitwise(_np(tp), phi)


def test_moments_match_numpy(device):
    announce("moments")
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    n, dx = 64, 1.0
    axis = np.arange(n) * dx
    xx, yy, zz = np.meshgrid(axis, axis, axis, indexing="ij")
    blob = np.exp(-((xx - 20) ** 2 + (yy - 34) ** 2 + (zz - 28) ** 2) / (2 * 6.0 ** 2))
    blob += 0.5 * np.exp(-((xx - 44) ** 2 + (yy - 30) ** 2 + (zz - 36) ** 2) / (2 * 4.0 ** 2))
    rho = (blob / blob.sum() * 3.2e11 / dx ** 3).astype(F32)
    solver = MultigridPoissonSolver(n, dx=dx)
    m, cm, s = solver._moments_device(_t(rho, device))
    ref = O.moments(rho, dx)
    tot = ref[0]
    assert m == pytest.approx(tot, rel=1e-12)
    assert np.allclose(cm, ref[1:4] / tot, rtol=1e-12)
    cx, cy, cz = ref[1:4] / tot
    s_ref = [ref[4] - tot * cx * cx, ref[5] - tot * cy * cy, ref[6] - tot * cz * cz,
             ref[7] - tot * cx * cy, ref[8] - tot * cx * cz, ref[9] - tot * cy * cz]
    assert np.allclose(s, s_ref, rtol=1e-8)


@pytest.mark.parametrize("n,dx,centre,warm", [
    (32, 1.0, None, False), (48, 0.5, (10.0, 13.0, 12.0), False), (64, 0.3125, None, True)])
def test_plain_solve_matches_oracle(device, n, dx, centre, warm):
    announce(f"plain solve n={n} warm={warm}")
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    G = 1.0
    rho = _gaussian_rho(n, dx, 3.0 * dx, 1.0, centre)
    phi0 = None
    if warm:
        phi0, _, _, _ = O.mg_solve(rho, dx, G, n_cycles=2, tol=0.0)
    ref, cycles, _, _ = O.mg_solve(rho, dx, G, phi0=phi0)
    solver = MultigridPoissonSolver(n, dx=dx, grav_constant=G)
    tphi = _t(phi0 if warm else np.zeros((n, n, n)), device)
    solver.solve(_t(rho, device), tphi, warm_start=warm)
    got = _np(tphi)
    assert solver.last_cycles == cycles
    assert np.abs(got - ref).max() <= 1e-6 * np.abs(ref).max()


@pytest.mark.parametrize("n,dx", [(32, 1.0), (64, 0.3125), (96, 0.5)])
def test_mixed_solve_reaches_1e5_of_exact_discrete_solution(device, n, dx):
    announce(f"mixed solve n={n}")
    from galaxy_collision_amd.solver.multigrid import MixedPrecisionMultigridSolver

    G = 1.0
    rho = _gaussian_rho(n, dx, 2.5 * dx, 1.0, centre=(0.4 * n * dx, 0.55 * n * dx, 0.5 * n * dx))
    rho += _gaussian_rho(n, dx, 1.5 * dx, 0.3, centre=(0.6 * n * dx, 0.45 * n * dx, 0.5 * n * dx))
    solver = MixedPrecisionMultigridSolver(n, dx=dx, grav_constant=G)  # tol: est. error 1e-6
    tphi = torch.zeros((n, n, n), device=device)
    solver.solve(_t(rho, device), tphi)
    phi = _np(tphi)
    f = (F32(4.0 * math.pi * G) * rho).astype(F32)
    exact = O.dirichlet_solve_f64(f, phi, dx)  # faces taken from the solve (they are exact)
    rel = np.linalg.norm(phi - exact) / np.linalg.norm(exact)
    stats = solver.last_stats
    assert rel <= 1e-5, f"algebraic phi rel-L2 {rel:.3e} ({stats})"
    assert stats["err_est"] <= 1e-6 and stats["cycles"] < 40, stats
    # the error estimate is not optimistic by more than a small factor
    assert rel <= 5 * max(stats["err_est"], 1e-7), (rel, stats)
    # the faces equal the oracle's float32 multipole values
    ref_faces = np.zeros_like(phi)
    O.set_faces(ref_faces, dx, O.face_params(O.moments(rho, dx), G))
    assert_bitwise(phi[0], ref_faces[0], 'face x=0')
    assert_bitwise(phi[:, :, -1], ref_faces[:, :, -1], 'face z=n-1')


def test_vcycle_reduces_residual(device):
    announce("vcycle reduces residual")
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    n = 32
    rho = _gaussian_rho(n, 1.0, 3.0, 1.0e10)
    solver = MultigridPoissonSolver(n, dx=1.0, n_cycles=1)
    phi = torch.zeros((n, n, n), device=device)
    solver.solve(_t(rho, device), phi)
    r1 = solver.residual_norm()
    for _ in range(6):
        solver._vcycle(0)
    assert solver.residual_norm() < 0.1 * r1


def test_point_mass_far_field(device):
    announce("point mass far field 96^3")
    from galaxy_collision_amd import units
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    n, mass, G = 96, 1.0e10, units.G
    c = n // 2
    rho = np.zeros((n, n, n), F32)
    rho[c, c, c] = mass
    solver = MultigridPoissonSolver(n, dx=1.0, grav_constant=G, n_cycles=80, tol=0.0)
    tphi = torch.zeros((n, n, n), device=device)
    solver.solve(_t(rho, device), tphi)
    rhs_norm = 4 * np.pi * G * float(np.sqrt((rho.astype(np.float64) ** 2).sum()))
    assert solver.residual_norm() / rhs_norm < 1e-4
    phi = _np(tphi)
    prev = np.inf
    for r in (4, 6, 8, 12, 16, 20):
        analytic = -G * mass / r
        vals = [phi[c + r, c, c], phi[c - r, c, c], phi[c, c + r, c], phi[c, c - r, c],
                phi[c, c, c + r], phi[c, c, c - r]]
        rel = abs(float(np.mean(vals)) - analytic) / abs(analytic)
        assert rel < 0.4 / r ** 2
        assert rel < prev * (1.0 + 1e-4) + 1e-7
        prev = rel


def test_adaptive_warm_start(device):
    announce("adaptive warm start")
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    n = 64
    rho = _t(_gaussian_rho(n, 1.0, 4.0, 1.0e10), device)
    pa, pf = torch.zeros((n, n, n), device=device), torch.zeros((n, n, n), device=device)
    adaptive = MultigridPoissonSolver(n, n_cycles=30, tol=1e-3)
    fixed = MultigridPoissonSolver(n, n_cycles=30, tol=0.0)
    adaptive.solve(rho, pa)
    fixed.solve(rho, pf)
    adaptive.solve(rho, pa, warm_start=True)
    fixed.solve(rho, pf, warm_start=True)
    assert adaptive.last_cycles <= 3 and fixed.last_cycles == 30
    a, f = _np(pa), _np(pf)
    assert np.sqrt(((a - f) ** 2).sum() / (f ** 2).sum()) < 1e-2


def test_256_cold_residual_regression(device):
    """Reference tests/test_solver_multigrid.py:73-95: cold 256^3, 20 cycles, ratio < 5e-2."""
    announce("256^3 cold regression")
    from galaxy_collision_amd import units
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    n = 256
    rho = _gaussian_rho(n, 1.0, 8.0, 2.32e11)
    solver = MultigridPoissonSolver(n, dx=1.0, n_cycles=20)
    phi = torch.zeros((n, n, n), device=device)
    solver.solve(_t(rho, device), phi)
    rhs_norm = 4 * np.pi * units.G * float(np.sqrt((rho.astype(np.float64) ** 2).sum()))
    assert solver.residual_norm() / rhs_norm < 5e-2


def test_mg_matches_device_fft_solver(device):
    announce("mg vs fft 48^3")
    from galaxy_collision_amd.solver import make_solver

    n, G = 48, 1.0
    rho = _t(_gaussian_rho(n, 1.0, 4.0, 2.0e3), device)
    pm, pf = torch.zeros((n, n, n), device=device), torch.zeros((n, n, n), device=device)
    make_solver("multigrid", n, dx=1.0, grav_constant=G, n_cycles=40).solve(rho, pm)
    make_solver("fft", n, dx=1.0, grav_constant=G).solve(rho, pf)
    a, b = _np(pm), _np(pf)
    ref = O.fft_solve_f64(_np(rho), 1.0, G)
    assert np.abs(b - ref).max() <= 1e-6 * np.abs(ref).max()
    core = (slice(6, n - 6),) * 3
    peak = np.abs(b).max()
    assert np.abs(a[core] - b[core]).max() / peak < 0.03
    assert np.abs(a[core] - b[core]).mean() / peak < 0.005


def test_solver_factory():
    from galaxy_collision_amd.solver import make_solver
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    assert isinstance(make_solver("multigrid", 16), MultigridPoissonSolver)
    with pytest.raises(ValueError, match="unknown solver"):
        make_solver("nope", 16)


@pytest.mark.parametrize("n,warm", [(128, False), (256, True)])
def test_plain_solve_large_matches_c_port(device, n, warm):
    """Large grids use the fused sweep and the bottom solver: the full PLAIN solve still
    reproduces the oracle's C port of the reference solve (same cycles, phi to 1e-6)."""
    announce(f"plain solve vs C port n={n}")
    from galaxy_collision_amd.solver.multigrid import MultigridPoissonSolver

    G, dx = 1.0, 40.0 / n
    rho = _gaussian_rho(n, dx, 6.0 * dx, 1.0, centre=(0.45 * n * dx, 0.52 * n * dx, 0.5 * n * dx))
    L = O.lib()
    h = L.cpu_mg_create(n, dx, G, 20, 2, 2, 40, 1, 1e-3)
    phi_c = np.zeros((n, n, n), F32)
    if warm:
        L.cpu_mg_solve(h, O._fp(np.ascontiguousarray(rho)), O._fp(phi_c), 0)
    phi0 = phi_c.copy()
    cycles = L.cpu_mg_solve(h, O._fp(np.ascontiguousarray(rho)), O._fp(phi_c), 1 if warm else 0)
    L.cpu_mg_destroy(h)
    solver = MultigridPoissonSolver(n, dx=dx, grav_constant=G)
    tphi = _t(phi0, device)
    solver.solve(_t(rho, device), tphi, warm_start=warm)
    assert solver.last_cycles == cycles
    got = _np(tphi)
    assert np.abs(got - phi_c).max() <= 1e-6 * np.abs(phi_c).max()


@pytest.mark.parametrize("mode,n", [("mixed", 17), ("mixed", 33), ("mixed", 37), ("mixed", 48),
                                    ("mixed", 129), ("plain", 16), ("plain", 64)])
def test_bottom_solver_equals_per_level_kernels(device, mode, n):
    """The one-workgroup bottom solver (every level <= 17 in LDS: batched loads, the
    residual held in LDS for the restriction, one lane per coarse cell in the
    prolongation) gives bitwise the phi and the cycle counts of the same levels run by the
    per-level kernels, cold and warm, on odd (17 -> 9), even (10 -> 6, a restriction that
    reaches the far face) and halving (PLAIN 16 -> 8) hierarchies, with the bottom at the
    finest level (17, 16) or below it."""
    announce(f"bottom solver vs per-level kernels {mode} n={n}")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.solver.multigrid import (MixedPrecisionMultigridSolver,
                                                       MultigridPoissonSolver)

    dx = 40.0 / n
    rho = _t(_gaussian_rho(n, dx, 3.0 * dx, 1.0, centre=(0.47 * n * dx, 0.5 * n * dx,
                                                          0.53 * n * dx)), device)
    out = []
    for bottom in (1, 0):
        cls = MixedPrecisionMultigridSolver if mode == "mixed" else MultigridPoissonSolver
        s = cls(n, dx=dx, grav_constant=1.0)
        nat.call("pm_mg_set_option", s._h, nat.PM_MG_OPT_BOTTOM, bottom)
        phi = torch.zeros((n,) * 3, device=device)
        s.solve(rho, phi)
        c0 = s.last_cycles
        s.solve(rho * 1.01, phi, warm_start=True)
        out.append((phi.clone(), c0, s.last_cycles))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]


@pytest.mark.parametrize("mode,n,sweeps,bottom", [("mixed", 129, (1, 1), 1), ("mixed", 129, (2, 2), 1),
                                                  ("mixed", 65, (1, 1), 0), ("plain", 128, (2, 2), 1),
                                                  ("plain", 64, (1, 1), 0), ("mixed", 129, (1, 2), 1)])
def test_small_level_fused_sweep_equals_colour_passes(device, mode, n, sweeps, bottom):
    """One launch per red-black sweep on the levels <= 65 below the finest (out of place,
    colour-1 nodes recomputing their colour-0 neighbours) gives bitwise the phi and cycle
    counts of two colour passes per sweep, cold and warm, with the bottom solver on or off;
    an odd number of sweeps per cycle keeps the colour passes."""
    announce(f"small-level fused sweep {mode} n={n} sweeps={sweeps} bottom={bottom}")
    from galaxy_collision_amd import _native as nat
    from galaxy_collision_amd.solver.multigrid import (MixedPrecisionMultigridSolver,
                                                       MultigridPoissonSolver)

    dx = 40.0 / n
    rho = _t(_gaussian_rho(n, dx, 3.0 * dx, 1.0, centre=(0.46 * n * dx, 0.52 * n * dx,
                                                          0.5 * n * dx)), device)
    out = []
    for small in (1, 0):
        cls = MixedPrecisionMultigridSolver if mode == "mixed" else MultigridPoissonSolver
        s = cls(n, dx=dx, grav_constant=1.0, pre_sweeps=sweeps[0], post_sweeps=sweeps[1])
        nat.call("pm_mg_set_option", s._h, nat.PM_MG_OPT_SMALL_SWEEP, small)
        nat.call("pm_mg_set_option", s._h, nat.PM_MG_OPT_BOTTOM, bottom)
        phi = torch.zeros((n,) * 3, device=device)
        s.solve(rho, phi)
        c0 = s.last_cycles
        s.solve(rho * 1.01, phi, warm_start=True)
        out.append((phi.clone(), c0, s.last_cycles))
    assert torch.equal(out[0][0], out[1][0])
    assert out[0][1:] == out[1][1:]
