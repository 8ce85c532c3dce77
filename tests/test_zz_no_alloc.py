"""pm_deposit_cic_noalloc: the deposit whose scratch is a static device array.

Called through the C ABI, it must give bit for bit what pm_deposit_cic (with a plan) gives,
on consecutive calls that reuse the array (a larger input, then a smaller one on a smaller
grid, then a 256^3 grid), refuse inputs that do not fit with PM_ERR_CAPACITY before any
device work, and still work after such a refusal.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import pm_oracle as O
from conftest import announce, assert_bitwise

pytestmark = pytest.mark.gpu
F32 = np.float32


def _inputs(npart: int, n: int, dx: float, seed: int, dense: int = 0):
    rng = np.random.default_rng(seed)
    pos = rng.normal(n * dx / 2, n * dx / 6, size=(npart, 3)).astype(F32)
    if dense:
        pos[:dense] = rng.uniform(0.4 * n * dx, 0.4 * n * dx + 0.9 * dx, size=(dense, 3))
    pos[: npart // 50] = rng.uniform(-2 * dx, (n + 1) * dx, size=(npart // 50, 3))  # edges
    mass = rng.uniform(0.5, 2.0, npart).astype(F32)
    return pos, mass


def _device(pos, mass, device):
    cols = [torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(device) for a in range(3)]
    return cols, torch.from_numpy(mass).to(device)


def _noalloc(n, cols, m, dx, device):
    from galaxy_collision_amd import _native as nat

    npart = m.numel()
    rho = torch.full((n, n, n), -7.0, device=device)  # stale values must be overwritten
    keys = torch.empty(max(npart, 1), dtype=torch.int32, device=device)
    perm = torch.empty(max(npart, 1), dtype=torch.int32, device=device)
    nat.call("pm_deposit_cic_noalloc", n, *(nat.ptr(c) for c in cols), nat.ptr(m), npart,
             float(dx), nat.ptr(rho), nat.ptr(keys, torch.int32), nat.ptr(perm, torch.int32),
             nat.stream())
    torch.cuda.synchronize()
    return rho.cpu().numpy(), keys[:npart].cpu().numpy(), perm[:npart].cpu().numpy()


def _with_plan(n, cols, m, dx, device):
    from galaxy_collision_amd.deposit import DepositPlan

    npart = m.numel()
    plan = DepositPlan(n, max(npart, 1))
    rho = torch.empty((n, n, n), device=device)
    keys = torch.empty(max(npart, 1), dtype=torch.int32, device=device)
    perm = torch.empty(max(npart, 1), dtype=torch.int32, device=device)
    plan.deposit(*cols, m, npart, dx, rho, keys, perm)
    torch.cuda.synchronize()
    out = rho.cpu().numpy(), keys[:npart].cpu().numpy(), perm[:npart].cpu().numpy()
    plan.close()
    return out


def test_noalloc_equals_plan_deposit_on_consecutive_calls(device):
    cases = [(1_200_000, 128, 0.5, 1, 30_000), (300_000, 64, 0.3125, 2, 5_000),
             (589_312, 256, 0.15625, 3, 20_000), (1_799_896, 64, 1.0, 4, 0)]
    for npart, n, dx, seed, dense in cases:
        announce(f"noalloc deposit npart={npart} n={n} seed={seed}")
        pos, mass = _inputs(npart, n, dx, seed, dense)
        cols, m = _device(pos, mass, device)
        got = _noalloc(n, cols, m, dx, device)
        exp = _with_plan(n, cols, m, dx, device)
        for g, e, what in zip(got, exp, ("rho", "keys", "perm")):
            assert_bitwise(g, e, f"{what} (npart={npart}, n={n})")
        if npart <= 600_000:
            rho_ref, keys_ref, perm_ref = O.deposit_cic(pos, mass, n, dx)
            assert_bitwise(got[0], rho_ref, "rho vs oracle")
            assert_bitwise(got[1].view(np.uint32), keys_ref, "keys vs oracle")
            assert_bitwise(got[2], perm_ref, "perm vs oracle")


def test_noalloc_refuses_oversize_then_works(device):
    from galaxy_collision_amd import _native as nat

    announce("noalloc deposit capacity refusal")
    with pytest.raises(nat.PMChainError, match="static scratch holds") as err:
        nat.call("pm_deposit_cic_noalloc", 256, 1, 1, 1, 1, 589_313, 1.0, 1, None, None, None)
    assert err.value.status == nat.PM_ERR_CAPACITY
    pos, mass = _inputs(50_000, 32, 0.5, 5)
    cols, m = _device(pos, mass, device)
    got = _noalloc(32, cols, m, 0.5, device)
    rho_ref, _, _ = O.deposit_cic(pos, mass, 32, 0.5)
    assert_bitwise(got[0], rho_ref, "rho after a refused call")
    # empty input: rho is zeroed
    empty = [torch.empty(0, device=device) for _ in range(3)]
    rho, _, _ = _noalloc(32, empty, torch.empty(0, device=device), 0.5, device)
    assert not rho.any()
