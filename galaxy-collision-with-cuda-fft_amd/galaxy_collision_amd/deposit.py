"""The particle<->grid stages of the PM force chain on libpmchain (reference deposit.py).

- ``deposit_density(parts, rho, dx, mode="sorted")``: CIC deposit of mass density
  (M/dx^3), deterministic in every mode: "sorted" (cell-key sort + segmented reduction in a
  defined float32 order, also yields the keys and particle order; the default here),
  "fixed" (64-bit fixed-point integer sums with device-scope atomics, order independent) or
  "bricked" (the "fixed" result bit for bit, summed per 8^3-cell brick in LDS; the mode the
  benchmark workloads give ``PMChain``, which reuses one binning over several steps).
  Reference deposit.py:177-182.
- ``potential_to_accel(phi, ax, ay, az, dx)``: g = -grad(phi), central inside, one-sided on
  the faces. Reference deposit.py:204-207.
- ``gather_acceleration(parts, ax, ay, az, acc_x, acc_y, acc_z, dx)``: CIC interpolation of
  g to the particles. Reference deposit.py:185-201.

Node (i,j,k) sits at (i,j,k)*dx; out-of-box CIC corners are dropped (open boundaries).
All arguments are device fields/tensors; work is enqueued on torch's current stream.
"""

from __future__ import annotations

import ctypes
import threading

import torch

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.data import as_tensor


_MODES = {"sorted": nat.PM_DEPOSIT_SORTED, "fixed": nat.PM_DEPOSIT_FIXED,
          "bricked": nat.PM_DEPOSIT_BRICKED}


class DepositPlan:
    """Sort and reduction scratch for deposits on one grid size and device."""

    def __init__(self, grid_size: int, max_particles: int):
        nat.require_device()
        self.grid_size = int(grid_size)
        self.capacity = int(max_particles)
        self.device = torch.cuda.current_device()
        h = ctypes.c_void_p()
        nat.call("pm_deposit_plan_create", self.grid_size, self.capacity, ctypes.byref(h))
        self.handle = h
        self.mode = "sorted"

    def set_mode(self, mode: str) -> None:
        if mode not in _MODES:
            raise ValueError(f"deposit mode must be one of {sorted(_MODES)}, got {mode!r}")
        if mode != self.mode:
            nat.call("pm_deposit_plan_set_mode", self.handle, _MODES[mode])
            self.mode = mode

    def deposit(self, pos_x, pos_y, pos_z, mass, n: int, dx: float, rho,
                keys_out: torch.Tensor | None = None, perm_out: torch.Tensor | None = None):
        nat.call("pm_deposit_cic", self.handle, nat.ptr(as_tensor(pos_x)),
                 nat.ptr(as_tensor(pos_y)), nat.ptr(as_tensor(pos_z)), nat.ptr(as_tensor(mass)),
                 int(n), float(dx), nat.ptr(as_tensor(rho)),
                 nat.ptr(keys_out, torch.int32) if keys_out is not None else None,
                 nat.ptr(perm_out, torch.int32) if perm_out is not None else None,
                 nat.stream())

    def deposit_keyed(self, pos_x, pos_y, pos_z, mass, n: int, dx: float, rho,
                      keys: torch.Tensor, perm_out: torch.Tensor | None = None):
        """``deposit`` for particles whose cell keys ``keys`` (as ``keys_out`` would hold
        them) are already known, e.g. written by the fused drift pass: no key pass."""
        nat.call("pm_deposit_cic_keyed", self.handle, nat.ptr(as_tensor(pos_x)),
                 nat.ptr(as_tensor(pos_y)), nat.ptr(as_tensor(pos_z)), nat.ptr(as_tensor(mass)),
                 int(n), float(dx), nat.ptr(as_tensor(rho)), nat.ptr(keys, torch.int32),
                 nat.ptr(perm_out, torch.int32) if perm_out is not None else None,
                 nat.stream())

    def deposit_step(self, pos_x, pos_y, pos_z, mass, n: int, dx: float, rho,
                     bin: bool = False, rho_kept: bool = False) -> None:
        """"bricked" mode: the deposit of one step of a run, reusing the last binning
        unless ``bin`` (same result as ``deposit``, bit for bit). ``rho_kept``: ``rho`` is
        the array of this plan's last deposit, unchanged since (its far-away zeros are kept)."""
        flags = (nat.PM_STEP_BIN if bin else 0) | (nat.PM_STEP_RHO_KEPT if rho_kept else 0)
        nat.call("pm_deposit_cic_step", self.handle, nat.ptr(as_tensor(pos_x)),
                 nat.ptr(as_tensor(pos_y)), nat.ptr(as_tensor(pos_z)), nat.ptr(as_tensor(mass)),
                 int(n), float(dx), nat.ptr(as_tensor(rho)), flags, nat.stream())

    def mark_ordered(self) -> None:
        """"bricked" mode: the particles were moved into ``order()`` of the last binning."""
        nat.call("pm_deposit_plan_mark_ordered", self.handle)

    def brick_list(self) -> tuple[int, int, int] | None:
        """(list address, device count address, max count) of the last deposit's marked node
        bricks, compacted; None after a fixed-mode deposit."""
        lst, cnt, mx = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int64()
        nat.call("pm_deposit_plan_brick_list", self.handle, ctypes.byref(lst), ctypes.byref(cnt),
                 ctypes.byref(mx))
        return (lst.value, cnt.value, mx.value) if lst.value else None

    def order(self) -> int | None:
        """Device address of the last deposit's particle order (sorted position -> particle
        index), or None after a fixed-mode or empty deposit; valid until the next deposit."""
        ptr = ctypes.c_void_p()
        nat.call("pm_deposit_plan_order", self.handle, ctypes.byref(ptr))
        return ptr.value

    def node_bricks(self) -> int | None:
        """Device address of the last deposit's node-brick marks (8^3 node bricks holding a
        CIC corner of some deposited particle), or None after a fixed-mode deposit."""
        ptr, edge = ctypes.c_void_p(), ctypes.c_int32()
        nat.call("pm_deposit_plan_node_bricks", self.handle, ctypes.byref(ptr),
                 ctypes.byref(edge))
        return ptr.value

    def set_timing(self, on: bool) -> None:
        nat.call("pm_deposit_plan_set_timing", self.handle, 1 if on else 0)

    def last_timing(self) -> dict:
        out = (ctypes.c_float * 4)()
        nat.call("pm_deposit_last_timing", self.handle, out)
        return dict(zip(("key", "sort", "reduce", "assemble"), list(out)))

    def close(self) -> None:
        if getattr(self, "handle", None):
            nat.call("pm_deposit_plan_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_plans: dict[tuple[int, int], DepositPlan] = {}
_plans_lock = threading.Lock()


def deposit_plan(grid_size: int, n_particles: int) -> DepositPlan:
    """Cached plan for (grid_size, current device), grown when more particles arrive."""
    key = (int(grid_size), torch.cuda.current_device())
    with _plans_lock:
        plan = _plans.get(key)
        if plan is None or plan.capacity < n_particles:
            if plan is not None:
                torch.cuda.current_stream().synchronize()
                plan.close()
            plan = DepositPlan(grid_size, max(int(n_particles), 1))
            _plans[key] = plan
        return plan


def _grid_n(field) -> int:
    shape = tuple(as_tensor(field).shape)
    if len(shape) != 3 or len(set(shape)) != 1:
        raise ValueError(f"expected a cubic (n, n, n) grid, got {shape}")
    return shape[0]


def deposit_density(parts, rho, dx: float = 1.0, mode: str = "sorted") -> None:
    """Deposit particle masses onto ``rho`` (mass per dx^3) by CIC. Overwrites ``rho``."""
    n = _grid_n(rho)
    plan = deposit_plan(n, parts.n)
    plan.set_mode(mode)
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho)


def deposit_keys(parts, rho, dx: float = 1.0):
    """Deposit and also return (cell keys in particle order, stable key-order permutation)."""
    n = _grid_n(rho)
    t = as_tensor(rho)
    keys = torch.empty(parts.n, dtype=torch.int32, device=t.device)
    perm = torch.empty(parts.n, dtype=torch.int32, device=t.device)
    plan = deposit_plan(n, parts.n)
    plan.set_mode("sorted")
    plan.deposit(parts.pos_x, parts.pos_y, parts.pos_z, parts.mass, parts.n, dx, rho, keys, perm)
    return keys, perm


def potential_to_accel(phi, ax_g, ay_g, az_g, dx: float = 1.0) -> None:
    """Node acceleration field g = -grad(phi) (the reference's "grad" step)."""
    n = _grid_n(phi)
    nat.call("pm_grad", nat.ptr(as_tensor(phi)), nat.ptr(as_tensor(ax_g)), nat.ptr(as_tensor(ay_g)),
             nat.ptr(as_tensor(az_g)), n, float(dx), nat.stream())


def gather_acceleration(parts, ax_g, ay_g, az_g, acc_x, acc_y, acc_z, dx: float = 1.0) -> None:
    """CIC-interpolate the node acceleration field to the particles."""
    n = _grid_n(ax_g)
    nat.call("pm_gather", nat.ptr(as_tensor(parts.pos_x)), nat.ptr(as_tensor(parts.pos_y)),
             nat.ptr(as_tensor(parts.pos_z)), nat.ptr(as_tensor(ax_g)), nat.ptr(as_tensor(ay_g)),
             nat.ptr(as_tensor(az_g)), nat.ptr(as_tensor(acc_x)), nat.ptr(as_tensor(acc_y)),
             nat.ptr(as_tensor(acc_z)), int(parts.n), n, float(dx), nat.stream())


__all__ = ["deposit_density", "deposit_keys", "gather_acceleration", "potential_to_accel",
           "DepositPlan", "deposit_plan"]
