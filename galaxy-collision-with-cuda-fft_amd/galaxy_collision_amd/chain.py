"""The fused PM step: one KDK leapfrog step of the whole force chain on the device.

``PMChain.step(dt)`` runs, on torch's current stream,

    kick(dt/2) + drift(dt)     (pm_kick_drift, one pass over the particles)
    deposit                    (sort-based deterministic CIC deposit)
    solve                      (the configured Poisson solver, warm-started)
    g = -grad(phi)             (pm_grad4_listed: interleaved (gx, gy, gz, 0) at the nodes of
                               the 8^3 bricks the deposit marked, i.e. every node the
                               gather reads; pm_grad4 on every node after a fixed-mode
                               deposit)
    gather + kick(dt/2)        (pm_gather_kick4, one pass over the particles)

which is bit for bit the reference step kick -> drift -> accel_fn -> kick
(reference integrator.py:100-131 with the accel_fn of sim.py:318-325), with the two
particle passes fused. ``PMChain.run(steps, dt)`` runs consecutive steps with the closing
kick of one step, the opening kick and the drift of the next in one particle pass
(pm_gather_kick_drift4), which leaves every particle and field value bitwise as ``steps``
calls of ``step`` would. The interleaved gradient turns the gather's 24 scattered loads per
particle into 8 16-byte loads; it lives in the chain's own buffer, so the grid's ax/ay/az
fields hold the gradient of the last prime()/accel() call, not of the last step.
"""

from __future__ import annotations

import ctypes

import torch

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.data import GridState, as_tensor
from galaxy_collision_amd.deposit import DepositPlan


class PMChain:
    """deposit_mode: "sorted" (default: cell-key sort + ordered segmented reduction),
    "fixed" (64-bit fixed-point integer atomics, order independent) or "bricked" (the
    "fixed" result bit for bit, summed per 8^3-cell brick in LDS from a binning of the
    particles that is redone only when the particles are reordered, or every 4 deposits
    without reordering); all deterministic. At 256^3 / 10M particles "fixed" is slower (its
    64-bit device-scope atomics), see the deposit mode notes in pmchain.h. The fixed-point
    modes do not depend on the particles' order, so reordering leaves their runs bitwise
    unchanged (up to the slots).

    reorder_every: K > 0 moves the particles into the last deposit's cell order in the first
    drift of a run that follows K or more deposits since the last move
    (pm_gather_kick_drift4_reorder; sorted deposits only),
    so that particles of one cell stay neighbours in memory: as they drift out of order the
    deposit's and the gather's loads scatter and slow down (at 256^3 / 10M particles the
    piece sums take 112 us after 10 steps, 145 after 25 and 265 after 320 unordered).
    Particle slots then change (``parts.pid`` keeps each particle's load index; the fields
    keep their ``DeviceField`` objects but swap the tensors behind them, so read
    ``field.tensor`` after a run rather than holding it across one), and since
    a cell's particles are summed in slot order the density changes at rounding level
    against an unordered run; every step stays deterministic. A reorder needs a following
    drift, so only multi-step ``run`` calls reorder (``step`` never does)."""

    def __init__(self, parts, grid_size: int, dx: float, solver, grid: GridState | None = None,
                 deposit_mode: str = "sorted", reorder_every: int = 0):
        self.parts = parts
        self.n = int(grid_size)
        self.dx = float(dx)
        self.solver = solver
        self.grid = grid or GridState(self.n)
        dev = as_tensor(parts.pos_x).device
        self.acc = [torch.zeros(parts.n, device=dev) for _ in range(3)]
        self.g4 = torch.empty((self.n, self.n, self.n, 4), device=dev)
        # the chain's own plan: the shared per-grid plan of deposit_density() is re-created
        # when a larger particle set arrives, which would pull the scratch from under a chain
        self.plan = DepositPlan(self.n, max(parts.n, 1))
        # cell keys of the current positions, written by the fused drift pass for the
        # next deposit (valid while keys_valid)
        self.keys = torch.empty(max(parts.n, 1), dtype=torch.int32, device=dev)
        self.keys_valid = False
        self.deposit_mode = deposit_mode
        if int(reorder_every) < 0:
            raise ValueError(f"reorder_every must be >= 0, got {reorder_every}")
        self.reorder_every = int(reorder_every)
        self.steps_done = 0
        self._since_order = 0  # deposits since the particles were last put in cell order
        # "bricked" mode: deposits since the last binning, binning cadence without reorders,
        # and a request to bin at the next deposit (set when a reorder follows it)
        self._since_bin = 0
        self._bin_every = self.reorder_every if self.reorder_every > 0 else 4
        self._bin_next = False
        self.warm = False
        self.last_cycles: list[int] = []

    def _p(self, f):
        return nat.ptr(as_tensor(f))

    def deposit(self) -> None:
        p = self.parts
        self.plan.set_mode(self.deposit_mode)
        if self.deposit_mode == "bricked":
            # bin again when the drift after this deposit moves the particles into the
            # binning's order, or when the binning has aged without such a move
            due = self._bin_next or self._since_bin + 1 >= self._bin_every
            # the grid's rho is this chain's own: only the chain's deposits write it
            self.plan.deposit_step(p.pos_x, p.pos_y, p.pos_z, p.mass, p.n, self.dx,
                                   self.grid.rho, bin=due, rho_kept=True)
            self._since_bin = 0 if due else self._since_bin + 1
            self._bin_next = False
            self.keys_valid = False
            return
        if self.keys_valid:
            self.plan.deposit_keyed(p.pos_x, p.pos_y, p.pos_z, p.mass, p.n, self.dx,
                                    self.grid.rho, self.keys)
        else:
            self.plan.deposit(p.pos_x, p.pos_y, p.pos_z, p.mass, p.n, self.dx, self.grid.rho)
        self.keys_valid = False

    def solve(self) -> None:
        self.solver.solve(self.grid.rho, self.grid.phi, warm_start=self.warm)
        self.warm = True
        if hasattr(self.solver, "last_cycles"):
            self.last_cycles.append(int(self.solver.last_cycles))

    def grad(self) -> None:
        g = self.grid
        nat.call("pm_grad", self._p(g.phi), self._p(g.ax), self._p(g.ay), self._p(g.az), self.n,
                 self.dx, nat.stream())

    def gather(self) -> None:
        p, g = self.parts, self.grid
        nat.call("pm_gather", self._p(p.pos_x), self._p(p.pos_y), self._p(p.pos_z),
                 self._p(g.ax), self._p(g.ay), self._p(g.az), *(nat.ptr(a) for a in self.acc),
                 p.n, self.n, self.dx, nat.stream())

    def accel(self) -> None:
        """deposit -> solve -> grad -> gather: acc = a(x) for the current positions."""
        self.deposit()
        self.solve()
        self.grad()
        self.gather()

    prime = accel

    def step(self, dt: float) -> None:
        self.run(1, dt)

    def run(self, steps: int, dt: float) -> None:
        """``steps`` consecutive KDK steps of size dt; the state afterwards (positions,
        velocities, acc, rho, phi) is bitwise that of ``steps`` calls of ``step`` unless the
        run moves the particles into cell order (``reorder_every``)."""
        if steps <= 0:
            return
        p, g = self.parts, self.grid
        h = 0.5 * dt
        pos = (self._p(p.pos_x), self._p(p.pos_y), self._p(p.pos_z))
        vel = (self._p(p.vel_x), self._p(p.vel_y), self._p(p.vel_z))
        nat.call("pm_kick_drift", *pos, *vel, *(nat.ptr(a) for a in self.acc), p.n, h, dt,
                 nat.stream())
        bricked = self.deposit_mode == "bricked"
        for s in range(steps):
            if bricked:
                self._bin_next = self._reorder_due(s, steps)
            self.deposit()
            self.solve()
            listed = self.plan.brick_list()
            if listed:  # only the nodes the gather below reads
                lst, cnt, mx = listed
                nat.call("pm_grad4_listed", self._p(g.phi), nat.ptr(self.g4), self.n, self.dx,
                         ctypes.c_void_p(lst), ctypes.c_void_p(cnt), mx, nat.stream())
            else:
                nat.call("pm_grad4", self._p(g.phi), nat.ptr(self.g4), self.n, self.dx,
                         nat.stream())
            order = self._reorder_order() if s + 1 < steps else None
            if order:
                # the drift also moves the particles into this deposit's cell order
                nat.call("pm_gather_kick_drift4_reorder", ctypes.c_void_p(order),
                         ctypes.byref(p.arrays()), ctypes.byref(p.arrays(shadow=True)),
                         nat.ptr(self.g4), p.n, self.n, self.dx, h, h, dt,
                         nat.ptr(self.keys, torch.int32), nat.stream())
                p.swap_shadow()
                if bricked:
                    self.plan.mark_ordered()
                self._since_order = -1  # this step's deposit is the order's own
                pos = (self._p(p.pos_x), self._p(p.pos_y), self._p(p.pos_z))
                vel = (self._p(p.vel_x), self._p(p.vel_y), self._p(p.vel_z))
                self.keys_valid = True
            elif s + 1 < steps:
                # also writes the drifted particles' cell keys for the next sorted deposit
                keys = None if bricked else nat.ptr(self.keys, torch.int32)
                nat.call("pm_gather_kick_drift4", *pos, nat.ptr(self.g4), *vel, p.n, self.n,
                         self.dx, h, h, dt, keys, nat.stream())
                self.keys_valid = not bricked
            else:
                nat.call("pm_gather_kick4", *pos, nat.ptr(self.g4),
                         *(nat.ptr(a) for a in self.acc), *vel, p.n, self.n, self.dx, h,
                         nat.stream())
            self.steps_done += 1
            self._since_order += 1

    def _reorder_due(self, s: int, steps: int) -> bool:
        """Whether step s of a run of ``steps`` moves the particles into cell order."""
        return (s + 1 < steps and self.reorder_every > 0
                and self._since_order + 1 >= self.reorder_every)

    def _reorder_order(self) -> int | None:
        """Device address of the particle order to move to in this step's drift, or None."""
        if self.reorder_every <= 0 or self._since_order + 1 < self.reorder_every:
            return None
        return self.plan.order()


__all__ = ["PMChain"]
