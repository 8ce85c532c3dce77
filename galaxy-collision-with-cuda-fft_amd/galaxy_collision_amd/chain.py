"""The fused PM step: one KDK leapfrog step of the whole force chain on the device.

``PMChain.step(dt)`` runs, on torch's current stream,

    kick(dt/2) + drift(dt)     (pm_kick_drift, one pass over the particles)
    deposit                    (sort-based deterministic CIC deposit)
    solve                      (the configured Poisson solver, warm-started)
    g = -grad(phi)             (pm_grad4_bricks: interleaved (gx, gy, gz, 0) at the nodes of
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
from galaxy_collision_amd.deposit import deposit_plan


class PMChain:
    """deposit_mode: "sorted" (default: cell-key sort + ordered segmented reduction) or
    "fixed" (64-bit fixed-point integer atomics, order independent); both deterministic.
    At 256^3 / 10M particles "fixed" is slower (its 64-bit device-scope atomics), see the
    deposit mode notes in pmchain.h."""

    def __init__(self, parts, grid_size: int, dx: float, solver, grid: GridState | None = None,
                 deposit_mode: str = "sorted"):
        self.parts = parts
        self.n = int(grid_size)
        self.dx = float(dx)
        self.solver = solver
        self.grid = grid or GridState(self.n)
        dev = as_tensor(parts.pos_x).device
        self.acc = [torch.zeros(parts.n, device=dev) for _ in range(3)]
        self.g4 = torch.empty((self.n, self.n, self.n, 4), device=dev)
        self.plan = deposit_plan(self.n, parts.n)
        self.deposit_mode = deposit_mode
        self.warm = False
        self.last_cycles: list[int] = []

    def _p(self, f):
        return nat.ptr(as_tensor(f))

    def deposit(self) -> None:
        p = self.parts
        self.plan.set_mode(self.deposit_mode)
        self.plan.deposit(p.pos_x, p.pos_y, p.pos_z, p.mass, p.n, self.dx, self.grid.rho)

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
        velocities, acc, rho, phi) is bitwise that of ``steps`` calls of ``step``."""
        if steps <= 0:
            return
        p, g = self.parts, self.grid
        h = 0.5 * dt
        pos = (self._p(p.pos_x), self._p(p.pos_y), self._p(p.pos_z))
        vel = (self._p(p.vel_x), self._p(p.vel_y), self._p(p.vel_z))
        nat.call("pm_kick_drift", *pos, *vel, *(nat.ptr(a) for a in self.acc), p.n, h, dt,
                 nat.stream())
        for s in range(steps):
            self.deposit()
            self.solve()
            bricks = self.plan.node_bricks()
            if bricks:  # only the nodes the gather below reads
                nat.call("pm_grad4_bricks", self._p(g.phi), nat.ptr(self.g4), self.n, self.dx,
                         ctypes.c_void_p(bricks), nat.stream())
            else:
                nat.call("pm_grad4", self._p(g.phi), nat.ptr(self.g4), self.n, self.dx,
                         nat.stream())
            if s + 1 < steps:
                nat.call("pm_gather_kick_drift4", *pos, nat.ptr(self.g4), *vel, p.n, self.n,
                         self.dx, h, h, dt, nat.stream())
            else:
                nat.call("pm_gather_kick4", *pos, nat.ptr(self.g4),
                         *(nat.ptr(a) for a in self.acc), *vel, p.n, self.n, self.dx, h,
                         nat.stream())


__all__ = ["PMChain"]
