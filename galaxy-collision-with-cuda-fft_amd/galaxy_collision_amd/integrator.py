"""Kick-drift-kick leapfrog on libpmchain (reference integrator.py).

``kdk_step(parts, acc_x, acc_y, acc_z, accel_fn, dt)`` runs kick(dt/2) -> drift(dt) ->
accel_fn() -> kick(dt/2) with one force evaluation per step: ``acc_*`` must hold a(x) for
the current positions on entry (prime once with ``accel_fn()``) and hold a(x_new) on exit
(reference integrator.py:100-131). Any out-of-band position change requires calling
``accel_fn()`` again before the next step.
"""

from __future__ import annotations

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.data import as_tensor


def _p(f):
    return nat.ptr(as_tensor(f))


def kick(vel_x, vel_y, vel_z, acc_x, acc_y, acc_z, n: int, dt: float) -> None:
    """v += a*dt (reference integrator.py:27-42)."""
    nat.call("pm_kick", _p(vel_x), _p(vel_y), _p(vel_z), _p(acc_x), _p(acc_y), _p(acc_z), int(n),
             float(dt), nat.stream())


def drift(pos_x, pos_y, pos_z, vel_x, vel_y, vel_z, n: int, dt: float) -> None:
    """x += v*dt (reference integrator.py:45-60)."""
    nat.call("pm_drift", _p(pos_x), _p(pos_y), _p(pos_z), _p(vel_x), _p(vel_y), _p(vel_z), int(n),
             float(dt), nat.stream())


def direct_accel(pos_x, pos_y, pos_z, mass, acc_x, acc_y, acc_z, n: int, grav: float,
                 eps2: float) -> None:
    """Direct O(N^2) Plummer-softened acceleration for small N (reference integrator.py:63-97)."""
    nat.call("pm_direct_accel", _p(pos_x), _p(pos_y), _p(pos_z), _p(mass), _p(acc_x), _p(acc_y),
             _p(acc_z), int(n), float(grav), float(eps2), nat.stream())


def kdk_step(parts, acc_x, acc_y, acc_z, accel_fn, dt: float) -> None:
    """One KDK leapfrog step with a single acceleration evaluation (see module docstring)."""
    n = parts.n
    half = 0.5 * dt
    kick(parts.vel_x, parts.vel_y, parts.vel_z, acc_x, acc_y, acc_z, n, half)
    drift(parts.pos_x, parts.pos_y, parts.pos_z, parts.vel_x, parts.vel_y, parts.vel_z, n, dt)
    accel_fn()
    kick(parts.vel_x, parts.vel_y, parts.vel_z, acc_x, acc_y, acc_z, n, half)


__all__ = ["kick", "drift", "direct_accel", "kdk_step"]
