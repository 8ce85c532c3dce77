"""Device-resident float64 conservation diagnostics (reference diagnostics_device.py:222-322).

``DeviceDiagnostics(grid_size).sample(parts, grid, dx)`` returns the reference's keys:
``M``, ``kinetic``, ``potential`` (grid PE = 1/2 dx^3 sum rho*phi), ``energy``,
``momentum`` and ``ang_momentum`` (about the origin, float64 3-vectors),
``center_of_mass`` and, with ``half_mass=True``, ``half_mass_radius`` from a float64 mass
histogram of radius about the centre of mass (last bin a catch-all) with linear
interpolation inside the crossing bin. Only 11 + 1 scalars and the histogram reach the host.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.data import as_tensor


class DeviceDiagnostics:
    def __init__(self, grid_size: int, *, nbins: int = 2048, r_max: float | None = None,
                 dx: float = 1.0):
        nat.require_device()
        self.grid_size = int(grid_size)
        self.nbins = int(nbins)
        self.r_max = float(r_max) if r_max is not None else float(grid_size) * dx * np.sqrt(3.0)
        self.binw = self.r_max / self.nbins
        self._hist = torch.zeros(self.nbins, dtype=torch.float64,
                                 device=torch.device("cuda", torch.cuda.current_device()))

    def _half_mass_radius(self, parts, cx: float, cy: float, cz: float, total: float) -> float:
        p = lambda f: nat.ptr(as_tensor(f))  # noqa: E731
        nat.call("pm_radial_histogram_f64", p(parts.pos_x), p(parts.pos_y), p(parts.pos_z),
                 p(parts.mass), int(parts.n), float(np.float32(cx)), float(np.float32(cy)),
                 float(np.float32(cz)), float(np.float32(1.0 / self.binw)), self.nbins,
                 nat.ptr(self._hist, torch.float64), nat.stream())
        hist = self._hist.cpu().numpy()
        cum = np.cumsum(hist)
        target = 0.5 * total
        b = int(np.searchsorted(cum, target))
        if b >= self.nbins:
            return self.r_max
        cum_lo = cum[b - 1] if b > 0 else 0.0
        frac = (target - cum_lo) / hist[b] if hist[b] > 0.0 else 0.0
        return float((b + frac) * self.binw)

    def sample(self, parts, grid, dx: float, *, half_mass: bool = True) -> dict:
        p = lambda f: nat.ptr(as_tensor(f))  # noqa: E731
        s = (ctypes.c_double * 11)()
        nat.call("pm_reduce_particles_f64", p(parts.pos_x), p(parts.pos_y), p(parts.pos_z),
                 p(parts.vel_x), p(parts.vel_y), p(parts.vel_z), p(parts.mass), int(parts.n), s,
                 nat.stream())
        raw_pe = ctypes.c_double()
        rho, phi = as_tensor(grid.rho), as_tensor(grid.phi)
        nat.call("pm_reduce_grid_pe_f64", nat.ptr(rho), nat.ptr(phi), rho.numel(),
                 ctypes.byref(raw_pe), nat.stream())
        s = np.array(list(s))
        pe = 0.5 * float(dx) ** 3 * raw_pe.value
        total = float(s[0])
        ke = float(s[1])
        com = s[8:11] / total if total > 0.0 else np.zeros(3)
        out = {"M": total, "kinetic": ke, "potential": pe, "energy": ke + pe,
               "momentum": s[2:5].copy(), "ang_momentum": s[5:8].copy(), "center_of_mass": com}
        if half_mass:
            out["half_mass_radius"] = self._half_mass_radius(parts, *map(float, com), total)
        return out


def virial_ratio(kinetic: float, potential: float) -> float:
    """T/|W| (reference diagnostics.py:161-167)."""
    return kinetic / abs(potential) if potential != 0.0 else float("nan")


__all__ = ["DeviceDiagnostics", "virial_ratio"]
