"""Device 2D density projection (reference viz/project.py:26-83).

``scatter_density(parts, img, extent, axes)`` fills the device image ``img`` (float32
``(bins, bins)``, indexed ``[a, b]``) with the mass per nearest bin of the in-plane axes
``a, b`` and returns the bin widths; ``project_density`` adds the host copy and the
conversion to surface density, oriented for ``imshow(origin="lower", extent=extent)`` (the
transpose), as the reference does. Out-of-extent particles are dropped. The sums are exact
fixed-point integers on the device, so a frame is identical run to run.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from galaxy_collision_amd import _native as nat
from galaxy_collision_amd.data import as_tensor

_AXIS = {0: "pos_x", 1: "pos_y", 2: "pos_z"}
_scratch: dict[tuple[int, int], torch.Tensor] = {}


def _scratch_for(bins: int, dev) -> torch.Tensor:
    key = (bins, dev.index if dev.index is not None else torch.cuda.current_device())
    buf = _scratch.get(key)
    if buf is None:
        n = ctypes.c_int64()
        nat.call("pm_project_scratch_bytes", int(bins), ctypes.byref(n))
        buf = torch.empty((n.value + 7) // 8, dtype=torch.int64, device=dev)
        _scratch[key] = buf
    return buf


def scatter_density(parts, img, extent, axes=(0, 1)) -> tuple[float, float]:
    """Device only: mass per bin into ``img``; returns ``(binw_a, binw_b)``."""
    a, b = axes
    t = as_tensor(img)
    if t.dim() != 2 or t.shape[0] != t.shape[1]:
        raise ValueError(f"img must be a square 2D field, got shape {tuple(t.shape)}")
    bins = int(t.shape[0])
    lo_a, hi_a, lo_b, hi_b = (float(v) for v in extent)
    binw_a = (hi_a - lo_a) / bins
    binw_b = (hi_b - lo_b) / bins
    pa, pb = as_tensor(getattr(parts, _AXIS[a])), as_tensor(getattr(parts, _AXIS[b]))
    m = as_tensor(parts.mass)
    scratch = _scratch_for(bins, t.device)
    nat.call("pm_project_mass", nat.ptr(pa), nat.ptr(pb), nat.ptr(m), int(parts.n),
             lo_a, lo_b, binw_a, binw_b, bins, nat.ptr(t), nat.ptr(scratch, torch.int64),
             nat.stream())
    return binw_a, binw_b


def project_density(parts, img, extent, axes=(0, 1)) -> np.ndarray:
    """Surface density image (mass / bin area), transposed for imshow(origin="lower")."""
    binw_a, binw_b = scatter_density(parts, img, extent, axes)
    return as_tensor(img).cpu().numpy().T / (binw_a * binw_b)


__all__ = ["scatter_density", "project_density"]
