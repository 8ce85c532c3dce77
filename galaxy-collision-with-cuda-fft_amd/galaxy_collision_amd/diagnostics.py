"""Host-side conservation formulas (reference diagnostics.py) for small-N validation.

The production diagnostics are reduced on the device (``diagnostics_device``); these float64
host formulas serve the clean energy metric of small validation runs.
"""

from __future__ import annotations

import numpy as np

from galaxy_collision_amd.diagnostics_device import virial_ratio


def potential_energy_direct(mass, pos, softening: float = 0.0, grav: float = 1.0,
                            chunk: int = 2048) -> float:
    """-1/2 G sum_{i != j} m_i m_j / sqrt(r_ij^2 + eps^2), float64, O(N^2) in row chunks
    (reference diagnostics.py:120-132)."""
    m = np.asarray(mass, dtype=np.float64)
    p = np.asarray(pos, dtype=np.float64)
    eps2 = float(softening) ** 2
    total = 0.0
    for i0 in range(0, len(m), chunk):
        d = p[i0:i0 + chunk, None, :] - p[None, :, :]
        r2 = (d * d).sum(-1) + eps2
        rows = np.arange(i0, min(i0 + chunk, len(m)))
        r2[rows - i0, rows] = np.inf  # no self-interaction
        total += float((m[i0:i0 + chunk, None] * m[None, :] / np.sqrt(r2)).sum())
    return -0.5 * float(grav) * total


__all__ = ["potential_energy_direct", "virial_ratio"]
