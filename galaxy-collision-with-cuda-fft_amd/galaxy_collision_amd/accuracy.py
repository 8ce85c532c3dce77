"""Potential accuracy measurements (host float64; a diagnostic, not a solver path).

``phi_accuracy`` compares a solved phi with
(a) the exact float64 solution of the same discrete problem (7-point Laplacian, the solve's
    own Dirichlet faces, the same float32 right-hand side 4*pi*G*rho), diagonalised by the
    type-I discrete sine transform (scipy.fft.dstn): the algebraic error of the solver;
(b) the zero-padded continuum-kernel FFT solve (float64, on the device), the reference's
    validation solver (solver/fft_oracle.py:54-109): algebraic plus discretization error.
"""

from __future__ import annotations

import math

import numpy as np
import torch


def dst_exact(f: np.ndarray, boundary: np.ndarray, dx: float) -> np.ndarray:
    """Exact float64 solution of lap7(phi) = f inside with phi = boundary on the faces."""
    from scipy.fft import dstn, idstn

    n = f.shape[0]
    b = f[1:-1, 1:-1, 1:-1].astype(np.float64) * dx * dx
    bd = boundary.astype(np.float64).copy()
    bd[1:-1, 1:-1, 1:-1] = 0.0
    b -= (bd[2:, 1:-1, 1:-1] + bd[:-2, 1:-1, 1:-1] + bd[1:-1, 2:, 1:-1] + bd[1:-1, :-2, 1:-1]
          + bd[1:-1, 1:-1, 2:] + bd[1:-1, 1:-1, :-2])
    m = n - 2
    lam1 = 2.0 * (np.cos(np.pi * np.arange(1, m + 1) / (m + 1)) - 1.0)
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    u = idstn(dstn(b, type=1, workers=-1) / lam, type=1, workers=-1)
    out = boundary.astype(np.float64).copy()
    out[1:-1, 1:-1, 1:-1] = u
    return out


def phi_rel_error_discrete(rho: np.ndarray, phi: np.ndarray, dx: float, grav: float) -> float:
    """||phi - exact|| / ||exact|| for the discrete problem phi was solved for."""
    f = (np.float32(4.0 * math.pi * grav) * rho).astype(np.float32)
    exact = dst_exact(f, phi.astype(np.float64), dx)
    return float(np.linalg.norm(phi.astype(np.float64) - exact) / np.linalg.norm(exact))


def phi_accuracy(rho_field, phi_field, n: int, dx: float, grav: float) -> dict:
    """Both errors of the solved phi (fields: DeviceField or device tensors)."""
    from galaxy_collision_amd.data import as_tensor
    from galaxy_collision_amd.solver.fft import FFTPoissonSolver

    rho_t, phi_t = as_tensor(rho_field), as_tensor(phi_field)
    rho = rho_t.cpu().numpy()
    phi = phi_t.cpu().numpy().astype(np.float64)
    alg = phi_rel_error_discrete(rho, phi, dx, grav)
    fft = FFTPoissonSolver(n, dx=dx, grav_constant=grav)
    out = torch.empty_like(phi_t)
    fft.solve(rho_t, out)
    ref = out.cpu().numpy().astype(np.float64)
    del fft, out
    torch.cuda.empty_cache()
    tot = float(np.linalg.norm(phi - ref) / np.linalg.norm(ref))
    return {"phi_rel_l2_vs_exact_discrete": alg, "phi_rel_l2_vs_fft_continuum": tot}


__all__ = ["dst_exact", "phi_rel_error_discrete", "phi_accuracy"]
