"""Zero-padded FFT Poisson solve on the device (reference solver/fft_oracle.py:54-109).

phi = rho (*) g with the continuum kernel g = -G*dx^3/r (self term 0), computed as a
circular convolution on a grid padded to ``pad_factor*n`` per axis so no periodic images
reach the box. Float64 throughout, rocFFT through torch.fft; the kernel transform is built
once. A validation solver (the reference's oracle role), not the production path.
"""

from __future__ import annotations

import torch

from galaxy_collision_amd import units
from galaxy_collision_amd.data import as_tensor
from galaxy_collision_amd.solver.base import PoissonSolver


class FFTPoissonSolver(PoissonSolver):
    def __init__(self, grid_size: int, dx: float = 1.0, pad_factor: int = 2,
                 grav_constant: float | None = None, device=None):
        if pad_factor < 2:
            raise ValueError(f"pad_factor must be >= 2 to avoid circular wrap, got {pad_factor}")
        if not torch.cuda.is_available():
            raise RuntimeError("FFTPoissonSolver runs on the HIP device")
        self.grid_size = int(grid_size)
        self.dx = float(dx)
        self.pad = pad_factor * self.grid_size
        self.G = units.G if grav_constant is None else float(grav_constant)
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self._greens_ft = self._build_greens_ft()

    def _build_greens_ft(self) -> torch.Tensor:
        m = self.pad
        i = torch.arange(m, device=self.device, dtype=torch.float64)
        d = torch.minimum(i, m - i)
        d2 = d * d
        r2 = d2[:, None, None] + d2[None, :, None] + d2[None, None, :]
        r2[0, 0, 0] = 1.0
        green = (-self.G * self.dx ** 3 / self.dx) / torch.sqrt(r2)
        green[0, 0, 0] = 0.0
        out = torch.fft.rfftn(green)
        del green, r2
        return out

    def solve(self, rho, phi, warm_start: bool = False) -> None:
        r, p = as_tensor(rho), as_tensor(phi)
        n, m = self.grid_size, self.pad
        padded = torch.zeros((m, m, m), dtype=torch.float64, device=self.device)
        padded[:n, :n, :n] = r.to(torch.float64)
        spec = torch.fft.rfftn(padded)
        del padded
        spec *= self._greens_ft
        conv = torch.fft.irfftn(spec, s=(m, m, m))
        p.copy_(conv[:n, :n, :n])  # rounded once to phi's dtype (float32 fields; float64 kept)


__all__ = ["FFTPoissonSolver"]
