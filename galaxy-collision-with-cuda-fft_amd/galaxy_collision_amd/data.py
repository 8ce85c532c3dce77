"""Structure-of-arrays particle and grid state in HBM (reference data.py:84-138).

Every field is one contiguous device tensor (float32, grids C-order with k fastest). A
``DeviceField`` wraps it with the small field API the reference's callers use
(``to_numpy``, ``from_numpy``, ``fill``, ``shape``, ``[i]``), so drivers written against the
reference run unchanged; ``.tensor`` exposes the torch tensor.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

_F32 = 4
_I32 = 4

# Particle fields: pos(x,y,z) + vel(x,y,z) + mass f32, gid i32 (reference data.py:33-34).
BYTES_PER_PARTICLE = 6 * _F32 + _F32 + _I32
# rho + phi, the node acceleration field, and the multigrid hierarchy (phi/rhs/res per
# level, 8/7 of the finest) in units of n^3 float32 arrays.
_GRID_FIELDS = 2 + 3 + 3.0 * 8.0 / 7.0


@dataclass(frozen=True)
class MemoryEstimate:
    n_particles: int
    grid_size: int
    particle_bytes: int
    grid_bytes: int

    @property
    def total_bytes(self) -> int:
        return self.particle_bytes + self.grid_bytes

    @property
    def total_gb(self) -> float:
        return self.total_bytes / 1e9


def estimate_memory(n_particles: int, grid_size: int = 256) -> MemoryEstimate:
    """SoA footprint of the production path (same accounting as reference data.py:61-81)."""
    if n_particles <= 0:
        raise ValueError(f"n_particles must be positive, got {n_particles}")
    if grid_size <= 0:
        raise ValueError(f"grid_size must be positive, got {grid_size}")
    return MemoryEstimate(n_particles, grid_size, n_particles * BYTES_PER_PARTICLE,
                          int(_GRID_FIELDS * grid_size ** 3 * _F32))


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("a HIP device is required for device-resident state")
    return torch.device("cuda", torch.cuda.current_device())


class DeviceField:
    """A contiguous device tensor with the field methods reference callers use."""

    __slots__ = ("tensor",)

    def __init__(self, tensor: torch.Tensor):
        if not tensor.is_contiguous():
            raise ValueError("DeviceField needs a contiguous tensor")
        self.tensor = tensor

    @classmethod
    def zeros(cls, shape, dtype=torch.float32, device=None) -> "DeviceField":
        if isinstance(shape, int):
            shape = (shape,)
        return cls(torch.zeros(tuple(shape), dtype=dtype, device=device or default_device()))

    @property
    def shape(self) -> tuple[int, ...]:
        return tuple(self.tensor.shape)

    @property
    def dtype(self):
        return self.tensor.dtype

    def to_numpy(self) -> np.ndarray:
        return self.tensor.detach().cpu().numpy()

    def from_numpy(self, arr: np.ndarray) -> None:
        arr = np.asarray(arr)
        if tuple(arr.shape) != self.shape:
            raise ValueError(f"shape mismatch: field {self.shape}, array {tuple(arr.shape)}")
        src = torch.from_numpy(np.ascontiguousarray(arr)).to(self.tensor.dtype)
        self.tensor.copy_(src)

    def fill(self, value) -> None:
        self.tensor.fill_(value)

    def __getitem__(self, idx):
        v = self.tensor[idx]
        return v.item() if v.dim() == 0 else v

    def __setitem__(self, idx, value) -> None:
        self.tensor[idx] = value

    def __len__(self) -> int:
        return self.tensor.shape[0]


def as_tensor(f) -> torch.Tensor:
    """The tensor behind a DeviceField (or the tensor itself)."""
    return f.tensor if isinstance(f, DeviceField) else f


class ParticleState:
    """SoA particle fields on the device (reference data.py:84-113).

    ``pid`` holds each particle's index in load order. A chain that reorders the particles
    (``PMChain(reorder_every=...)``) moves every field, ``pid`` included, into a second set
    of arrays and swaps the sets, so ``pid`` keeps identifying particles.
    """

    def __init__(self, n: int, device=None):
        if n <= 0:
            raise ValueError(f"particle count must be positive, got {n}")
        dev = device or default_device()
        self.n = n
        mk = lambda dtype=torch.float32: DeviceField.zeros(n, dtype, dev)  # noqa: E731
        self.pos_x, self.pos_y, self.pos_z = mk(), mk(), mk()
        self.vel_x, self.vel_y, self.vel_z = mk(), mk(), mk()
        self.mass = mk()
        self.gid = mk(torch.int32)
        self.pid = DeviceField(torch.arange(n, dtype=torch.int32, device=dev))
        self._shadow: list[torch.Tensor] | None = None

    def fields(self):
        return (self.pos_x, self.pos_y, self.pos_z, self.vel_x, self.vel_y, self.vel_z,
                self.mass, self.gid, self.pid)

    def arrays(self, shadow: bool = False):
        """pm_particle_arrays of the fields (or of the second set, allocated on first use)."""
        from galaxy_collision_amd._native import ParticleArrays

        if shadow:
            if self._shadow is None:
                self._shadow = [torch.empty_like(f.tensor) for f in self.fields()]
            ptrs = [t.data_ptr() for t in self._shadow]
        else:
            ptrs = [f.tensor.data_ptr() for f in self.fields()]
        return ParticleArrays(*ptrs)

    def swap_shadow(self) -> None:
        """Make the second set of arrays the fields (after it was filled)."""
        if self._shadow is None:
            raise RuntimeError("no second set of particle arrays to swap in")
        for k, f in enumerate(self.fields()):
            f.tensor, self._shadow[k] = self._shadow[k], f.tensor

    def slots_of(self, indices) -> torch.Tensor:
        """Current slots (device int64) of the particles with the given load indices."""
        pid = self.pid.tensor
        slot = torch.empty(pid.numel(), dtype=torch.int64, device=pid.device)
        slot[pid.long()] = torch.arange(pid.numel(), dtype=torch.int64, device=pid.device)
        return slot[torch.as_tensor(np.asarray(indices), dtype=torch.int64, device=pid.device)]

    def in_load_order(self) -> np.ndarray | None:
        """Host permutation that puts host copies of the fields back into load order, or None
        when the particles are in load order."""
        pid = self.pid.to_numpy()
        if np.array_equal(pid, np.arange(pid.shape[0], dtype=pid.dtype)):
            return None
        return np.argsort(pid, kind="stable")

    def load(self, pos: np.ndarray, vel: np.ndarray | None = None, mass: np.ndarray | None = None,
             gid: np.ndarray | None = None) -> None:
        """Copy host (N,3)/(N,) arrays into the device fields."""
        pos = np.asarray(pos, dtype=np.float32)
        for a, f in enumerate((self.pos_x, self.pos_y, self.pos_z)):
            f.from_numpy(np.ascontiguousarray(pos[:, a]))
        if vel is not None:
            vel = np.asarray(vel, dtype=np.float32)
            for a, f in enumerate((self.vel_x, self.vel_y, self.vel_z)):
                f.from_numpy(np.ascontiguousarray(vel[:, a]))
        if mass is not None:
            self.mass.from_numpy(np.asarray(mass, dtype=np.float32))
        if gid is not None:
            self.gid.from_numpy(np.asarray(gid, dtype=np.int32))
        t = self.pid.tensor
        t.copy_(torch.arange(t.numel(), dtype=torch.int32, device=t.device))

    def load_tensors(self, pos: torch.Tensor, vel: torch.Tensor | None = None,
                     mass: torch.Tensor | None = None, gid: torch.Tensor | None = None) -> None:
        """``load`` from (N,3)/(N,) torch tensors (device to device when they are on it)."""
        if pos.shape != (self.n, 3):
            raise ValueError(f"pos must be ({self.n}, 3), got {tuple(pos.shape)}")
        for a, f in enumerate((self.pos_x, self.pos_y, self.pos_z)):
            f.tensor.copy_(pos[:, a])
        if vel is not None:
            for a, f in enumerate((self.vel_x, self.vel_y, self.vel_z)):
                f.tensor.copy_(vel[:, a])
        if mass is not None:
            self.mass.tensor.copy_(mass)
        if gid is not None:
            self.gid.tensor.copy_(gid)
        t = self.pid.tensor
        t.copy_(torch.arange(t.numel(), dtype=torch.int32, device=t.device))

    def positions(self) -> np.ndarray:
        return np.stack([f.to_numpy() for f in (self.pos_x, self.pos_y, self.pos_z)], axis=1)

    def velocities(self) -> np.ndarray:
        return np.stack([f.to_numpy() for f in (self.vel_x, self.vel_y, self.vel_z)], axis=1)


class GridState:
    """Device grids rho, phi and g = -grad(phi) (reference data.py:116-138)."""

    def __init__(self, grid_size: int = 256, device=None):
        if grid_size <= 0:
            raise ValueError(f"grid_size must be positive, got {grid_size}")
        dev = device or default_device()
        self.grid_size = grid_size
        shape = (grid_size,) * 3
        self.rho = DeviceField.zeros(shape, device=dev)
        self.phi = DeviceField.zeros(shape, device=dev)
        self.ax = DeviceField.zeros(shape, device=dev)
        self.ay = DeviceField.zeros(shape, device=dev)
        self.az = DeviceField.zeros(shape, device=dev)


__all__ = ["BYTES_PER_PARTICLE", "MemoryEstimate", "estimate_memory", "DeviceField",
           "ParticleState", "GridState", "as_tensor"]
