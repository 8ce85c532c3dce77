"""Snapshot I/O in the reference's on-disk format (reference io.py:25-200).

A :class:`Snapshot` holds NumPy arrays: ``pos`` (N, 3), ``vel`` (N, 3), ``mass`` (N,),
``gid`` (N,) int32, optional ``rho`` / ``phi`` (G, G, G), scalars ``step``, ``time``, ``dx``
and a dict of diagnostics. ``.npz`` files store every array under its own name, the
scalars as 0-d arrays and each diagnostic under ``diag__<name>`` (compressed), so files
written here and by the reference read back identically. ``.h5``/``.hdf5`` use h5py when
it is installed (datasets per array, scalars as attributes, diagnostics in a group).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

PARTICLE_KEYS = ("pos", "vel", "mass", "gid")
GRID_KEYS = ("rho", "phi")
META_KEYS = ("step", "time", "dx")
DIAG_PREFIX = "diag__"


@dataclass
class Snapshot:
    step: int
    time: float
    pos: np.ndarray
    vel: np.ndarray
    mass: np.ndarray
    gid: np.ndarray
    dx: float = 1.0
    rho: np.ndarray | None = None
    phi: np.ndarray | None = None
    diagnostics: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def snapshot_from_states(step, time, parts, grid=None, diagnostics=None, dx=1.0) -> Snapshot:
    """Copy a device ParticleState (and optionally GridState) into a Snapshot, particles in
    load order (a chain that reorders particles moves them; ``parts.pid`` maps them back)."""
    pos = np.stack([parts.pos_x.to_numpy(), parts.pos_y.to_numpy(), parts.pos_z.to_numpy()], 1)
    vel = np.stack([parts.vel_x.to_numpy(), parts.vel_y.to_numpy(), parts.vel_z.to_numpy()], 1)
    mass, gid = parts.mass.to_numpy(), parts.gid.to_numpy()
    back = parts.in_load_order() if hasattr(parts, "in_load_order") else None
    if back is not None:
        pos, vel, mass, gid = pos[back], vel[back], mass[back], gid[back]
    return Snapshot(
        step=int(step), time=float(time), pos=pos, vel=vel, mass=mass,
        gid=gid, dx=float(dx),
        rho=grid.rho.to_numpy() if grid is not None else None,
        phi=grid.phi.to_numpy() if grid is not None else None,
        diagnostics={k: np.asarray(v) for k, v in (diagnostics or {}).items()})


def write_snapshot(snap: Snapshot, path) -> Path:
    """Write by suffix (.npz, or .h5/.hdf5 with h5py); creates parent directories."""
    path = Path(path)
    suffix = path.suffix.lower()
    if suffix not in (".npz", ".h5", ".hdf5"):
        raise ValueError(f"unsupported snapshot extension {path.suffix!r} (use .npz/.h5/.hdf5)")
    path.parent.mkdir(parents=True, exist_ok=True)
    if suffix == ".npz":
        arrays = {k: np.asarray(getattr(snap, k)) for k in META_KEYS}
        arrays.update({k: getattr(snap, k) for k in PARTICLE_KEYS})
        arrays.update({k: getattr(snap, k) for k in GRID_KEYS if getattr(snap, k) is not None})
        arrays.update({DIAG_PREFIX + k: np.asarray(v) for k, v in snap.diagnostics.items()})
        # through a file object: given a path, numpy appends ".npz" unless the name ends in
        # exactly that (lower case), and the file would not be where the caller asked
        with open(path, "wb") as fh:
            np.savez_compressed(fh, **arrays)
    else:
        h5py = _h5py()
        with h5py.File(path, "w") as f:
            for k in META_KEYS:
                f.attrs[k] = getattr(snap, k)
            for k in PARTICLE_KEYS:
                f.create_dataset(k, data=getattr(snap, k), compression="gzip")
            for k in GRID_KEYS:
                if getattr(snap, k) is not None:
                    f.create_dataset(k, data=getattr(snap, k), compression="gzip")
            if snap.diagnostics:
                g = f.create_group("diagnostics")
                for k, v in snap.diagnostics.items():
                    g.create_dataset(k, data=np.asarray(v))
    return path


def read_snapshot(path) -> Snapshot:
    path = Path(path)
    if not path.is_file():
        raise FileNotFoundError(f"snapshot not found: {path}")
    suffix = path.suffix.lower()
    if suffix == ".npz":
        with np.load(path, allow_pickle=False) as d:
            return Snapshot(
                step=int(d["step"]), time=float(d["time"]), pos=d["pos"], vel=d["vel"],
                mass=d["mass"], gid=d["gid"], dx=float(d["dx"]),
                rho=d["rho"] if "rho" in d.files else None,
                phi=d["phi"] if "phi" in d.files else None,
                diagnostics={k[len(DIAG_PREFIX):]: np.asarray(d[k]) for k in d.files
                             if k.startswith(DIAG_PREFIX)})
    if suffix in (".h5", ".hdf5"):
        h5py = _h5py()
        with h5py.File(path, "r") as f:
            meta = {k: f.attrs[k] for k in META_KEYS}
            arr = {k: f[k][()] for k in PARTICLE_KEYS}
            grids = {k: (f[k][()] if k in f else None) for k in GRID_KEYS}
            diag = ({k: np.asarray(f["diagnostics"][k][()]) for k in f["diagnostics"]}
                    if "diagnostics" in f else {})
        return Snapshot(step=int(meta["step"]), time=float(meta["time"]), dx=float(meta["dx"]),
                        diagnostics=diag, **arr, **grids)
    raise ValueError(f"unsupported snapshot extension {path.suffix!r} (use .npz/.h5/.hdf5)")


def _h5py():
    try:
        import h5py
    except ImportError as e:  # pragma: no cover - depends on the environment
        raise ImportError("HDF5 snapshots need h5py, which is not installed; use .npz") from e
    return h5py


__all__ = ["Snapshot", "snapshot_from_states", "write_snapshot", "read_snapshot"]
