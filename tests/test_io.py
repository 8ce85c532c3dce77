"""Snapshot I/O (reference io.py; tests mirror reference tests/test_io.py:16-110).

npz round-trips a Snapshot losslessly (values and dtypes of particles, optional grids,
diagnostics always read back as ndarrays, scalar metadata) and uses the reference's key
layout (`diag__` prefix), so files are interchangeable. HDF5 is exercised only where h5py
is installed. CPU only.
"""

from __future__ import annotations

import numpy as np
import pytest

from galaxy_collision_amd.io import Snapshot, read_snapshot, write_snapshot

try:
    import h5py  # noqa: F401

    EXTS = [".npz", ".h5", ".hdf5"]
except ImportError:
    EXTS = [".npz"]


def _make_snapshot(rng, n=200, g=8, with_grid=True, with_diag=True):
    return Snapshot(
        step=12, time=0.12,
        pos=rng.uniform(0, g, size=(n, 3)).astype(np.float32),
        vel=rng.normal(size=(n, 3)).astype(np.float32),
        mass=rng.uniform(1.0, 2.0, size=n).astype(np.float32),
        gid=rng.integers(0, 2, size=n).astype(np.int32),
        dx=1.5,
        rho=rng.random((g, g, g)).astype(np.float32) if with_grid else None,
        phi=rng.random((g, g, g)).astype(np.float32) if with_grid else None,
        diagnostics=({"energy_total": np.float64(-1.23e10), "momentum": np.array([0.1, -0.2, 0.3]),
                      "ang_momentum": np.array([1.0, 2.0, 3.0])} if with_diag else {}),
    )


def _assert_equal(a: Snapshot, b: Snapshot):
    assert a.step == b.step and a.time == pytest.approx(b.time) and a.dx == pytest.approx(b.dx)
    for key in ("pos", "vel", "mass", "gid"):
        va, vb = getattr(a, key), getattr(b, key)
        np.testing.assert_array_equal(va, vb)
        assert va.dtype == vb.dtype, f"{key} dtype drift: {va.dtype} -> {vb.dtype}"
    for key in ("rho", "phi"):
        ga, gb = getattr(a, key), getattr(b, key)
        assert (ga is None) == (gb is None)
        if ga is not None:
            np.testing.assert_array_equal(ga, gb)
            assert ga.dtype == gb.dtype
    assert set(a.diagnostics) == set(b.diagnostics)
    for key in a.diagnostics:
        assert isinstance(b.diagnostics[key], np.ndarray)
        np.testing.assert_allclose(a.diagnostics[key], b.diagnostics[key])
        assert np.asarray(a.diagnostics[key]).dtype == b.diagnostics[key].dtype


@pytest.mark.parametrize("ext", EXTS)
def test_round_trip_full(tmp_path, ext):
    snap = _make_snapshot(np.random.default_rng(0))
    _assert_equal(snap, read_snapshot(write_snapshot(snap, tmp_path / f"s{ext}")))


@pytest.mark.parametrize("ext", EXTS)
def test_round_trip_minimal_no_grid_no_diag(tmp_path, ext):
    snap = _make_snapshot(np.random.default_rng(1), with_grid=False, with_diag=False)
    back = read_snapshot(write_snapshot(snap, tmp_path / f"m{ext}"))
    _assert_equal(snap, back)
    assert back.rho is None and back.phi is None and back.diagnostics == {}


def test_npz_key_layout_matches_reference_format(tmp_path):
    """Arrays under their own names, scalars as 0-d arrays, diagnostics as diag__<name>
    (reference io.py:157-175)."""
    path = write_snapshot(_make_snapshot(np.random.default_rng(2)), tmp_path / "k.npz")
    with np.load(path, allow_pickle=False) as d:
        assert set(d.files) == {"step", "time", "dx", "pos", "vel", "mass", "gid", "rho", "phi",
                                "diag__energy_total", "diag__momentum", "diag__ang_momentum"}
        assert d["step"].shape == () and int(d["step"]) == 12
        assert d["pos"].shape == (200, 3) and d["gid"].dtype == np.int32


def test_write_creates_parent_dirs(tmp_path):
    path = write_snapshot(_make_snapshot(np.random.default_rng(3)), tmp_path / "a" / "b" / "s.npz")
    assert path.is_file()


def test_unsupported_extension(tmp_path):
    with pytest.raises(ValueError, match="unsupported snapshot extension"):
        write_snapshot(_make_snapshot(np.random.default_rng(4)), tmp_path / "s.txt")
    (tmp_path / "x.txt").write_text("x")
    with pytest.raises(ValueError, match="unsupported snapshot extension"):
        read_snapshot(tmp_path / "x.txt")


def test_read_missing_file(tmp_path):
    with pytest.raises(FileNotFoundError):
        read_snapshot(tmp_path / "nope.npz")


def test_sim_config_validation():
    from galaxy_collision_amd.sim import SimConfig

    SimConfig(solver="multigrid", ic_preset="two_galaxy")
    for bad in ({"solver": "cg"}, {"ic_preset": "hello"}, {"dt": 0.0}, {"steps": -1},
                {"n_particles": 0}, {"output_cadence": -2}):
        with pytest.raises(ValueError):
            SimConfig(**bad)
