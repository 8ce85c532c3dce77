"""The accuracy diagnostics' exact discrete solve against the oracle's DST-I (scipy). CPU only:
the torch transform runs on the host here; the same code runs on the device in the GPU tests.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import pm_oracle as O
from galaxy_collision_amd.accuracy import dst_exact, dst_exact_device


@pytest.mark.parametrize("n,dx", [(3, 1.0), (6, 0.5), (17, 0.3125), (40, 2.0)])
def test_device_dst_solve_matches_oracle(n, dx):
    rng = np.random.default_rng(n)
    f = rng.normal(size=(n, n, n)).astype(np.float32)
    b = rng.normal(size=(n, n, n)).astype(np.float32)
    ref = O.dirichlet_solve_f64(f, b, dx)
    got = dst_exact_device(torch.from_numpy(f), torch.from_numpy(b), dx).numpy()
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.abs(dst_exact(f, b, dx) - ref).max() <= 1e-12 * np.abs(ref).max()
    # the faces are the boundary values, unchanged
    assert np.array_equal(got[0], b[0].astype(np.float64))
    assert np.array_equal(got[:, :, -1], b[:, :, -1].astype(np.float64))


def test_device_dst_solve_is_a_solution():
    """7-point residual of the result vanishes to float64 round-off."""
    n, dx = 24, 0.7
    rng = np.random.default_rng(3)
    f = rng.normal(size=(n, n, n))
    b = rng.normal(size=(n, n, n))
    u = dst_exact_device(torch.from_numpy(f), torch.from_numpy(b), dx).numpy()
    lap = (u[2:, 1:-1, 1:-1] + u[:-2, 1:-1, 1:-1] + u[1:-1, 2:, 1:-1] + u[1:-1, :-2, 1:-1]
           + u[1:-1, 1:-1, 2:] + u[1:-1, 1:-1, :-2] - 6.0 * u[1:-1, 1:-1, 1:-1]) / dx ** 2
    assert np.abs(lap - f[1:-1, 1:-1, 1:-1]).max() <= 1e-10 * np.abs(f).max()
