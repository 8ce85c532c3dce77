"""Shared pytest setup: import paths and the ``gpu`` marker.

``-m gpu`` tests need a HIP device and the built libpmchain.so; ``-m "not gpu"`` tests run
on the CPU (the oracle against its closed forms, host logic, and that the library loads
and exports every declared entry point).
"""

from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG_DIR = ROOT / "galaxy-collision-with-cuda-fft_amd"
for p in (str(PKG_DIR), str(ROOT / "oracle"), str(ROOT)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device and the built libpmchain.so")


@pytest.fixture(scope="session")
def device():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)


def announce(msg: str) -> None:
    """Print and flush a line naming the case about to run on the GPU."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()


def assert_bitwise(got, exp, what: str = "") -> None:
    """Bitwise equality with a short report (count and first differing index) on failure."""
    import numpy as np

    got, exp = np.asarray(got), np.asarray(exp)
    if got.shape != exp.shape:
        raise AssertionError(f"{what}: shape {got.shape} != {exp.shape}")
    if got.dtype.kind == "f":
        same = (got == exp) | (np.isnan(got) & np.isnan(exp))
    else:
        same = got == exp
    if not same.all():
        bad = np.argwhere(~same)
        idx = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} of {got.size} values differ; first at {idx}: "
                             f"{got[idx]!r} vs {exp[idx]!r}")
