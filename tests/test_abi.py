"""The C ABI library loads and exports every entry point include/pmchain.h declares.

CPU-only: loads libpmchain.so with ctypes (no kernel is launched).
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "pmchain.h"
LIB = ROOT / "galaxy-collision-with-cuda-fft_amd" / "galaxy_collision_amd" / "libpmchain.so"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "pm_deposit_cic" in names and "pm_mg_solve" in names
    assert len(names) >= 25


@pytest.mark.skipif(not LIB.exists(), reason="libpmchain.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="libpmchain.so not built")
def test_python_binding_covers_header():
    from galaxy_collision_amd import _native as nat

    assert set(declared_functions()) <= set(nat.exported_names())
    assert nat.lib().pm_abi_version() == 2


@pytest.mark.skipif(not LIB.exists(), reason="libpmchain.so not built")
def test_error_path_without_device_work():
    """Argument validation fails before any HIP call: a null plan is rejected with a message."""
    from galaxy_collision_amd import _native as nat

    with pytest.raises(nat.PMChainError, match="null plan"):
        nat.call("pm_deposit_cic", None, None, None, None, None, 0, 1.0, None, None, None, None)
    with pytest.raises(nat.PMChainError, match="grid_size"):
        out = ctypes.c_void_p()
        nat.call("pm_deposit_plan_create", 0, 10, ctypes.byref(out))
    # the reordering drift is out of place: a destination overlapping a source is refused
    arr = nat.ParticleArrays(*[0x10000000 + k * 0x100000 for k in range(9)])
    with pytest.raises(nat.PMChainError, match="overlaps a source"):
        nat.call("pm_gather_kick_drift4_reorder", 0x20000000, ctypes.byref(arr),
                 ctypes.byref(arr), 0x30000000, 1000, 16, 1.0, 0.1, 0.1, 0.1, 0x40000000, None)
    with pytest.raises(nat.PMChainError, match="mode must be"):
        nat.call("pm_timing_enable", 3)
    nat.call("pm_timing_enable", 0)
