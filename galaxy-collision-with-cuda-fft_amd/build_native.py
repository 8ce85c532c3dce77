"""Build libpmchain.so (the HIP kernels + C ABI) for gfx950 with hipcc.

Usage: ``python build_native.py [--force] [--jobs N]``; ``build()`` from Python.

Flags that matter for correctness:
- ``-ffp-contract=off``: no implicit FMA contraction, so every kernel keeps one rounding
  per operation and matches its numpy float32 restatement bit for bit;
- ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""

from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
BUILD = HERE / "_build"
LIB = HERE / "galaxy_collision_amd" / "libpmchain.so"
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-I", str(INCLUDE),
    "-I", str(CSRC),
]


# Per-source additions. The bricked deposit's unit pass claims its next unit with one
# lane's atomic whose result is only read at the unit's end; the atomic optimizer would
# broadcast that result at once (a wait on every outstanding load, prefetches included).
FILE_FLAGS = {
    "pm_deposit_bricked.hip": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"],
}


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found (ROCm is required to build libpmchain.so)")
    return exe


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [_hipcc(), *COMMON_FLAGS, *FILE_FLAGS.get(src.name, []), "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{proc.stdout}\n{proc.stderr}")
    if verbose and proc.stderr.strip():
        print(proc.stderr, file=sys.stderr)
    return obj


def build(force: bool = True, jobs: int = 8, verbose: bool = False) -> Path:
    """Compile every ``csrc/*.hip`` and link ``libpmchain.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    if not force and LIB.exists():
        newest = max(p.stat().st_mtime for p in [*sources, *CSRC.glob("*.h"), INCLUDE / "pmchain.h"])
        if LIB.stat().st_mtime >= newest:
            return LIB
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as pool:
        objs = list(pool.map(lambda s: _compile(s, verbose), sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--force", action="store_true", help="rebuild even if up to date")
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    args = ap.parse_args()
    path = build(force=args.force, jobs=args.jobs, verbose=args.verbose)
    print(path)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
