"""
Build ``libheatgeo.so`` in-tree with hipcc for gfx950.

Every ``.hip`` and ``.cpp`` under ``csrc/`` is compiled to an object in ``_build/`` (in parallel)
and linked into ``ordito_amd/libheatgeo.so``. Runs on a machine without a GPU: hipcc
cross-compiles the device code.
"""

from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
from pathlib import Path

PACKAGE = Path(__file__).resolve().parent
CSRC = PACKAGE / "csrc"
BUILD = PACKAGE.parent / "_build"
LIBRARY = PACKAGE / "libheatgeo.so"
ARCH = os.environ.get("HEATGEO_ARCH", "gfx950")

# Packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have given silently wrong
# results on this hardware. -O3 for device and host.
COMMON_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-but-set-variable",
    "-Wno-unknown-pragmas",
]


def hipcc() -> str:
    candidate = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return candidate


def _compile(source: Path) -> Path:
    obj = BUILD / (source.name + ".o")
    cmd = [hipcc(), *COMMON_FLAGS, "-c", str(source), "-o", str(obj)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"hipcc failed on {source.name}:\n{result.stdout}\n{result.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every source and link ``libheatgeo.so``; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted([*CSRC.glob("*.hip"), *CSRC.glob("*.cpp")])
    workers = min(len(sources), max(1, (os.cpu_count() or 2)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIBRARY.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objects)]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"link failed:\n{result.stdout}\n{result.stderr}")
    os.replace(tmp, LIBRARY)
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
