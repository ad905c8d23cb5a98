"""
Build the oracle's C restatement (``heat_oracle.c``) into ``oracle/liboracle.so`` with gcc.

Test infrastructure: only the tests and the CPU-baseline leg of ``bench.py`` load it.
"""

from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "liboracle.so"


def build(verbose: bool = False) -> Path | None:
    sources = sorted(HERE.glob("*.c"))
    if not sources:
        return None
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-o", str(LIBRARY), *map(str, sources), "-lm"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{result.stdout}\n{result.stderr}")
    if verbose:
        print(f"built {LIBRARY}")
    return LIBRARY


if __name__ == "__main__":
    build(verbose=True)
