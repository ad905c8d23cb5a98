"""
Write the golden fixtures: FP64 oracle distance fields on small seeded meshes.

The reference (ordito) cannot be imported here (warp-lang absent, Python 3.10 < 3.11) and holds no
numeric golden files for this path, so these fixtures freeze the oracle's own outputs after the
oracle is pinned (tests/test_oracle.py): they let the GPU tests compare against stored vectors, and
the CPU tests catch any drift in the oracle. Run: python tests/golden/make_golden.py
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import heat_oracle as ho  # noqa: E402
from ordito_amd import creation  # noqa: E402
from tests import meshes  # noqa: E402

CASES = {
    "icosphere3_jitter": (lambda: creation.jittered_icosphere(3, 0.01, 1), [[0], [100, 200]]),
    "torus48x24": (lambda: creation.torus(48, 24, 1.0, 0.3, 0.1, 3), [[5], [0, 600]]),
    "grid30x24": (lambda: creation.planar_grid(30, 24, 1.0, 0.2, 4), [[0], [371]]),
    "bumpy500": (lambda: meshes.bumpy_hull(), [[0], [3, 77, 250]]),
}


def main() -> None:
    for name, (make, source_sets) in CASES.items():
        vertices, faces = make()
        distances = np.stack([ho.geodesic(vertices, faces, s).distance for s in source_sets])
        lengths = np.array([len(s) for s in source_sets], dtype=np.int32)
        flat = np.concatenate([np.asarray(s, dtype=np.int32) for s in source_sets])
        np.savez_compressed(
            HERE / f"{name}.npz",
            vertices=vertices,
            faces=faces,
            sources=flat,
            source_lengths=lengths,
            distance=distances,
        )
        print(name, vertices.shape, distances.shape)


if __name__ == "__main__":
    main()
