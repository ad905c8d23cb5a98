"""Shared fixtures: seeded meshes and the `gpu` marker (tests that need an MI355X)."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (gfx950) and the built libheatgeo")


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch


MESHES = {
    "icosphere3": lambda: __import__("ordito_amd.creation", fromlist=["x"]).jittered_icosphere(3, 0.01, 1),
    "icosphere4": lambda: __import__("ordito_amd.creation", fromlist=["x"]).jittered_icosphere(4, 0.01, 0),
    "torus": lambda: __import__("ordito_amd.creation", fromlist=["x"]).torus(48, 24, 1.0, 0.3, 0.1, 3),
    "grid": lambda: __import__("ordito_amd.creation", fromlist=["x"]).planar_grid(30, 24, 1.0, 0.2, 4),
    "bumpy": lambda: __import__("tests.meshes", fromlist=["x"]).bumpy_hull(),
    "saddle_graded": lambda: __import__("tests.meshes", fromlist=["x"]).graded_saddle(),
    "two_bodies": lambda: _two_bodies(),
    "hub_disk": lambda: __import__("tests.meshes", fromlist=["x"]).hub_disk(),
    "pinched_grid": lambda: __import__("tests.meshes", fromlist=["x"]).pinched_grid(),
}


def _two_bodies() -> tuple[np.ndarray, np.ndarray]:
    """One mesh of two connected components: a jittered icosphere and a torus beside it."""
    from ordito_amd import creation

    v0, f0 = creation.jittered_icosphere(3, 0.01, 2)
    v1, f1 = creation.torus(24, 12, 1.0, 0.3, 0.1, 5)
    v1 = v1 + np.array([3.0, 0.0, 0.0], dtype=v1.dtype)
    faces = np.concatenate([np.asarray(f0).reshape(-1, 3), np.asarray(f1).reshape(-1, 3) + len(v0)])
    return np.concatenate([v0, v1]).astype(np.float32), faces.astype(np.int32)


def mesh(name: str) -> tuple[np.ndarray, np.ndarray]:
    return MESHES[name]()
