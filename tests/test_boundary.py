"""
The drop-in boundary on the CPU: the C ABI library loads, exports every function
``include/heatgeo.h`` declares, and the Python mirror of ``ordito.heat`` raises the reference's
exceptions on bad arguments before any device work (``heat.py:335-377``, ``3732-3808``).
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared_functions() -> list[str]:
    header = (ROOT / "include" / "heatgeo.h").read_text()
    body = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    return sorted(set(re.findall(r"\b(hg_[a-z_0-9]+)\s*\(", body)))


def test_library_exports_every_declared_function():
    from ordito_amd import _lib

    lib = _lib.load()
    declared = _declared_functions()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.EXPORTED) == declared


def test_version_and_empty_error():
    from ordito_amd import _lib

    lib = _lib.load()
    assert lib.hg_version().decode().startswith("heatgeo")


def test_null_handles_are_runtime_errors():
    from ordito_amd import _lib

    lib = _lib.load()
    out = ctypes.c_int64()
    assert lib.hg_nbytes(None, ctypes.byref(out)) == _lib.HG_ERR_RUNTIME
    assert "null" in lib.hg_last_error().decode()
    with pytest.raises(RuntimeError):
        _lib.call("hg_release", None)


def test_unknown_method_is_a_value_error():
    from ordito_amd import heat

    with pytest.raises(ValueError):
        heat.heat_geodesic(np.zeros((3, 3), np.float32), np.array([0, 1, 2]), [0], method="nope")
    with pytest.raises(ValueError):
        heat.HeatSolver.__init__(object.__new__(heat.HeatSolver), None, method="nope")


def test_cot_entries_and_use_robust_are_exclusive():
    from ordito_amd import heat

    with pytest.raises(ValueError):
        heat.heat_operators(
            np.zeros((3, 3), np.float32), np.array([0, 1, 2]), cot_entries=np.zeros((1, 3)),
            use_robust=True,
        )


def test_unknown_precision_is_a_value_error():
    from ordito_amd import heat

    with pytest.raises(ValueError):
        heat.heat_geodesic(np.zeros((3, 3), np.float32), np.array([0, 1, 2]), [0], precision="fp16")


def test_compute_without_a_gpu_fails_loudly():
    """No CPU fallback: the product path refuses to run where libheatgeo has no GPU."""
    import torch

    from ordito_amd import heat

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no CPU fallback|GPU"):
        heat.heat_geodesic(np.eye(3, dtype=np.float32), np.array([0, 1, 2]), [0])


def test_generators_are_closed_or_rimmed_as_documented():
    from ordito_amd import creation

    for k in (0, 1, 3):
        v, f = creation.icosphere(k)
        assert v.shape == (10 * 4**k + 2, 3) and f.size == 3 * 20 * 4**k
        assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-6)
        assert np.allclose(v[0], [-0.5257311, 0.8506508, 0.0], atol=1e-6)
    v, f = creation.torus(16, 8)
    tri = f.reshape(-1, 3)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    _, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)  # closed
    assert v.shape[0] - len(counts) + tri.shape[0] == 0  # Euler characteristic of a torus
    v, f = creation.planar_grid(5, 4)
    assert v.shape == (20, 3) and f.size == 3 * 2 * 4 * 3
