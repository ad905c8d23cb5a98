"""
The device-built operator pattern (``mesh_operator_pattern``, ``laplacian.py:1071-1389``) against
the oracle's sorted-key restatement, entry for entry: row bounds, columns, each entry's corner run
in ascending slot order, and each vertex's ascending corners (the stable sort of the face buffer,
``heat.py:3170-3190``). Bit-exact: the pattern is integer work.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import heat_oracle as ho
from tests.conftest import MESHES, mesh

pytestmark = pytest.mark.gpu


def _expected_vertex_corners(n: int, faces: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    flat = np.asarray(faces, dtype=np.int64).reshape(-1)
    order = np.argsort(flat, kind="stable")
    ptr = np.concatenate([[0], np.cumsum(np.bincount(flat, minlength=n))])
    return ptr, order


def _check(n: int, faces: np.ndarray) -> None:
    from ordito_amd.laplacian import mesh_operator_pattern

    got = mesh_operator_pattern(np.asarray(faces, dtype=np.int32).reshape(-1), n)
    want = ho.operator_pattern(n, faces)
    assert got.n_vertices == n and got.operator == "cotmatrix"
    assert np.array_equal(got.offsets, want.indptr)
    assert np.array_equal(got.columns, want.indices)
    assert np.array_equal(got.corner_ptr, want.corner_ptr)
    assert np.array_equal(got.corner_slots, want.corner_slots)
    ptr, corners = _expected_vertex_corners(n, faces)
    assert np.array_equal(got.vertex_ptr, ptr)
    assert np.array_equal(got.vertex_corners, corners)


@pytest.mark.parametrize("name", sorted(MESHES))
def test_pattern_matches_oracle(cuda, name):
    v, f = mesh(name)
    _check(len(v), np.asarray(f).reshape(-1, 3))


def test_pattern_unreferenced_and_repeated_edges(cuda):
    # Two triangles sharing an edge, a vertex no face references (no diagonal), a fan with a
    # non-manifold edge shared by three faces (a run of three corners).
    faces = np.array([[0, 1, 2], [2, 1, 3], [1, 2, 5], [5, 6, 0]], dtype=np.int32)
    _check(8, faces)


def test_pattern_no_faces(cuda):
    _check(5, np.zeros((0, 3), dtype=np.int32))


def test_pattern_at_full_size(cuda):
    # 1024 x 512 torus: 524,288 vertices, 1,048,576 faces, 6.8 M sorted keys.
    from ordito_amd import creation

    v, f = creation.torus(1024, 512, 1.0, 0.3, 0.0, 0)
    _check(len(v), np.asarray(f).reshape(-1, 3))


def test_pattern_rejects_bad_index(cuda):
    from ordito_amd.laplacian import mesh_operator_pattern

    with pytest.raises(ValueError, match="vertex index"):
        mesh_operator_pattern(np.array([0, 1, 7], dtype=np.int32), 4)
    with pytest.raises(ValueError, match="vertex index"):
        mesh_operator_pattern(np.array([0, -1, 2], dtype=np.int32), 4)


def test_pattern_rejects_other_operators(cuda):
    from ordito_amd.laplacian import mesh_operator_pattern

    with pytest.raises(ValueError):
        mesh_operator_pattern(np.array([0, 1, 2], dtype=np.int32), 3, operator="laplacian_directed")
    with pytest.raises(ValueError):
        mesh_operator_pattern(np.array([0, 1, 2], dtype=np.int32), 3, operator="nope")


def test_mesh_pattern_is_the_device_pattern(cuda):
    # A bound mesh's operators use the pattern the device built.
    from ordito_amd import heat

    v, f = mesh("hub_disk")
    ops = heat.heat_operators(cuda.from_numpy(v).cuda(), cuda.from_numpy(np.asarray(f).reshape(-1)).cuda())
    want = ho.operator_pattern(len(v), np.asarray(f).reshape(-1, 3))
    assert np.array_equal(ops[2].offsets.cpu().numpy(), want.indptr)
    assert np.array_equal(ops[2].columns.cpu().numpy(), want.indices)


def test_pattern_from_host_buffers(cuda):
    # The C ABI takes the faces and the outputs in host memory too (copied through the library).
    import ctypes

    from ordito_amd import _lib

    v, f = mesh("grid")
    faces = np.ascontiguousarray(np.asarray(f, dtype=np.int32).reshape(-1))
    n, n_faces = len(v), faces.size // 3
    cap = 6 * n_faces + n
    offsets = np.zeros(n + 1, dtype=np.int32)
    columns = np.zeros(cap, dtype=np.int32)
    run_ptr = np.zeros(cap + 1, dtype=np.int32)
    slots = np.zeros(6 * n_faces, dtype=np.int32)
    sizes = (ctypes.c_int64 * 2)()
    ptr = lambda a: a.ctypes.data  # noqa: E731
    _lib.call("hg_operator_pattern", ptr(faces), n_faces, n, None, ptr(offsets), ptr(columns),
              ptr(run_ptr), ptr(slots), None, None, sizes)
    want = ho.operator_pattern(n, faces.reshape(-1, 3))
    nnz = int(sizes[0])
    assert nnz == want.indices.size and int(sizes[1]) == want.corner_slots.size
    assert np.array_equal(offsets, want.indptr)
    assert np.array_equal(columns[:nnz], want.indices)
    assert np.array_equal(run_ptr[: nnz + 1], want.corner_ptr)
    assert np.array_equal(slots, want.corner_slots)


def test_pattern_degenerate_faces(cuda):
    # A face repeating a vertex: its self-edge lands on that vertex's diagonal entry with the
    # corner slots it carries, and its two other edges both ways, as the sorted keys give them.
    faces = np.array([[0, 1, 2], [2, 2, 3], [3, 1, 1], [0, 4, 0]], dtype=np.int32)
    _check(5, faces)
