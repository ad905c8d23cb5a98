"""
The cotangent operator's sparsity pattern, built on the GPU (ordito ``laplacian.py:1022-1141``).

``mesh_operator_pattern(faces, n_vertices)`` returns the pattern ``cotmatrix`` and the heat
method's operators share: every face edge both ways plus a diagonal per referenced vertex, rows
sorted by column, and each entry's run of contributing corner slots ``3f + e`` (corner ``e``
opposite the edge) in ascending order, the order its half-cotangents are summed in. The runs take
the place of the reference's ``run_start`` / ``keys`` / ``count`` / ``order`` arrays, which spell
the same runs through the sorted keys. Built by ``hg_operator_pattern`` (radix sorts on the
device); a face index outside ``[0, n_vertices)`` is a ``ValueError``.
"""

from __future__ import annotations

import ctypes
from typing import Any, NamedTuple

from . import _arrays, _lib

_OPERATORS = ("cotmatrix", "laplacian_symmetric", "laplacian_directed")


class MeshOperatorPattern(NamedTuple):
    """The sparsity of the cotangent operator on one face buffer (``laplacian.py:1022-1068``)."""

    operator: str
    n_vertices: int
    offsets: Any  # (V+1,) int32 row bounds
    columns: Any  # (nnz,) int32 sorted columns per row
    corner_ptr: Any  # (nnz+1,) int32: entry e's corner slots are corner_slots[ptr[e]:ptr[e+1]]
    corner_slots: Any  # int32 corner slots 3f+e, ascending per entry; none for a diagonal
    vertex_ptr: Any  # (V+1,) int32: vertex v's corners are vertex_corners[ptr[v]:ptr[v+1]]
    vertex_corners: Any  # (3F,) int32 ascending per vertex (heat.py:3170-3190)


def mesh_operator_pattern(
    faces: Any, n_vertices: int, *, operator: str = "cotmatrix"
) -> MeshOperatorPattern:
    """
    Sparsity of the cotangent operator on ``faces`` (``laplacian.py:1071-1141``).

    ``faces`` is the ``(3 * n_faces,)`` (or ``(n_faces, 3)``) triangle index buffer; the arrays
    come back as int32 device tensors, or numpy arrays when ``faces`` is one. Only the
    ``"cotmatrix"`` pattern is served; the reference's other names are a ``ValueError`` here, as an
    unknown name is there.
    """
    if operator not in _OPERATORS:
        raise ValueError(f"operator must be one of {list(_OPERATORS)}, got {operator!r}")
    if operator != "cotmatrix":
        raise ValueError(
            f"operator={operator!r} serves ordito.laplacian.laplacian, outside the heat path; "
            "only 'cotmatrix' is built"
        )
    n_vertices = int(n_vertices)
    if n_vertices < 0:
        raise ValueError(f"n_vertices must be non-negative, got {n_vertices}")
    torch = _arrays.torch_module()
    f = _arrays.faces_tensor(faces)
    n_faces = f.numel() // 3
    capacity = 6 * n_faces + n_vertices

    def buffer(count: int):
        return torch.empty(max(count, 1), dtype=torch.int32, device="cuda")

    offsets, columns = buffer(n_vertices + 1), buffer(capacity)
    corner_ptr, corner_slots = buffer(capacity + 1), buffer(6 * n_faces)
    vertex_ptr, vertex_corners = buffer(n_vertices + 1), buffer(3 * n_faces)
    sizes = (ctypes.c_int64 * 2)()
    _lib.call(
        "hg_operator_pattern",
        _arrays.pointer(f),
        n_faces,
        n_vertices,
        _arrays.current_stream(),
        offsets.data_ptr(),
        columns.data_ptr(),
        corner_ptr.data_ptr(),
        corner_slots.data_ptr(),
        vertex_ptr.data_ptr(),
        vertex_corners.data_ptr(),
        sizes,
    )
    nnz, n_slots = int(sizes[0]), int(sizes[1])
    arrays = (
        offsets[: n_vertices + 1],
        columns[:nnz],
        corner_ptr[: nnz + 1],
        corner_slots[:n_slots],
        vertex_ptr[: n_vertices + 1],
        vertex_corners[: 3 * n_faces],
    )
    arrays = tuple(_arrays.like_input(a, faces) for a in arrays)
    return MeshOperatorPattern(operator, n_vertices, *arrays)
