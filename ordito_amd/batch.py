"""
Many meshes packed into one block-diagonal system (``hg_mesh_create_packed``).

ordito solves one mesh per call (``heat.py:370-609``). Here a ``MeshBatch`` binds many meshes of
uneven size as one system: one analysis (one nested-dissection tree per mesh, every mesh's leaves
on level 0), one factorization whose tile launches cover every mesh's fronts of a level, and
queries answered together -- each mesh's queries fill columns 0, 1, ... of one batched refined
solve, so a level sweep is one launch over every mesh. Each mesh keeps what it would have alone:
its diffusion time rule (``heat.py:3714-3736``), its components and pins, its shift. Solves go
through the factors (the batch factors on first use).
"""

from __future__ import annotations

import ctypes
import math
import weakref
from typing import Any, Sequence

import numpy as np

from . import _arrays, _lib
from .heat import _METHOD_CODES, _require_precision


def _destroy(ptr: int | None) -> None:
    if ptr:
        try:
            _lib.load().hg_mesh_destroy(ctypes.c_void_p(ptr))
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


class MeshBatch:
    """
    ``meshes``: a sequence of ``(vertices, faces)`` pairs, each mesh's faces indexing its own
    vertices. ``precision`` is the factored solves' ("fp64" or "mixed").
    """

    def __init__(self, meshes: Sequence[tuple[Any, Any]], *, precision: str = "fp64",
                 method: str = "meshlab") -> None:
        if len(meshes) == 0:
            raise ValueError("MeshBatch needs at least one mesh")
        if method not in _METHOD_CODES:
            raise ValueError(f"method must be 'meshlab' or 'potpourri3d', got {method!r}")
        _arrays.require_gpu()
        torch = _arrays.torch_module()
        self.precision = precision
        self._precision_code = _require_precision(precision)
        verts = [_arrays.vertices_tensor(v) for v, _ in meshes]
        faces = [_arrays.faces_tensor(f) for _, f in meshes]
        self.sizes = [int(v.shape[0]) for v in verts]
        self.face_counts = [int(f.numel() // 3) for f in faces]
        self.vertex_offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.face_offsets = np.concatenate([[0], np.cumsum(self.face_counts)]).astype(np.int64)
        packed_v = torch.cat(verts) if self.vertex_offsets[-1] > 0 else torch.zeros((1, 3), dtype=torch.float32, device="cuda")
        packed_f = torch.cat(faces) if self.face_offsets[-1] > 0 else torch.zeros(3, dtype=torch.int32, device="cuda")
        handle = ctypes.c_void_p()
        _lib.check(
            _lib.load().hg_mesh_create_packed(
                _arrays.pointer(packed_v),
                self.vertex_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                _arrays.pointer(packed_f),
                self.face_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                len(meshes),
                _arrays.current_stream(),
                ctypes.byref(handle),
            )
        )
        self._ptr = handle
        self._finalizer = weakref.finalize(self, _destroy, handle.value)
        _lib.call("hg_operators", handle, math.nan, _METHOD_CODES[method], 1)
        self.last_stats: dict[str, int] = {}

    def __len__(self) -> int:
        return len(self.sizes)

    @property
    def times(self) -> np.ndarray:
        """Each mesh's diffusion time."""
        out = np.zeros(len(self.sizes))
        _lib.call("hg_packed_times", self._ptr, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def factor(self) -> bool:
        """Factor both systems of every mesh now (one set of launches for all of them)."""
        _lib.call("hg_mesh_set_stream", self._ptr, _arrays.current_stream())
        out = ctypes.c_int32()
        _lib.call(
            "hg_factor",
            self._ptr,
            _lib.HG_SYSTEM_HEAT | _lib.HG_SYSTEM_POISSON,
            self._precision_code,
            ctypes.byref(out),
        )
        return bool(out.value)

    def set_refinement(self, heat: float | None = None, poisson: float | None = None) -> None:
        """Refinement targets of the batch's factored solves (see ``HeatSolver.set_refinement``)."""
        _lib.call(
            "hg_set_refinement",
            self._ptr,
            self._precision_code,
            math.nan if heat is None else float(heat),
            math.nan if poisson is None else float(poisson),
        )

    @property
    def nbytes(self) -> int:
        out = ctypes.c_int64()
        _lib.call("hg_nbytes", self._ptr, ctypes.byref(out))
        return int(out.value)

    @property
    def device_bytes(self) -> int:
        """Every device byte the batch's state holds: operators, plan, factors and workspaces."""
        out = ctypes.c_int64()
        _lib.call("hg_device_bytes", self._ptr, ctypes.byref(out))
        return int(out.value)

    def release(self) -> None:
        _lib.call("hg_release", self._ptr)

    def geodesic(self, queries: Sequence[tuple[int, Any]]) -> list:
        """
        Distance fields for ``queries``, each ``(mesh index, sources)`` with the sources indexing
        that mesh's vertices; returns one float64 GPU tensor of that mesh's size per query. The
        whole batch is one library call.
        """
        torch = _arrays.torch_module()
        _lib.call("hg_mesh_set_stream", self._ptr, _arrays.current_stream())
        n_q = len(queries)
        if n_q == 0:
            return []
        owners = np.array([int(m) for m, _ in queries], dtype=np.int32)
        if np.any((owners < 0) | (owners >= len(self.sizes))):
            raise ValueError(f"query mesh index outside [0, {len(self.sizes)})")
        tensors = [_arrays.sources_tensor(s) for _, s in queries]
        offsets = [0]
        for s in tensors:
            offsets.append(offsets[-1] + int(s.numel()))
        flat = torch.cat(tensors) if offsets[-1] > 0 else torch.zeros(1, dtype=torch.int32, device="cuda")
        lengths = [self.sizes[m] for m in owners]
        out = torch.zeros(max(sum(lengths), 1), dtype=torch.float64, device="cuda")
        stats = _lib.SolveStats()
        _lib.call(
            "hg_geodesic_packed",
            self._ptr,
            _arrays.pointer(flat),
            _lib.offsets_array(offsets),
            owners.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            n_q,
            _arrays.pointer(out),
            self._precision_code,
            ctypes.byref(stats),
        )
        self.last_stats = stats.as_dict()
        fields = []
        at = 0
        for length in lengths:
            fields.append(out[at : at + length])
            at += length
        return fields

    def close(self) -> None:
        self._finalizer()
