"""
``Trimesh``: the per-mesh cache that keeps the heat method's solvers and factorizations.

Mirrors the part of ``ordito.mesh.Trimesh`` the heat path reaches (``mesh.py:2187-2265``
``heat_solver``, ``mesh.py:3457-3495`` ``release_factorizations``, ``mesh.py:1045`` ``bounds``,
``mesh.py:1179`` ``mean_edge_length``): the mesh owns its vertices and faces and one
``HeatSolver`` per ``(t, use_robust, method)``, whose factorizations stay resident on the GPU
until released.
"""

from __future__ import annotations

from typing import Any

import numpy as np

from . import _arrays


class Trimesh:
    """A triangle mesh on the GPU: ``(V, 3)`` float32 vertices and a flat ``(3F,)`` int32 face buffer."""

    def __init__(self, vertices: Any, faces: Any) -> None:
        self._template = vertices
        self.vertices = _arrays.vertices_tensor(vertices)
        self.faces = _arrays.faces_tensor(faces)
        self._solvers: dict[tuple[float | None, bool, str, str], Any] = {}

    @property
    def n_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.faces.numel() // 3)

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """Axis-aligned bounding box ``(min, max)``; ``(+inf, -inf)`` for an empty mesh."""
        if self.n_vertices == 0:
            return np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
        v = self.vertices
        return v.min(dim=0).values.cpu().numpy(), v.max(dim=0).values.cpu().numpy()

    @property
    def mean_edge_length(self) -> float:
        """Mean over every face's three edges (an interior edge twice); ``0.0`` with no faces."""
        if self.n_faces == 0:
            return 0.0
        tri = self.faces.reshape(-1, 3).long()
        p = self.vertices
        total = 0.0
        for a, b in ((0, 1), (1, 2), (2, 0)):
            d = p[tri[:, b]] - p[tri[:, a]]
            total += float(d.norm(dim=1).double().sum())
        return total / (3 * self.n_faces)

    def heat_solver(
        self,
        t: float | None = None,
        *,
        use_robust: bool = False,
        method: str = "potpourri3d",
        precision: str = "fp64",
    ):
        """
        The mesh's ``HeatSolver`` at ``t`` (made on first request and kept): ordito's
        ``Trimesh.heat_solver`` (``mesh.py:2187-2265``). ``precision`` ("fp64" or "mixed") is this
        build's addition: the factored solves' solve-block precision.
        """
        from .heat import HeatSolver

        key = (None if t is None else float(t), bool(use_robust), method, precision)
        solver = self._solvers.get(key)
        if solver is None:
            solver = self._solvers[key] = HeatSolver(
                self, t, use_robust=use_robust, method=method, precision=precision
            )
        return solver

    def release_factorizations(self) -> None:
        """Drop every solver's factorizations (``mesh.py:3457-3495``)."""
        for solver in self._solvers.values():
            solver.release()

    @property
    def factorization_nbytes(self) -> int:
        return sum(solver.nbytes for solver in self._solvers.values())
