"""
The heat-method geodesic path with ``ordito.heat``'s array-in/array-out signatures.

``heat_geodesic`` (ordito ``heat.py:370-609``), ``heat_operators`` (``heat.py:200-367``) and
``HeatSolver`` (``heat.py:3491-4456``) over the MI355X library ``libheatgeo`` (C ABI in
``include/heatgeo.h``). The solver policy is the reference's: a ``HeatSolver`` iterates a system
the first time it solves it (settled Jacobi-CG for the heat step, Jacobi-Chebyshev CG for the
Poisson step), factors it on its second solve (rent or buy, ``heat.py:4394-4404``) and at once when
a settled iterate fails its backward-error check (``linalg.py:2687-2727``), and the Poisson
system is factored whenever the heat system is (``heat.py:4118-4151``). Every solve runs on the GPU;
there is no CPU path.
"""

from __future__ import annotations

import ctypes
import logging
import math
import weakref
from dataclasses import dataclass
from typing import Any, Sequence

import numpy as np

from . import _arrays, _lib
from .mesh import Trimesh

_LOGGER = logging.getLogger(__name__)

_METHODS = ("meshlab", "igl", "potpourri3d")
_METHOD_CODES = {
    "meshlab": _lib.HG_METHOD_MESHLAB,
    "igl": _lib.HG_METHOD_IGL,
    "potpourri3d": _lib.HG_METHOD_POTPOURRI3D,
}
_PRECISIONS = {"fp64": _lib.HG_PRECISION_FP64, "mixed": _lib.HG_PRECISION_MIXED}


def _require_method(method: str, *, geodesic: bool) -> None:
    if method not in _METHODS:
        raise ValueError(f"method must be 'meshlab', 'igl' or 'potpourri3d', got {method!r}")


def _method_code(method: str, use_robust: bool) -> int:
    """The library's method code, with ``use_robust`` (mollified side lengths for the
    half-cotangents, ordito ``heat.py:229-243``) as its flag. At ``"potpourri3d"`` the reference's
    ``use_robust`` runs geometry-central's fully intrinsic method on an intrinsic Delaunay
    triangulation of the tufted cover (``heat.py:425-430``), which this library does not build."""
    if use_robust and method == "potpourri3d":
        raise NotImplementedError(
            "use_robust at method='potpourri3d' (geometry-central's intrinsic Delaunay heat "
            "method) is not implemented; method='meshlab' and 'igl' take use_robust"
        )
    return _METHOD_CODES[method] | (_lib.HG_METHOD_ROBUST if use_robust else 0)


def _require_precision(precision: str) -> int:
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    return _PRECISIONS[precision]


@dataclass
class CsrMatrix:
    """A square CSR matrix on the GPU: ``offsets`` (n+1) int32, ``columns`` and ``values``."""

    offsets: Any
    columns: Any
    values: Any

    @property
    def nrow(self) -> int:
        return int(self.offsets.numel() - 1)

    @property
    def shape(self) -> tuple[int, int]:
        return (self.nrow, self.nrow)

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix(
            (
                self.values.cpu().numpy(),
                self.columns.cpu().numpy(),
                self.offsets.cpu().numpy(),
            ),
            shape=self.shape,
        )


@dataclass
class JacobiPreconditioner:
    """``D^-1`` of a system (``linalg.py:1591-1647``); the CG kernel applies it on the device."""

    inverse_diagonal: Any


@dataclass
class ChebyshevPreconditioner:
    """The degree-12 Jacobi-Chebyshev polynomial of a system over its Gershgorin interval
    (``linalg.py:5135-5232``); the CG kernel applies it on the device."""

    matrix: CsrMatrix
    steps: int = 12


HeatOperators = tuple[
    CsrMatrix, JacobiPreconditioner, CsrMatrix, CsrMatrix, ChebyshevPreconditioner, Any, Any, Any
]


class _Handle:
    """An ``hg_mesh`` of libheatgeo, destroyed with its owner."""

    def __init__(self, vertices, faces) -> None:
        _arrays.require_gpu()
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(
            lib.hg_mesh_create(
                _arrays.pointer(vertices),
                int(vertices.shape[0]),
                _arrays.pointer(faces),
                int(faces.numel() // 3),
                _arrays.current_stream(),
                ctypes.byref(handle),
            )
        )
        self.ptr = handle
        self._finalizer = weakref.finalize(self, _destroy, handle.value)

    def sync_stream(self) -> None:
        _lib.call("hg_mesh_set_stream", self.ptr, _arrays.current_stream())

    def close(self) -> None:
        self._finalizer()


def _destroy(ptr: int | None) -> None:
    if ptr:
        try:
            _lib.load().hg_mesh_destroy(ctypes.c_void_p(ptr))
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


def _views(handle: _Handle) -> _lib.OperatorView:
    view = _lib.OperatorView()
    _lib.call("hg_operator_views", handle.ptr, ctypes.byref(view))
    return view


def _copy_device(pointer: int, count: int, dtype):
    """A torch tensor holding a copy of ``count`` elements of library device memory."""
    torch = _arrays.torch_module()
    out = torch.empty(count, dtype=dtype, device="cuda")
    if count:
        torch.cuda.synchronize()
        _arrays.device_copy(out.data_ptr(), int(pointer), count * out.element_size())
    return out


def _operators_from_views(view: _lib.OperatorView) -> HeatOperators:
    torch = _arrays.torch_module()
    n, f, nnz = int(view.n_vertices), int(view.n_faces), int(view.nnz)
    offsets = _copy_device(view.offsets, n + 1, torch.int32)
    columns = _copy_device(view.columns, nnz, torch.int32)
    laplacian = CsrMatrix(offsets, columns, _copy_device(view.laplacian, nnz, torch.float64))
    heat = CsrMatrix(offsets, columns, _copy_device(view.heat, nnz, torch.float64))
    poisson = CsrMatrix(offsets, columns, _copy_device(view.poisson, nnz, torch.float64))
    diagonal = _diagonal(heat)
    inverse = torch.where(diagonal != 0, 1.0 / torch.where(diagonal != 0, diagonal, 1.0), 0.0)
    cot = _copy_device(view.cot, 3 * f, torch.float32).reshape(f, 3)
    normals = _copy_device(view.normals, 3 * f, torch.float32).reshape(f, 3)
    areas = _copy_device(view.areas, f, torch.float32)
    return (
        heat,
        JacobiPreconditioner(inverse),
        laplacian,
        poisson,
        ChebyshevPreconditioner(poisson),
        cot,
        normals,
        areas,
    )


def _diagonal(matrix: CsrMatrix):
    torch = _arrays.torch_module()
    n = matrix.nrow
    rows = torch.repeat_interleave(
        torch.arange(n, device="cuda"), (matrix.offsets[1:] - matrix.offsets[:-1]).long()
    )
    mask = matrix.columns.long() == rows
    out = torch.zeros(n, dtype=torch.float64, device="cuda")
    out[rows[mask]] = matrix.values[mask]
    return out


def heat_operators(
    vertices: Any,
    faces: Any,
    t: float | None = None,
    *,
    use_robust: bool = False,
    cot_entries: Any = None,
    pattern: Any = None,
    method: str = "potpourri3d",
) -> HeatOperators:
    """
    Assemble the source-independent operators of the heat method (ordito ``heat.py:200-367``).

    Returns ``(heat_system, heat_preconditioner, laplacian, poisson_system,
    poisson_preconditioner, cot_entries, face_normals, face_areas)``: ``M - tL``, its Jacobi
    preconditioner, the float64 cotangent Laplacian ``L`` (negative diagonal), ``-L``, its
    Jacobi-Chebyshev preconditioner, the ``(F, 3)`` float32 half-cotangents, unit normals and
    areas. ``t=None`` is the squared mean edge length of ``method`` (MeshLab: every face's three
    edges; igl and potpourri3d: the unique edges); the lumped mass is barycentric, or Voronoi at
    ``method="igl"``. A ``cot_entries`` table, when given, must have shape
    ``(n_faces, 3)``; the table is formed from the positions either way (what
    ``cotmatrix_entries`` forms), or with ``use_robust`` from the mollified side lengths.
    ``pattern`` is accepted for signature compatibility.
    """
    if cot_entries is not None and use_robust:
        raise ValueError(
            "cot_entries and use_robust are mutually exclusive: use_robust rebuilds the "
            "half-cotangent table from mollified edge lengths."
        )
    _require_method(method, geodesic=False)
    code = _method_code(method, use_robust)
    v = _arrays.vertices_tensor(vertices)
    f = _arrays.faces_tensor(faces)
    if cot_entries is not None:
        shape = tuple(np.shape(cot_entries)) if not _arrays.is_torch(cot_entries) else tuple(
            cot_entries.shape
        )
        if shape != (f.numel() // 3, 3):
            raise ValueError(
                f"cot_entries must have shape (n_faces, 3) = ({f.numel() // 3}, 3), got {shape}"
            )
    handle = _Handle(v, f)
    try:
        _lib.call(
            "hg_operators",
            handle.ptr,
            math.nan if t is None else float(t),
            code,
            0,
        )
        return _operators_from_views(_views(handle))
    finally:
        handle.close()


class HeatSolver:
    """
    The heat method's operators and solves on one mesh at one diffusion time, kept by the mesh
    (ordito ``heat.py:3491-4456``): operators assembled on first use; a system factored on its
    second solve, or at once when the heat step's settled iterate fails its backward-error check;
    the Poisson system factored whenever the heat system is. ``precision`` ("fp64", the reference's,
    or "mixed": FP32 solve blocks with FP64 refinement to componentwise 1e-6 for the heat step and
    normwise 1e-6 for the Poisson step, a served accuracy) selects the factored solves' precision.
    """

    def __init__(
        self,
        mesh: Trimesh,
        t: float | None = None,
        *,
        use_robust: bool = False,
        method: str = "potpourri3d",
        precision: str = "fp64",
    ) -> None:
        _require_method(method, geodesic=False)
        self._method_code = _method_code(method, use_robust)
        self.t = None if t is None else float(t)
        self.use_robust = bool(use_robust)
        self.method = method
        self.precision = precision
        self._precision_code = _require_precision(precision)
        self._mesh = weakref.ref(mesh)
        self._handle: _Handle | None = None
        self._operators: HeatOperators | None = None
        self._solves = {"heat": 0, "poisson": 0, "pinned": 0, "poisson_regularized": 0}
        self._rim: int | None = None
        self._ill_conditioned = False
        self.last_stats: dict[str, int] = {}

    # -- state -----------------------------------------------------------------------------------

    def _bound_mesh(self) -> Trimesh:
        mesh = self._mesh()
        if mesh is None:
            raise RuntimeError("HeatSolver: its Trimesh no longer exists")
        return mesh

    def _native(self) -> _Handle:
        if self._handle is None:
            mesh = self._bound_mesh()
            handle = _Handle(mesh.vertices, mesh.faces)
            _lib.call(
                "hg_operators",
                handle.ptr,
                math.nan if self.t is None else float(self.t),
                self._method_code,
                1,
            )
            self._handle = handle
        self._handle.sync_stream()
        return self._handle

    @property
    def operators(self) -> HeatOperators:
        """``heat_operators`` for the mesh at this solver's time (the default time's floor applied)."""
        if self._operators is None:
            self._operators = _operators_from_views(_views(self._native()))
        return self._operators

    @property
    def time(self) -> float:
        """The diffusion time the systems are built at."""
        return float(_views(self._native()).t)

    @property
    def settle_rounds(self) -> int:
        return int(_views(self._native()).settle_rounds)

    @property
    def mollify_delta(self) -> float:
        """With ``use_robust``, the constant added to every side length (0 on a clean mesh)."""
        return float(_views(self._native()).mollify_delta)

    @property
    def ill_conditioned(self) -> bool:
        """Whether a settled iterate of this mesh's heat system failed its backward-error check."""
        return self._ill_conditioned

    @property
    def nbytes(self) -> int:
        """Device memory the kept factorizations hold, in bytes (``heat.py:3838-3847``)."""
        if self._handle is None:
            return 0
        out = ctypes.c_int64()
        _lib.call("hg_nbytes", self._handle.ptr, ctypes.byref(out))
        return int(out.value)

    @property
    def device_bytes(self) -> int:
        """Every device byte the solver's mesh state holds: operators, plan, factors and the
        per-call workspaces (this build's addition, for admission by memory)."""
        if self._handle is None:
            return 0
        out = ctypes.c_int64()
        _lib.call("hg_device_bytes", self._handle.ptr, ctypes.byref(out))
        return int(out.value)

    def _factored(self) -> int:
        return self._factored_on(self._native())

    @staticmethod
    def _factored_on(native: _Handle) -> int:
        out = ctypes.c_int32()
        _lib.call("hg_factored", native.ptr, ctypes.byref(out))
        return int(out.value)

    def release(self) -> None:
        """Drop every kept factorization and forget the solve counts (``heat.py:3849-3861``)."""
        if self._handle is not None:
            _lib.call("hg_release", self._handle.ptr)
        self._solves = {"heat": 0, "poisson": 0, "pinned": 0, "poisson_regularized": 0}

    @property
    def unfactorable(self) -> bool:
        """Whether a factorization was refused under the memory budget (its solves iterate)."""
        return bool(self._factored() >> 4)

    def set_graphs(self, enable: bool = True) -> None:
        """Replay the factored FP64 geodesic path as a HIP graph per column count (the default),
        or launch it kernel by kernel. (No counterpart in the reference.)"""
        _lib.call("hg_set_graphs", self._native().ptr, int(bool(enable)))

    def graph_state(self) -> tuple[int, bool]:
        """(graphs held, whether a capture failed and the mesh now solves eagerly)."""
        cached, broken = ctypes.c_int64(), ctypes.c_int32()
        _lib.call("hg_graph_state", self._native().ptr, ctypes.byref(cached), ctypes.byref(broken))
        return int(cached.value), bool(broken.value)

    def set_memory_budget(self, nbytes: int) -> None:
        """
        The factorizations' memory budget in bytes (ordito's ``CHOLESKY_MEMORY_BUDGET``,
        ``cholesky.py:72``): one whose peak would exceed it is refused and its system iterated. 0
        (the default) takes 90 % of the device memory free when a factorization starts. A heat
        system refused under it is marched band by band (``linalg.py:2765-2939``), and a refused
        Poisson system is solved under a multigrid V-cycle.
        """
        _lib.call("hg_set_memory_budget", self._native().ptr, int(nbytes))

    def set_budget_probe(self, probe: bool = True) -> None:
        """Whether a system is asked of its analysis, before it iterates, if it fits an explicitly
        set budget (default); off, a refusal is met only when a rejected settled iterate falls back
        to a factorization (ordito's size test skipping a small operator, ``linalg.py:2729``)."""
        _lib.call("hg_set_budget_probe", self._native().ptr, 1 if probe else 0)

    def set_settle_rule(
        self, check_rounds: int = 16, change_tolerance: float = 1e-6, settle_rounds: int = 0
    ) -> None:
        """The heat solves' settle rule (``solve_spd_settled``'s parameters, ``linalg.py:2490``):
        every ``check_rounds`` rounds, stop once every entry changed by under ``change_tolerance``
        of itself, or once no entry was reached for ``settle_rounds`` rounds (0: the window the
        time gives)."""
        _lib.call(
            "hg_set_settle_rule", self._native().ptr, int(check_rounds), float(change_tolerance),
            int(settle_rounds),
        )

    def factor(self) -> bool:
        """Factor the heat and Poisson systems now (``heat.py:3863-3888``); False when one is
        refused under the memory budget, whose solves then keep iterating."""
        out = ctypes.c_int32()
        _lib.call(
            "hg_factor",
            self._native().ptr,
            _lib.HG_SYSTEM_HEAT | _lib.HG_SYSTEM_POISSON,
            self._precision_code,
            ctypes.byref(out),
        )
        return bool(out.value)

    def set_refinement(self, heat: float | None = None, poisson: float | None = None) -> None:
        """
        Refinement targets of this solver's factored solves at its precision: the heat step's
        componentwise tolerance and the Poisson step's normwise one, each in (0, 1); None keeps
        the current value. (No counterpart in the reference, whose targets are fixed at
        ``linalg.py:181`` and ``heat.py:108``.)
        """
        _lib.call(
            "hg_set_refinement",
            self._native().ptr,
            self._precision_code,
            math.nan if heat is None else float(heat),
            math.nan if poisson is None else float(poisson),
        )

    def enable_timing(self, enable: bool = True) -> None:
        """Bracket every phase of later solves with HIP events on the solver's stream."""
        _lib.call("hg_timing_enable", self._native().ptr, int(bool(enable)))

    def read_timing(self, reset: bool = False) -> dict[str, tuple[float, int]]:
        """Accumulated device milliseconds and span counts per phase since the last reset."""
        ms = (ctypes.c_double * 8)()
        counts = (ctypes.c_int64 * 8)()
        _lib.call("hg_timing_read", self._native().ptr, ms, counts, int(bool(reset)))
        return {name: (float(ms[i]), int(counts[i])) for i, name in enumerate(_lib.PHASES)}

    def plan_stats(self) -> dict[str, int]:
        """Statistics of the Cholesky analysis: supernodes, levels, solve-block entries (stored,
        structural nonzeros, and those one sweep direction reads), flops."""
        out = (ctypes.c_int64 * 8)()
        _lib.call("hg_plan_stats", self._native().ptr, out)
        keys = (
            "supernodes",
            "levels",
            "solve_block_entries",
            "flops",
            "arena_entries",
            "max_front",
            "front_vector_entries",
            "rows",
        )
        stats = dict(zip(keys, (int(v) for v in out)))
        entries = (ctypes.c_int64 * 3)()
        _lib.call("hg_inverse_factor_entries", self._native().ptr, entries)
        stats["solve_block_nonzeros"] = int(entries[0])
        stats["solve_block_read"] = int(entries[1])
        return stats

    # -- solves ----------------------------------------------------------------------------------

    def _require_fields(self, caller: str, rhs, solution, per_column: int) -> int:
        n = self._bound_mesh().n_vertices
        if rhs.shape[-1] != n * per_column or tuple(solution.shape) != tuple(rhs.shape):
            raise ValueError(
                f"{caller}: rhs must hold {n} rows for this mesh and solution its shape, got "
                f"{tuple(rhs.shape)} and {tuple(solution.shape)}"
            )
        return 1 if rhs.ndim == 1 else int(rhs.shape[0])

    @property
    def rim_size(self) -> int:
        """The number of vertices with an edge on a single face (0 for a closed mesh)."""
        if self._rim is None:
            out = ctypes.c_int64()
            _lib.call("hg_rim_count", self._native().ptr, ctypes.byref(out))
            self._rim = int(out.value)
        return self._rim

    _SYSTEMS = {
        "heat": _lib.HG_SYSTEM_HEAT,
        "poisson": _lib.HG_SYSTEM_POISSON,
        "pinned": _lib.HG_SYSTEM_PINNED,
        "poisson_regularized": _lib.HG_SYSTEM_REGULARIZED,
    }

    def _count(self, role: str) -> None:
        """Count a solve of ``role``'s system; factor it on its second (rent or buy)."""
        self._solves[role] = self._solves.get(role, 0) + 1
        if self._solves[role] >= 2:
            system = self._SYSTEMS[role]
            state = self._factored()
            if not state & system and not state & (system << 4):
                out = ctypes.c_int32()
                _lib.call("hg_factor", self._native().ptr, system, self._precision_code, ctypes.byref(out))
                if out.value:
                    _LOGGER.debug("HeatSolver: factored the %s system on its second solve", role)
                else:
                    _LOGGER.debug("HeatSolver: the %s system is over the memory budget", role)

    def diffuse(self, rhs, solution) -> None:
        """Solve ``(M - tL) u = rhs`` to settled entries, ``rhs`` (V,) or (k, V) float64 on the GPU
        (``heat.py:3890-3914``); ``solution`` is overwritten."""
        columns = self._require_fields("HeatSolver.diffuse", rhs, solution, 1)
        self._count("heat")
        mode = _lib.HG_MODE_FACTORED if self._factored() & 1 else _lib.HG_MODE_ITERATE
        flags = ctypes.c_int32()
        _lib.call(
            "hg_diffuse",
            self._native().ptr,
            _arrays.pointer(rhs),
            _arrays.pointer(solution),
            columns,
            mode,
            self._precision_code,
            ctypes.byref(flags),
        )
        self._read_stats()
        if flags.value & 1:
            self._ill_conditioned = True
            _LOGGER.debug(
                "HeatSolver: the heat system's settled iterate failed its backward-error check; "
                "factored it"
            )

    def diffuse_pinned(self, rhs, solution) -> bool:
        """
        Solve ``(M - tL) u = rhs`` on the interior vertices with ``u = 0`` on every vertex with an
        edge on a single face, settled as ``diffuse`` is (``heat.py:4153-4200``); ``rhs`` (V,) or
        (k, V) float64 on the GPU, its rim entries ignored. Returns whether the mesh has a rim: a
        closed mesh has nothing to pin, and ``solution`` is then left as it was.
        """
        columns = self._require_fields("HeatSolver.diffuse_pinned", rhs, solution, 1)
        if self.rim_size == 0:
            return False
        self._count("pinned")
        mode = _lib.HG_MODE_FACTORED if self._factored() & _lib.HG_SYSTEM_PINNED else _lib.HG_MODE_ITERATE
        flags = ctypes.c_int32()
        has_rim = ctypes.c_int32()
        _lib.call(
            "hg_diffuse_pinned",
            self._native().ptr,
            _arrays.pointer(rhs),
            _arrays.pointer(solution),
            columns,
            mode,
            self._precision_code,
            ctypes.byref(flags),
            ctypes.byref(has_rim),
        )
        self._read_stats()
        if flags.value & 1:
            self._ill_conditioned = True
            _LOGGER.debug(
                "HeatSolver: the pinned system's settled iterate failed its backward-error check; "
                "factored it"
            )
        return bool(has_rim.value)

    def solve_poisson(self, rhs, solution) -> None:
        """Solve ``-L phi = rhs`` to the heat method's tolerance (``heat.py:4062-4151``)."""
        columns = self._require_fields("HeatSolver.solve_poisson", rhs, solution, 1)
        self._count("poisson")
        factored = self._factored()
        if (factored & 1 or self._ill_conditioned) and not factored & 2:
            out = ctypes.c_int32()
            _lib.call(
                "hg_factor", self._native().ptr, _lib.HG_SYSTEM_POISSON, self._precision_code,
                ctypes.byref(out),
            )
        mode = _lib.HG_MODE_FACTORED if self._factored() & 2 else _lib.HG_MODE_ITERATE
        _lib.call(
            "hg_solve_poisson",
            self._native().ptr,
            _arrays.pointer(rhs),
            _arrays.pointer(solution),
            columns,
            mode,
            self._precision_code,
        )
        self._read_stats()

    def _read_stats(self) -> None:
        stats = _lib.SolveStats()
        _lib.call("hg_last_stats", self._native().ptr, ctypes.byref(stats))
        self.last_stats = stats.as_dict()

    def geodesic(self, source_sets: Sequence[Any]):
        """
        Distance fields for many source sets at once, ``(k, V)`` float64 on the GPU: the whole path
        in one library call, the columns solved together. Counts as one solve of each system.
        """
        torch = _arrays.torch_module()
        mesh = self._bound_mesh()
        n = mesh.n_vertices
        k = len(source_sets)
        if k == 0 or n == 0 or mesh.n_faces == 0:
            return torch.zeros((k, n), dtype=torch.float64, device="cuda")
        # One-dimensional tensors concatenate as they are (one conversion of the result); anything
        # else is converted set by set.
        same_kind = all(type(s) is torch.Tensor and s.dim() == 1 for s in source_sets) and len(
            {(s.device, s.dtype) for s in source_sets}
        ) == 1
        if same_kind:
            counts = [s.numel() for s in source_sets]
            flat = torch.cat(source_sets) if sum(counts) > 0 else None
            if flat is not None:
                flat = flat.to(device="cuda", dtype=torch.int32).contiguous()
        else:
            tensors = [_arrays.sources_tensor(s) for s in source_sets]
            counts = [s.numel() for s in tensors]
            flat = torch.cat(tensors) if sum(counts) > 0 else None
        if flat is None:
            flat = torch.zeros(1, dtype=torch.int32, device="cuda")
        offsets = [0]
        for c in counts:
            offsets.append(offsets[-1] + c)
        return self.geodesic_flat(flat, offsets)

    def geodesic_flat(self, sources, offsets):
        """
        ``geodesic`` for source sets given as one flat ``(S,)`` int32 tensor on the GPU and
        ``k + 1`` host offsets (a sequence of ints or a ``ctypes`` array of ``c_int64``): set ``q``
        is ``sources[offsets[q]:offsets[q + 1]]``. The serving form: a batch assembled once is
        solved without converting or concatenating its sets on every call.
        """
        torch = _arrays.torch_module()
        mesh = self._bound_mesh()
        n = mesh.n_vertices
        k = len(offsets) - 1
        if k < 0:
            raise ValueError("geodesic_flat: offsets must hold k + 1 >= 1 entries")
        if (
            type(sources) is not torch.Tensor
            or sources.dtype != torch.int32
            or sources.device.type != "cuda"
            or sources.dim() != 1
            or not sources.is_contiguous()
        ):
            raise ValueError("geodesic_flat: sources must be a contiguous (S,) int32 tensor on the GPU")
        if offsets[0] != 0 or offsets[k] > sources.numel():
            # (that they do not decrease the library checks)
            raise ValueError("geodesic_flat: offsets must start at 0 and end within sources")
        if k == 0 or n == 0 or mesh.n_faces == 0:
            return torch.zeros((k, n), dtype=torch.float64, device="cuda")
        if not isinstance(offsets, ctypes.Array):
            offsets = _lib.offsets_array(offsets)
        out = torch.empty((k, n), dtype=torch.float64, device="cuda")  # every entry is written
        # Rent or buy, per system (heat.py:4394-4404): the second solve factors each. One handle
        # and one factored-state query for the whole call. At "igl" a mesh with a rim also solves
        # the rim-pinned system (heat.py:594-597), which the library solves through its factor
        # whenever the heat system's is held.
        native = self._native()
        factored = self._factored_on(native)
        # At "potpourri3d" the Poisson step solves the regularized system (heat.py:3121-3128).
        if self.method == "potpourri3d":
            poisson = ("poisson_regularized", _lib.HG_SYSTEM_REGULARIZED)
        else:
            poisson = ("poisson", _lib.HG_SYSTEM_POISSON)
        roles = [("heat", _lib.HG_SYSTEM_HEAT), poisson]
        if self.method == "igl" and self.rim_size > 0:
            roles.append(("pinned", _lib.HG_SYSTEM_PINNED))
        for role, system in roles:
            self._solves[role] = self._solves.get(role, 0) + 1
            if self._solves[role] >= 2 and not factored & system and not factored & (system << 4):
                done = ctypes.c_int32()
                _lib.call("hg_factor", native.ptr, system, self._precision_code, ctypes.byref(done))
                if done.value:
                    factored |= system
                    _LOGGER.debug("HeatSolver: factored the %s system on its second solve", role)
                else:
                    factored |= system << 4
                    _LOGGER.debug("HeatSolver: the %s system is over the memory budget", role)
        mode = _lib.HG_MODE_FACTORED if factored & _lib.HG_SYSTEM_HEAT else _lib.HG_MODE_ITERATE
        stats = _lib.SolveStats()
        _lib.call(
            "hg_geodesic",
            native.ptr,
            _arrays.pointer(sources),
            offsets,
            k,
            _arrays.pointer(out),
            mode,
            self._precision_code,
            ctypes.byref(stats),
        )
        self.last_stats = stats.as_dict()
        if stats.backward_error_failed:
            self._ill_conditioned = True
        return out


def _resolve_arguments(mesh, faces, sources, t, vertices):
    """ordito's ``mesh_arguments`` (``mesh.py:3612``): a Trimesh, or vertices and faces."""
    if isinstance(mesh, Trimesh):
        if vertices is not None:
            raise ValueError("heat_geodesic: pass a Trimesh or vertices and faces, not both")
        # heat_geodesic(mesh, sources, /, *, t): the second positional argument is the sources.
        if faces is not None:
            if sources is not None:
                t = sources if t is None else t
            sources = faces
        return mesh, True, sources, t, mesh._template
    v = mesh if mesh is not None else vertices
    if v is None or faces is None:
        raise ValueError("heat_geodesic: pass a Trimesh, or vertices and faces")
    return Trimesh(v, faces), False, sources, t, v


def heat_geodesic(
    mesh: Any = None,
    faces: Any = None,
    sources: Any = None,
    t: float | None = None,
    *,
    vertices: Any = None,
    use_robust: bool = False,
    method: str = "meshlab",
    precision: str = "fp64",
):
    """
    Approximate geodesic distance to the nearest source vertex by the heat method (ordito
    ``heat.py:370-609``): MeshLab's discretization (``method="meshlab"``, the default),
    ``igl::heat_geodesics``' (``"igl"``: the Voronoi mass, ``h`` over the unique edges, and the
    Neumann heat averaged with the heat pinned to zero on the rim) or geometry-central's
    (``"potpourri3d"``: the barycentric mass, ``h`` over the unique edges, and the Poisson step
    regularized with ``1e-6`` times the identity, ``heat.py:3121-3128``). Takes a ``Trimesh`` (which
    keeps the operators and factorizations for later calls) or ``vertices, faces``; returns
    ``(V,)`` float64 -- a torch tensor on the GPU when the vertices were one, a numpy array
    otherwise.

    ``use_robust`` (``"meshlab"`` and ``"igl"``) builds the half-cotangents from side lengths
    mollified by the smallest constant that leaves every triangle's triangle-inequality margin
    ``1e-5`` of the mean side (``heat.py:229-243``, ``laplacian.py:779-874``): a mesh with
    degenerate triangles then solves; a clean mesh's weights move only by float32 rounding.

    Raises ``ValueError`` for an unknown ``method`` or a source outside ``[0, n_vertices)``, and
    ``NotImplementedError`` for ``use_robust`` at ``"potpourri3d"`` (geometry-central's intrinsic
    Delaunay method).
    """
    if method not in _METHODS:
        raise ValueError(f"method must be 'meshlab', 'igl' or 'potpourri3d', got {method!r}")
    _require_method(method, geodesic=True)
    _method_code(method, use_robust)
    _require_precision(precision)
    bound, owned, sources, t, template = _resolve_arguments(mesh, faces, sources, t, vertices)
    torch = _arrays.torch_module()
    n = bound.n_vertices
    src = _arrays.sources_tensor(sources if sources is not None else [])
    if n == 0 or bound.n_faces == 0 or src.numel() == 0:
        return _arrays.like_input(torch.zeros(n, dtype=torch.float64, device="cuda"), template)
    solver = bound.heat_solver(t, use_robust=use_robust, method=method, precision=precision)
    distance = solver.geodesic([src])[0]
    if not owned and solver.nbytes > 0:
        _LOGGER.info(
            "heat_geodesic: built a sparse Cholesky factorization (%d bytes) and is discarding it; "
            "pass an ordito_amd.mesh.Trimesh to keep it for later calls",
            solver.nbytes,
        )
    return _arrays.like_input(distance, template)
