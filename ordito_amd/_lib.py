"""
ctypes binding of ``libheatgeo.so`` (the C ABI declared in ``include/heatgeo.h``).

The library is loaded from this package's directory; there is no fallback: a missing library or a
missing GPU raises. Status codes map to the exceptions ordito raises on the same conditions
(``heat.py:5488-5503`` ValueError on a bad source, ``cholesky.py:90-118``
NotPositiveDefiniteError(ValueError), RuntimeError otherwise).
"""

from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIBRARY_PATH = Path(__file__).resolve().parent / "libheatgeo.so"

HG_OK = 0
HG_ERR_VALUE = 1
HG_ERR_RUNTIME = 2
HG_ERR_NOT_PD = 3
HG_ERR_BUDGET = 4

HG_METHOD_MESHLAB = 0
HG_METHOD_IGL = 1
HG_METHOD_POTPOURRI3D = 2
HG_METHOD_ROBUST = 0x100
HG_MODE_ITERATE = 0
HG_MODE_FACTORED = 1
HG_PRECISION_FP64 = 0
HG_PRECISION_MIXED = 1
HG_SYSTEM_HEAT = 1
HG_SYSTEM_POISSON = 2
HG_SYSTEM_PINNED = 4
HG_SYSTEM_REGULARIZED = 8

# The C ABI's exported symbols, in header order; the load test checks every one resolves.
EXPORTED = (
    "hg_last_error",
    "hg_version",
    "hg_mesh_create",
    "hg_mesh_destroy",
    "hg_mesh_set_stream",
    "hg_operators",
    "hg_operator_views",
    "hg_operator_pattern",
    "hg_factor",
    "hg_factored",
    "hg_set_graphs",
    "hg_graph_state",
    "hg_set_memory_budget",
    "hg_set_budget_probe",
    "hg_set_settle_rule",
    "hg_inverse_factor_entries",
    "hg_set_refinement",
    "hg_set_plan_options",
    "hg_diffuse",
    "hg_diffuse_pinned",
    "hg_rim_count",
    "hg_solve_poisson",
    "hg_unit_gradient_divergence",
    "hg_seed_sources",
    "hg_shift_to_sources",
    "hg_shift_to_sources_noalloc",
    "hg_geodesic",
    "hg_mesh_create_packed",
    "hg_geodesic_packed",
    "hg_packed_times",
    "hg_server_create",
    "hg_server_add_mesh",
    "hg_server_submit",
    "hg_server_wait",
    "hg_server_stats",
    "hg_server_destroy",
    "hg_last_stats",
    "hg_nbytes",
    "hg_release",
    "hg_device_bytes",
    "hg_plan_stats",
    "hg_plan_create",
    "hg_plan_sizes",
    "hg_plan_copy",
    "hg_plan_validate",
    "hg_plan_factor_tasks",
    "hg_plan_destroy",
    "hg_timing_enable",
    "hg_timing_read",
)


class NotPositiveDefiniteError(ValueError):
    """The operator met a pivot that is not positive (``cholesky.py:90-118``)."""


class OperatorView(ctypes.Structure):
    _fields_ = [
        ("n_vertices", ctypes.c_int64),
        ("n_faces", ctypes.c_int64),
        ("nnz", ctypes.c_int64),
        ("offsets", ctypes.c_void_p),
        ("columns", ctypes.c_void_p),
        ("laplacian", ctypes.c_void_p),
        ("heat", ctypes.c_void_p),
        ("poisson", ctypes.c_void_p),
        ("mass", ctypes.c_void_p),
        ("cot", ctypes.c_void_p),
        ("normals", ctypes.c_void_p),
        ("areas", ctypes.c_void_p),
        ("t", ctypes.c_double),
        ("settle_rounds", ctypes.c_int32),
        ("time_from_edges", ctypes.c_int32),
        ("mollify_delta", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("evictions", ctypes.c_int64),
        ("resident", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("resident_bytes", ctypes.c_int64),
        ("queries_per_s", ctypes.c_double),
        ("p50_ms", ctypes.c_double),
        ("p99_ms", ctypes.c_double),
        ("mean_ms", ctypes.c_double),
        ("now_seconds", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "reserved"}


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("heat_rounds", ctypes.c_int32),
        ("poisson_rounds", ctypes.c_int32),
        ("heat_refinements", ctypes.c_int32),
        ("poisson_refinements", ctypes.c_int32),
        ("backward_error_failed", ctypes.c_int32),
        ("pinned_solves", ctypes.c_int32),
        ("heat_marched", ctypes.c_int32),
        ("poisson_multigrid", ctypes.c_int32),
    ]

    def as_dict(self) -> dict[str, int]:
        return {name: int(getattr(self, name)) for name, _ in self._fields_}


_lib: ctypes.CDLL | None = None
_lock = threading.Lock()

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PI32 = ctypes.POINTER(ctypes.c_int32)

_SIGNATURES = {
    "hg_last_error": ([], ctypes.c_char_p),
    "hg_version": ([], ctypes.c_char_p),
    "hg_mesh_create": ([_P, _I64, _P, _I64, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "hg_mesh_destroy": ([_P], ctypes.c_int),
    "hg_mesh_set_stream": ([_P, _P], ctypes.c_int),
    "hg_operators": ([_P, ctypes.c_double, _I32, _I32], ctypes.c_int),
    "hg_operator_views": ([_P, ctypes.POINTER(OperatorView)], ctypes.c_int),
    "hg_operator_pattern": ([_P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _PI64], ctypes.c_int),
    "hg_factor": ([_P, _I32, _I32, _PI32], ctypes.c_int),
    "hg_factored": ([_P, _PI32], ctypes.c_int),
    "hg_set_refinement": ([_P, _I32, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "hg_set_plan_options": ([_P, _I32, _I32, _I32], ctypes.c_int),
    "hg_diffuse": ([_P, _P, _P, _I64, _I32, _I32, _PI32], ctypes.c_int),
    "hg_diffuse_pinned": ([_P, _P, _P, _I64, _I32, _I32, _PI32, _PI32], ctypes.c_int),
    "hg_rim_count": ([_P, _PI64], ctypes.c_int),
    "hg_set_memory_budget": ([_P, _I64], ctypes.c_int),
    "hg_set_budget_probe": ([_P, _I32], ctypes.c_int),
    "hg_set_settle_rule": ([_P, _I32, ctypes.c_double, _I32], ctypes.c_int),
    "hg_inverse_factor_entries": ([_P, _PI64], ctypes.c_int),
    "hg_set_graphs": ([_P, _I32], ctypes.c_int),
    "hg_graph_state": ([_P, _PI64, _PI32], ctypes.c_int),
    "hg_solve_poisson": ([_P, _P, _P, _I64, _I32, _I32], ctypes.c_int),
    "hg_unit_gradient_divergence": ([_P, _P, _P, _I64], ctypes.c_int),
    "hg_seed_sources": ([_P, _P, _PI64, _I64, _P], ctypes.c_int),
    "hg_shift_to_sources": ([_P, _P, _P, _PI64, _I64], ctypes.c_int),
    "hg_shift_to_sources_noalloc": ([_P, _P, _P, _PI64, _I64], ctypes.c_int),
    "hg_geodesic": ([_P, _P, _PI64, _I64, _P, _I32, _I32, ctypes.POINTER(SolveStats)], ctypes.c_int),
    "hg_mesh_create_packed": ([_P, _PI64, _P, _PI64, _I64, _P, ctypes.POINTER(_P)], ctypes.c_int),
    "hg_geodesic_packed": (
        [_P, _P, _PI64, _PI32, _I64, _P, _I32, ctypes.POINTER(SolveStats)],
        ctypes.c_int,
    ),
    "hg_packed_times": ([_P, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "hg_server_create": ([_I64, _I32, _I32, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "hg_server_add_mesh": ([_P, _P, _I64, _P, _I64, _PI64], ctypes.c_int),
    "hg_server_submit": ([_P, _I64, _P, _I64, _P, _PI64], ctypes.c_int),
    "hg_server_wait": ([_P, _I64, ctypes.c_double, _PI32], ctypes.c_int),
    "hg_server_stats": ([_P, ctypes.c_double, ctypes.POINTER(ServerStats)], ctypes.c_int),
    "hg_server_destroy": ([_P], ctypes.c_int),
    "hg_last_stats": ([_P, ctypes.POINTER(SolveStats)], ctypes.c_int),
    "hg_nbytes": ([_P, _PI64], ctypes.c_int),
    "hg_release": ([_P], ctypes.c_int),
    "hg_device_bytes": ([_P, _PI64], ctypes.c_int),
    "hg_plan_stats": ([_P, _PI64], ctypes.c_int),
    "hg_plan_create": ([_P, _P, _I64, _P, _I32, _I32, ctypes.POINTER(_P)], ctypes.c_int),
    "hg_plan_sizes": ([_P, _PI64], ctypes.c_int),
    "hg_plan_copy": ([_P, _I32, _P], ctypes.c_int),
    "hg_plan_validate": ([_P], ctypes.c_int),
    "hg_plan_factor_tasks": ([_P, _PI64, _P, _P], ctypes.c_int),
    "hg_plan_destroy": ([_P], ctypes.c_int),
    "hg_timing_enable": ([_P, _I32], ctypes.c_int),
    "hg_timing_read": ([_P, ctypes.POINTER(ctypes.c_double), _PI64, _I32], ctypes.c_int),
}

# hg_timing_read's phases, in order.
PHASES = (
    "factor",
    "heat_sweep",
    "poisson_sweep",
    "residual",
    "divergence",
    "shift",
    "cg",
    "assembly",
)


def load() -> ctypes.CDLL:
    """Load the library once; raise if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIBRARY_PATH.exists():
                raise RuntimeError(
                    f"{LIBRARY_PATH} is not built: run `python -c 'import __graft_entry__ as g; "
                    "g.build()'` (hipcc, gfx950); there is no CPU fallback"
                )
            lib = ctypes.CDLL(str(LIBRARY_PATH))
            for name, (argtypes, restype) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = argtypes
                fn.restype = restype
            _lib = lib
        return _lib


def check(status: int) -> None:
    """Raise the exception a status code names, with the library's message."""
    if status == HG_OK:
        return
    message = load().hg_last_error().decode(errors="replace")
    if status == HG_ERR_VALUE:
        raise ValueError(message)
    if status == HG_ERR_NOT_PD:
        raise NotPositiveDefiniteError(message)
    if status == HG_ERR_BUDGET:
        raise MemoryError(message)
    raise RuntimeError(message)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def offsets_array(offsets) -> ctypes.Array:
    values = [int(v) for v in offsets]
    return (ctypes.c_int64 * len(values))(*values)
