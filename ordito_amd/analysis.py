"""
Host-side view of the sparse Cholesky analysis (nested dissection, supernodes, level schedule).

Wraps ``hg_plan_*`` of the C ABI: the same analysis the library runs before factoring, read out
as numpy arrays for inspection, tests and the tile autotuner. Needs no GPU.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

_INT32_FIELDS = {
    "perm": (0, "n"),
    "col_begin": (1, "s"),
    "ncol": (2, "s"),
    "nrow": (3, "s"),
    "parent": (4, "s"),
    "level": (5, "s"),
    "row_ptr": (6, "s1"),
    "rows": (7, "rows"),
    "rel": (8, "rows"),
    "child_ptr": (9, "s1"),
    "children": (10, "children"),
    "level_ptr": (11, "levels1"),
    "level_nodes": (12, "s"),
    "entry_src": (17, "entries"),
    "entry_row": (18, "entries"),
    "entry_col": (19, "entries"),
}
_INT64_FIELDS = {"front_off": (13, "s"), "w_off": (14, "s"), "vec_off": (15, "s"), "entry_dst": (16, "entries")}


@dataclass
class HostPlan:
    n: int
    n_super: int
    n_levels: int
    arena_size: int
    w_size: int
    vec_size: int
    arrays: dict[str, np.ndarray]

    def __getattr__(self, name: str) -> np.ndarray:
        arrays = self.__dict__.get("arrays", {})
        if name in arrays:
            return arrays[name]
        raise AttributeError(name)


def analyse(
    offsets, columns, xyz, *, leaf_rows: int = 0, top_rows: int = 0, validate: bool = False
) -> HostPlan:
    """
    Run the library's analysis on a host CSR pattern (sorted rows) and float32 (n, 3) xyz. With
    ``validate`` the library also checks every index its device sweeps form from the plan.
    """
    lib = _lib.load()
    offsets = np.ascontiguousarray(offsets, dtype=np.int32)
    columns = np.ascontiguousarray(columns, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    n = offsets.size - 1
    handle = ctypes.c_void_p()
    _lib.check(
        lib.hg_plan_create(
            offsets.ctypes.data, columns.ctypes.data, n, xyz.ctypes.data, leaf_rows, top_rows,
            ctypes.byref(handle),
        )
    )
    try:
        if validate:
            _lib.check(lib.hg_plan_validate(handle))
        sizes = (ctypes.c_int64 * 8)()
        _lib.check(lib.hg_plan_sizes(handle, sizes))
        n_super, n_levels, n_rows, n_children, n_entries, arena, w_size, vec = (int(v) for v in sizes)
        count = {
            "n": n,
            "s": n_super,
            "s1": n_super + 1,
            "rows": n_rows,
            "children": n_children,
            "levels1": n_levels + 1,
            "entries": n_entries,
        }
        arrays: dict[str, np.ndarray] = {}
        for name, (field, size) in _INT32_FIELDS.items():
            out = np.zeros(max(count[size], 1), dtype=np.int32)
            _lib.check(lib.hg_plan_copy(handle, field, out.ctypes.data))
            arrays[name] = out[: count[size]]
        for name, (field, size) in _INT64_FIELDS.items():
            out = np.zeros(max(count[size], 1), dtype=np.int64)
            _lib.check(lib.hg_plan_copy(handle, field, out.ctypes.data))
            arrays[name] = out[: count[size]]
        # The numeric factorization's tile launches (kind, level, first, count) and tasks.
        counts = (ctypes.c_int64 * 2)()
        _lib.check(lib.hg_plan_factor_tasks(handle, counts, None, None))
        launches = np.zeros((max(int(counts[0]), 1), 4), dtype=np.int32)
        tasks = np.zeros((max(int(counts[1]), 1), 4), dtype=np.int32)
        _lib.check(lib.hg_plan_factor_tasks(handle, counts, launches.ctypes.data, tasks.ctypes.data))
        arrays["factor_launches"] = launches[: int(counts[0])]
        arrays["factor_tasks"] = tasks[: int(counts[1])]
        return HostPlan(n, n_super, n_levels, arena, w_size, vec, arrays)
    finally:
        lib.hg_plan_destroy(handle)
