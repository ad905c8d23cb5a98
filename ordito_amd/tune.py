"""
Autotuner of the sparse Cholesky's tunables for a given mesh or packed batch.

What it tunes (``hg_set_plan_options``): the nested dissection's leaf size and top merge (the
front sizes the tile launches see, so how full the 64-row tiles and the FP64 matrix-core steps
are) and the task count a sweep level's products are split towards (how many workgroups a
level with few, large fronts gets). The reference fixes these (leaf 32, top 512,
``cholesky.py:538-660``); here they are measured per workload: every candidate ordering is
analysed, factored and timed on the served operation -- a batch of queries through the factors --
and the split target is then tuned on the best ordering, which only rebuilds the sweep tasks.
The fastest setting is left in place.
"""

from __future__ import annotations

import time
from typing import Any, Sequence

import numpy as np

from . import _lib


def _handle(target) -> Any:
    if hasattr(target, "_native"):
        return target._native().ptr
    return target._ptr


def _default_batch(target, columns: int, seed: int):
    """Single-source queries: `columns` on a mesh, one per packed mesh on a batch."""
    rng = np.random.default_rng(seed)
    if hasattr(target, "sizes"):
        return [(m, np.array([rng.integers(n)], dtype=np.int32)) for m, n in enumerate(target.sizes)]
    n = target._bound_mesh().n_vertices
    return [np.array([s], dtype=np.int32) for s in rng.choice(n, size=min(columns, n), replace=False)]


def _time(target, batch, repeats: int) -> float:
    import torch

    target.geodesic(batch)  # warm: workspaces sized, first-use costs paid
    torch.cuda.synchronize()
    samples = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        target.geodesic(batch)
        torch.cuda.synchronize()
        samples.append(time.perf_counter() - t0)
    return float(np.median(samples))


def autotune(
    target,
    *,
    columns: int = 64,
    leaf: Sequence[int] = (32, 48, 64, 96),
    top: Sequence[int] = (256, 512, 1024),
    split: Sequence[int] = (512, 1024, 2048, 4096),
    repeats: int = 3,
    seed: int = 0,
) -> dict:
    """
    Tune ``target`` (a ``HeatSolver`` or a ``MeshBatch``) and leave the fastest setting in place.
    Returns the best ``{"leaf", "top", "split", "seconds"}`` and every trial.
    """
    handle = _handle(target)
    batch = _default_batch(target, columns, seed)
    trials = []
    best = None
    for leaf_rows in leaf:
        for top_rows in top:
            _lib.call("hg_set_plan_options", handle, int(leaf_rows), int(top_rows), 0)
            factor_start = time.perf_counter()
            target.factor()
            factor_seconds = time.perf_counter() - factor_start
            seconds = _time(target, batch, repeats)
            trial = {"leaf": int(leaf_rows), "top": int(top_rows), "split": 0,
                     "seconds": seconds, "factor_seconds": factor_seconds}
            trials.append(trial)
            if best is None or seconds < best["seconds"]:
                best = dict(trial)
    _lib.call("hg_set_plan_options", handle, best["leaf"], best["top"], 0)
    target.factor()
    for target_tasks in split:
        _lib.call("hg_set_plan_options", handle, best["leaf"], best["top"], int(target_tasks))
        seconds = _time(target, batch, repeats)
        trial = {"leaf": best["leaf"], "top": best["top"], "split": int(target_tasks), "seconds": seconds}
        trials.append(trial)
        if seconds < best["seconds"]:
            best = dict(trial)
    _lib.call("hg_set_plan_options", handle, best["leaf"], best["top"], best["split"])
    target.factor()
    return {"best": best, "trials": trials}
