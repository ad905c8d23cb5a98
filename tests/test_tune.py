"""
The autotuner (``ordito_amd.tune.autotune``): it tries every candidate ordering and split target,
leaves the fastest in place, and the answers through the tuned factors are still the oracle's.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import heat_oracle as ho
from tests.conftest import mesh

pytestmark = pytest.mark.gpu


def test_autotune_keeps_answers(cuda):
    from ordito_amd import heat
    from ordito_amd.tune import autotune

    v, f = mesh("torus")
    tm = heat.Trimesh(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda())
    solver = tm.heat_solver(method="meshlab", precision="fp64")
    report = autotune(solver, columns=8, leaf=(16, 48), top=(64, 512), split=(256, 1024), repeats=1)
    assert len(report["trials"]) == 2 * 2 + 2
    best = report["best"]
    assert best["seconds"] == min(t["seconds"] for t in report["trials"])
    sources = np.array([3], dtype=np.int32)
    got = solver.geodesic([cuda.from_numpy(sources).cuda()])[0].cpu().numpy()
    worst, _ = ho.relative_error(got, ho.geodesic(v, f, sources).distance)
    assert worst <= 1e-9, worst


def test_autotune_a_batch(cuda):
    from ordito_amd.batch import MeshBatch
    from ordito_amd.tune import autotune

    meshes = [mesh(name) for name in ("icosphere3", "grid")]
    batch = MeshBatch([(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda()) for v, f in meshes])
    report = autotune(batch, leaf=(32,), top=(128, 512), split=(1024,), repeats=1)
    assert report["best"]["top"] in (128, 512)
    v, f = meshes[1]
    got = batch.geodesic([(1, np.array([5], dtype=np.int32))])[0].cpu().numpy()
    worst, _ = ho.relative_error(got, ho.geodesic(v, f, np.array([5])).distance)
    assert worst <= 1e-9, worst
