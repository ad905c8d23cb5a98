"""
Packed meshes (``ordito_amd.batch.MeshBatch``, ``hg_mesh_create_packed``): the analysis gives each
mesh its own tree (CPU), and on the GPU every query's field is the one its mesh gives alone, and the
oracle's, within the stated bars.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import heat_oracle as ho
from tests.conftest import mesh

NAMES = ["icosphere3", "torus", "grid", "bumpy"]


def _packed_pattern(names):
    """The operator pattern of the meshes side by side (block-diagonal), and their coordinates."""
    indptr, indices, xyz = [np.zeros(1, dtype=np.int64)], [], []
    base = 0
    for name in names:
        v, f = mesh(name)
        p = ho.operator_pattern(len(v), f)
        indptr.append(p.indptr[1:] + indptr[-1][-1])
        indices.append(p.indices + base)
        xyz.append(v)
        base += len(v)
    return np.concatenate(indptr), np.concatenate(indices), np.concatenate(xyz)


def test_packed_analysis_keeps_each_mesh_tree():
    """Each connected component is dissected alone: the packed plan's supernodes are exactly the
    meshes' own (sizes, structure sizes, levels), and it is as deep as the deepest mesh."""
    from ordito_amd import analysis

    alone = []
    for name in NAMES:
        v, f = mesh(name)
        p = ho.operator_pattern(len(v), f)
        alone.append(analysis.analyse(p.indptr, p.indices, v))
    indptr, indices, xyz = _packed_pattern(NAMES)
    packed = analysis.analyse(indptr, indices, xyz, validate=True)
    assert packed.n_levels == max(a.n_levels for a in alone)
    key = lambda plan: sorted(zip(plan.ncol.tolist(), plan.nrow.tolist(), plan.level.tolist()))
    expected = sorted(sum((key(a) for a in alone), []))
    assert key(packed) == expected
    # Level 0 holds every mesh's leaves.
    assert packed.level_ptr[1] == sum(a.level_ptr[1] for a in alone)


def _queries(sizes, seed=0):
    rng = np.random.default_rng(seed)
    queries = []
    for m, n in enumerate(sizes):
        for _ in range(1 + m % 3):
            count = int(rng.integers(1, 4))
            queries.append((m, rng.choice(n, size=count, replace=False).astype(np.int32)))
    rng.shuffle(queries)
    return queries


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_matches_each_mesh_alone(cuda, precision):
    """Every query's field is the one its mesh's own solver gives (fp64: 1e-10 of the range, the
    refinement's agreement; mixed: 1e-6) and the oracle's (1e-9; mixed 1e-4, the served bar)."""
    from ordito_amd import heat
    from ordito_amd.batch import MeshBatch

    meshes = [mesh(name) for name in NAMES]
    batch = MeshBatch([(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda()) for v, f in meshes],
                      precision=precision)
    assert batch.factor()
    queries = _queries(batch.sizes)
    fields = batch.geodesic([(m, cuda.from_numpy(s).cuda()) for m, s in queries])
    assert len(fields) == len(queries)
    times = batch.times
    bar_alone, bar_oracle = (1e-10, 1e-9) if precision == "fp64" else (1e-6, 1e-4)
    solvers = {}
    for (m, s), got in zip(queries, fields):
        v, f = meshes[m]
        if m not in solvers:
            tm = heat.Trimesh(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda())
            solver = tm.heat_solver(method="meshlab", precision=precision)
            assert solver.factor()
            solver._solves = {"heat": 2, "poisson": 2}
            solvers[m] = (tm, solver)
            assert times[m] == solver.time
        got = got.cpu().numpy()
        alone = solvers[m][1].geodesic([cuda.from_numpy(s).cuda()])[0].cpu().numpy()
        worst, _ = ho.relative_error(got, alone)
        assert worst <= bar_alone, (m, worst)
        expected = ho.geodesic(v, f, s).distance
        worst, _ = ho.relative_error(got, expected)
        assert worst <= bar_oracle, (m, worst)


@pytest.mark.gpu
def test_batch_reports_bad_queries(cuda):
    from ordito_amd import _lib
    from ordito_amd.batch import MeshBatch

    meshes = [mesh(name) for name in NAMES[:2]]
    batch = MeshBatch([(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda()) for v, f in meshes])
    with pytest.raises(ValueError):
        batch.geodesic([(0, np.array([batch.sizes[0]], dtype=np.int32))])
    with pytest.raises(ValueError):
        batch.geodesic([(2, np.array([0], dtype=np.int32))])
    # A packed mesh answers through hg_geodesic_packed only.
    with pytest.raises(ValueError):
        _lib.call("hg_geodesic", batch._ptr, None, _lib.offsets_array([0, 0]), 1, None, 1, 0, None)
    # A query with no sources is a zeros field, as ordito's (heat.py:567-568).
    empty = batch.geodesic([(1, np.zeros(0, dtype=np.int32))])[0].cpu().numpy()
    assert empty.shape == (batch.sizes[1],) and not np.any(empty)


@pytest.mark.gpu
def test_batch_rejects_faces_outside_their_mesh(cuda):
    from ordito_amd.batch import MeshBatch

    v, f = mesh("icosphere3")
    bad = f.copy()
    bad[0] = len(v)  # valid in the packed numbering, outside its own mesh
    with pytest.raises(ValueError):
        MeshBatch([(cuda.from_numpy(v).cuda(), cuda.from_numpy(f).cuda()),
                   (cuda.from_numpy(v).cuda(), cuda.from_numpy(bad).cuda())])
