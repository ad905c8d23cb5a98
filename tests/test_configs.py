"""
The five benchmark configurations at their full sizes on the GPU, and the iterated solves past a
single workgroup.

- torus 1024 x 512 (524,288 vertices; R = 1, r = 0.3, normal jitter 0.1 h): the solver's time rule
  takes the (diag / 400)^2 floor over h^2 and widens the settle window to match (heat.py:3714-3736,
  5228-5243); four single-source queries and one 64-source query through the factors within 1e-9
  of the range of the FP64 oracle's (SuperLU) answers, and one one-shot (iterated) query within
  1e-5, the settled conjugate gradient running its full window on the multi-block cooperative
  kernel.
- icosphere(9) displaced by spherical harmonics (2,621,442 vertices), 16 queries: FP64 and mixed
  (FP32 inverse factors refined in FP64 to 1e-10) agree within 1e-4 of the range; every source reads
  exactly 0 and no distance is negative; the heat system solved against M 1 returns all ones
  (L 1 = 0) to 1e-10; on the same surface at icosphere(7) both match the oracle (1e-9, 1e-8).
- 256 meshes of 1e3-2e5 vertices packed into one system, a query of 1-16 sources each: sampled
  queries equal their mesh solved alone (1e-10) and the oracle on the smallest meshes (1e-9).
- 64 meshes of 1e4-1e6 vertices behind the continuous-batching server under Poisson arrivals:
  sampled requests equal the same request served alone afterwards (1e-10), and the oracle on the
  smallest mesh (1e-9).
- A 10,240-vertex torus (over the 8,192 rows the iterated solve runs as one workgroup): the
  multi-block settled CG and Chebyshev CG against the oracle (1e-5), at 1, 7 and 64 columns, and
  HeatSolver.diffuse over three columns against each column diffused alone.
"""

from __future__ import annotations

import gc
import math

import numpy as np
import pytest

from oracle import heat_oracle as ho

pytestmark = pytest.mark.gpu


def _gpu(cuda, array):
    return cuda.from_numpy(np.ascontiguousarray(array)).cuda()


@pytest.fixture()
def release(cuda):
    """Frees the test's device memory afterwards (the big configurations hold tens of GB)."""
    yield
    gc.collect()
    cuda.cuda.synchronize()
    cuda.cuda.empty_cache()


# --- the torus 1024 x 512 ------------------------------------------------------------------------


@pytest.fixture(scope="module")
def torus_c2():
    from ordito_amd import creation

    vertices, faces = creation.torus(1024, 512, 1.0, 0.3, jitter=0.1, seed=0)
    sources = creation.random_sources(len(vertices), 64, seed=1)
    oracle = ho.FactoredOracle(vertices, faces)
    return vertices, faces, sources, oracle


@pytest.mark.timeout(900)
def test_torus_factored_queries_match_the_oracle(cuda, torus_c2, release):
    from ordito_amd import heat

    vertices, faces, sources, oracle = torus_c2
    print(f"torus 1024x512 factored: {len(vertices)} vertices", flush=True)
    t, window = ho.resolve_time(vertices, faces, None)
    h = ho.mean_face_edge_length(vertices, faces)
    assert t is not None and t > h * h  # the floor decides
    mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
    solver = mesh.heat_solver(method="meshlab")
    assert math.isclose(solver.time, t, rel_tol=1e-12)
    assert solver.settle_rounds == window == math.ceil(192 * math.sqrt(t / (h * h)))
    assert solver.factor()
    picks = sources[[0, 17, 40, 63]]
    got = solver.geodesic([_gpu(cuda, np.array([s], np.int32)) for s in picks]).cpu().numpy()
    for q, s in enumerate(picks):
        worst, _ = ho.relative_error(got[q], oracle.distance(np.array([s])))
        assert worst <= 1e-9, (int(s), worst)
    many = solver.geodesic([_gpu(cuda, sources)])[0].cpu().numpy()
    worst, _ = ho.relative_error(many, oracle.distance(sources))
    assert worst <= 1e-9, worst
    # Many sources: the shift zeroes their mean (heat.py:3134-3162), not each source's value.
    assert abs(np.mean(many[sources])) <= 1e-12 * np.ptp(many)


@pytest.mark.timeout(900)
def test_torus_one_shot_iterates_its_full_window(cuda, torus_c2, release):
    from ordito_amd import heat

    vertices, faces, sources, oracle = torus_c2
    print("torus 1024x512 one-shot (iterated)", flush=True)
    _, window = ho.resolve_time(vertices, faces, None)
    mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
    source = sources[5:6]
    got = heat.heat_geodesic(mesh, _gpu(cuda, source)).cpu().numpy()
    solver = mesh.heat_solver(method="meshlab")
    stats = solver.last_stats
    print(f"one-shot stats {stats}, factored bytes {solver.nbytes}", flush=True)
    # The settled CG runs at least its widened window (the count of reached entries keeps
    # changing while the far field emerges, then the window passes).
    assert stats["heat_rounds"] >= window, stats
    if stats["backward_error_failed"]:
        # The iterate failed its componentwise check (solve_spd_settled, linalg.py:2687-2727):
        # the heat system was factored and solved again, and the Poisson system with it.
        assert stats["heat_refinements"] >= 1 and solver.nbytes > 0, stats
    else:
        assert solver.nbytes == 0 and stats["poisson_rounds"] > 0, stats
    worst, _ = ho.relative_error(got, oracle.distance(source))
    assert worst <= 1e-5, worst
    assert got[source[0]] == 0.0


# --- the displaced icosphere(9) ------------------------------------------------------------------


def _displaced_sphere_fields(cuda, vertices, faces, sources, precision):
    from ordito_amd import heat

    mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
    solver = mesh.heat_solver(method="meshlab", precision=precision)
    if precision == "mixed":
        solver.set_refinement(heat=1e-10, poisson=1e-10)
    assert solver.factor()
    fields = solver.geodesic([_gpu(cuda, np.array([s], np.int32)) for s in sources])
    return mesh, solver, fields.cpu().numpy()


@pytest.mark.timeout(900)
def test_displaced_sphere_full_size_fp64_and_mixed(cuda, release):
    from ordito_amd import creation

    vertices, faces = creation.displaced_sphere(9, lmax=16, amplitude=0.05, seed=0)
    n = len(vertices)
    assert n == 2_621_442
    sources = creation.random_sources(n, 16, seed=1)
    print(f"displaced icosphere(9): {n} vertices, FP64", flush=True)
    mesh, solver, fp64 = _displaced_sphere_fields(cuda, vertices, faces, sources, "fp64")
    # (M - tL) 1 = M 1 (L 1 = 0): the factored heat solve of the system's row sums returns ones.
    system = solver.operators[0]
    rows = cuda.repeat_interleave(
        cuda.arange(n, device="cuda"), (system.offsets[1:] - system.offsets[:-1]).long()
    )
    row_sums = cuda.zeros(n, dtype=cuda.float64, device="cuda").index_add_(0, rows, system.values)
    ones = cuda.empty((1, n), dtype=cuda.float64, device="cuda")
    solver.diffuse(row_sums.reshape(1, n).contiguous(), ones)
    assert float((ones - 1.0).abs().max()) <= 1e-10
    del mesh, solver, system, rows, row_sums, ones
    gc.collect()
    print("displaced icosphere(9): mixed", flush=True)
    mesh, solver, mixed = _displaced_sphere_fields(cuda, vertices, faces, sources, "mixed")
    del mesh, solver
    for q, s in enumerate(sources):
        assert fp64[q, s] == 0.0 and mixed[q, s] == 0.0
        assert fp64[q].min() >= 0.0 and mixed[q].min() >= 0.0
        assert np.all(np.isfinite(fp64[q]))
        worst, _ = ho.relative_error(mixed[q], fp64[q])
        assert worst <= 1e-4, (q, worst)


@pytest.mark.timeout(900)
def test_displaced_sphere_level7_matches_the_oracle(cuda, release):
    from ordito_amd import creation

    vertices, faces = creation.displaced_sphere(7, lmax=16, amplitude=0.05, seed=0)
    sources = creation.random_sources(len(vertices), 3, seed=1)
    print(f"displaced icosphere(7): {len(vertices)} vertices", flush=True)
    oracle = ho.FactoredOracle(vertices, faces, ordering="COLAMD")
    for precision, bar in (("fp64", 1e-9), ("mixed", 1e-8)):
        mesh, solver, fields = _displaced_sphere_fields(cuda, vertices, faces, sources, precision)
        for q, s in enumerate(sources):
            worst, _ = ho.relative_error(fields[q], oracle.distance(np.array([s])))
            assert worst <= bar, (precision, int(s), worst)
        del mesh, solver


# --- 256 packed meshes ---------------------------------------------------------------------------


@pytest.mark.timeout(900)
def test_packed_256_meshes_match_each_mesh_alone(cuda, release):
    from ordito_amd import creation, heat
    from ordito_amd.batch import MeshBatch

    rng = np.random.default_rng(0)
    meshes = creation.mesh_collection(256, 1e3, 2e5, rng)
    sizes = [len(meshes[i][0]) for i in range(256)]
    print(f"256 packed meshes: {sum(sizes)} vertices", flush=True)
    queries = []
    for i in range(256):
        count = min(int(rng.integers(1, 17)), sizes[i])
        queries.append((i, rng.choice(sizes[i], size=count, replace=False).astype(np.int32)))
    batch = MeshBatch([(_gpu(cuda, v), _gpu(cuda, f)) for v, f in meshes.values()])
    assert batch.factor()
    fields = batch.geodesic([(m, _gpu(cuda, s)) for m, s in queries])
    by_size = sorted(range(256), key=lambda i: sizes[i])
    sample = sorted({by_size[0], by_size[1], by_size[128], by_size[-1], int(rng.integers(256))})
    for i in sample:
        v, f = meshes[i]
        tm = heat.Trimesh(_gpu(cuda, v), _gpu(cuda, f))
        solver = tm.heat_solver(method="meshlab")
        assert solver.factor()
        alone = solver.geodesic([_gpu(cuda, queries[i][1])])[0].cpu().numpy()
        got = fields[i].cpu().numpy()
        worst, _ = ho.relative_error(got, alone)
        assert worst <= 1e-10, (i, sizes[i], worst)
        if sizes[i] <= 20_000:
            worst, _ = ho.relative_error(got, ho.geodesic(v, f, queries[i][1]).distance)
            assert worst <= 1e-9, (i, sizes[i], worst)
        del tm, solver
    batch.close()


# --- 64 served meshes ----------------------------------------------------------------------------


@pytest.mark.timeout(900)
def test_served_64_meshes_under_poisson_arrivals(cuda, release):
    import time

    from ordito_amd import creation
    from ordito_amd.serve import GeodesicServer

    rng = np.random.default_rng(0)
    meshes = creation.mesh_collection(64, 1e4, 1e6, rng)
    sizes = {i: len(v) for i, (v, _) in meshes.items()}
    print(f"64 served meshes: {sum(sizes.values())} vertices", flush=True)
    server = GeodesicServer(meshes, budget_bytes=200e9, workers=4)
    server.warm()
    print("served: warm", flush=True)
    smallest = min(sizes, key=sizes.get)
    sampled, futures = [], []
    start = time.perf_counter()
    arrival = start
    while arrival - start < 3.0:  # 3 s of arrivals at 2,000 queries/s
        arrival += rng.exponential(1.0 / 2000.0)
        delay = arrival - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        m = int(rng.integers(64))
        sources = rng.choice(sizes[m], size=min(int(rng.integers(1, 33)), sizes[m]), replace=False)
        future = server.submit(m, sources.astype(np.int32))
        futures.append(future)
        if (m == smallest or len(sampled) < 6) and len(sampled) < 8:
            sampled.append((m, sources.astype(np.int32), future))
    for future in futures:
        future.result(timeout=300)
    stats = server.stats(since=start)
    assert stats["completed"] == len(futures)
    assert stats["batches"] < len(futures)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0.0
    for m, sources, future in sampled:
        mixed_in = future.result().cpu().numpy()
        alone = server.submit(m, sources).result(timeout=300).cpu().numpy()
        worst, _ = ho.relative_error(mixed_in, alone)
        assert worst <= 1e-10, (m, sizes[m], worst)
        if m == smallest:
            v, f = meshes[m]
            worst, _ = ho.relative_error(mixed_in, ho.geodesic(v, f, sources).distance)
            assert worst <= 1e-9, (m, worst)
    server.close()


# --- iterated solves past one workgroup ----------------------------------------------------------


@pytest.fixture(scope="module")
def torus_10k():
    from ordito_amd import creation

    vertices, faces = creation.torus(128, 80, 1.0, 0.3, jitter=0.1, seed=9)
    assert len(vertices) > 8192
    return vertices, faces, ho.FactoredOracle(vertices, faces)


@pytest.mark.parametrize("columns", [1, 7, 64])
def test_multiblock_iterated_geodesic_matches_the_oracle(cuda, torus_10k, columns):
    from ordito_amd import heat

    vertices, faces, oracle = torus_10k
    print(f"iterated torus 128x80, {columns} columns", flush=True)
    sources = np.random.default_rng(columns).choice(len(vertices), size=columns, replace=False)
    mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
    solver = mesh.heat_solver(method="meshlab")
    got = solver.geodesic([_gpu(cuda, np.array([s], np.int32)) for s in sources]).cpu().numpy()
    assert solver.nbytes == 0 and solver.last_stats["heat_rounds"] > 0
    assert solver.last_stats["poisson_rounds"] > 0
    for q, s in enumerate(sources):
        worst, _ = ho.relative_error(got[q], oracle.distance(np.array([s])))
        assert worst <= 1e-5, (q, int(s), worst)


def test_multiblock_diffuse_columns_match_each_alone(cuda, torus_10k):
    """Three right-hand sides diffused together by the settled CG: each within the settle rule's
    agreement (1e-6 of its largest entry) of the same column diffused alone, and of the direct
    solve; the shared settle monitor may run a column a few rounds past its own stop."""
    from ordito_amd import heat

    vertices, faces, oracle = torus_10k
    n = len(vertices)
    rhs = np.zeros((3, n))
    rhs[0, 11] = 1.0
    rhs[1, [200, 4000]] = 1.0
    rhs[2] = np.random.default_rng(3).random(n)
    together = cuda.empty((3, n), dtype=cuda.float64, device="cuda")
    mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
    mesh.heat_solver(method="meshlab").diffuse(_gpu(cuda, rhs), together)
    together = together.cpu().numpy()
    for q in range(3):
        alone_mesh = heat.Trimesh(_gpu(cuda, vertices), _gpu(cuda, faces))
        alone = cuda.empty((1, n), dtype=cuda.float64, device="cuda")
        alone_mesh.heat_solver(method="meshlab").diffuse(_gpu(cuda, rhs[q : q + 1]), alone)
        alone = alone.cpu().numpy()[0]
        scale = np.max(np.abs(alone))
        assert np.max(np.abs(together[q] - alone)) <= 1e-6 * scale, q
        exact = oracle.heat_lu.solve(rhs[q])
        assert np.max(np.abs(together[q] - exact)) <= 1e-5 * np.max(np.abs(exact)), q
