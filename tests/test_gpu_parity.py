"""
Parity of the HIP path against the FP64 oracle, on the GPU through the C ABI.

Bars (each written in its test):
- face terms: bit-exact (float32, the same operations in the same order, no contraction);
- assembled operators: bit-exact off the diagonal and in ``L``/``-L``; ``M - tL`` to 1e-14
  relative (``t`` comes from a parallel reduction whose summation order differs from numpy's);
- factored FP64 distance: 1e-9 of the range (both are direct FP64 solves of the same systems);
- iterated (settled CG) distance: 1e-5 of the range (the settle rule stops at ~1e-6 per entry,
  ordito's iterated-vs-factored agreement is 4e-6 of the range, ``tests/test_heat.py:393-398``);
- mixed-precision distance: 1e-4 of the range (the served accuracy bar);
- unit-gradient divergence: 1e-12 of each column's largest entry, exact under sign and power-of-two
  scaling.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import heat_oracle as ho
from tests.conftest import mesh

pytestmark = pytest.mark.gpu


def _tensor(cuda, array):
    return cuda.from_numpy(np.ascontiguousarray(array)).cuda()


@pytest.mark.parametrize("name", ["icosphere3", "torus", "grid", "bumpy", "saddle_graded"])
def test_operators_match_oracle(cuda, name):
    from ordito_amd import heat

    vertices, faces = mesh(name)
    ops = heat.heat_operators(_tensor(cuda, vertices), _tensor(cuda, faces), method="meshlab")
    heat_sys, jacobi, laplacian, poisson, _, cot, normals, areas = ops
    ref = ho.assemble(vertices, faces, None)
    # Face terms: bit-exact.
    assert np.array_equal(cot.cpu().numpy(), ref.terms.cot)
    assert np.array_equal(normals.cpu().numpy(), ref.terms.normals)
    assert np.array_equal(areas.cpu().numpy(), ref.terms.areas)
    # Pattern: identical.
    assert np.array_equal(laplacian.offsets.cpu().numpy(), ref.pattern.indptr)
    assert np.array_equal(laplacian.columns.cpu().numpy(), ref.pattern.indices)
    # L and -L: bit-exact; M - tL within the time's rounding.
    assert np.array_equal(laplacian.values.cpu().numpy(), ref.laplacian)
    assert np.array_equal(poisson.values.cpu().numpy(), ref.poisson)
    got = heat_sys.values.cpu().numpy()
    assert np.allclose(got, ref.heat, rtol=1e-14, atol=0.0)


@pytest.mark.parametrize("name", ["icosphere4", "torus", "grid", "bumpy", "two_bodies"])
def test_factored_fp64_matches_oracle(cuda, name):
    from ordito_amd import heat

    vertices, faces = mesh(name)
    sources = np.array([0, len(vertices) // 2], dtype=np.int32)
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    assert solver.factor()
    got = solver.geodesic([_tensor(cuda, sources)])[0].cpu().numpy()
    expected = ho.geodesic(vertices, faces, sources).distance
    worst, _ = ho.relative_error(got, expected)
    assert worst <= 1e-9, worst


@pytest.mark.parametrize(
    "name, method",
    [("icosphere4", "meshlab"), ("bumpy", "meshlab"), ("two_bodies", "meshlab"),
     ("saddle_graded", "meshlab"), ("hub_disk", "meshlab"), ("hub_disk", "igl")],
)
def test_wide_batches_match_the_oracle(cuda, name, method):
    """40 queries in one factored call (the wide form of every by-rows kernel: a wave a row in the
    divergence, over a hub row of 400 corners on hub_disk): single sources spread over the mesh,
    two multi-source queries and an empty one, each against the factored oracle to 1e-9 of its
    range."""
    from ordito_amd import creation, heat

    vertices, faces = mesh(name)
    n = len(vertices)
    picks = creation.random_sources(n, 37, seed=11)
    picks[0] = 0
    sets = [np.array([p], np.int32) for p in picks]
    sets += [np.array([1, n // 3, n - 1], np.int32), np.array([n // 2, n // 2 + 1], np.int32)]
    sets.append(np.zeros(0, np.int32))
    print(f"wide batch: {name} {method}, {len(sets)} queries", flush=True)
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method=method)
    assert solver.factor()
    got = solver.geodesic([_tensor(cuda, q) for q in sets]).cpu().numpy()
    oracle = ho.FactoredOracle(vertices, faces, method=method)
    for q, sources in enumerate(sets):
        if sources.size == 0:
            assert not got[q].any()
            continue
        worst, _ = ho.relative_error(got[q], oracle.distance(sources))
        assert worst <= 1e-9, (q, worst)


def test_hub_disk_iterated_matches_oracle(cuda):
    """The one-shot path on a mesh with a vertex of degree 400: its long row in the settled CG,
    the Chebyshev Poisson solve and the divergence."""
    from ordito_amd import heat

    vertices, faces = mesh("hub_disk")
    for sources in (np.array([0], np.int32), np.array([4321], np.int32)):
        got = heat.heat_geodesic(vertices, faces, sources)
        worst, _ = ho.relative_error(got, ho.geodesic(vertices, faces, sources).distance)
        assert worst <= 1e-5, (sources, worst)


@pytest.mark.parametrize("name", ["icosphere4", "torus", "grid", "bumpy", "two_bodies"])
def test_iterated_matches_oracle(cuda, name):
    from ordito_amd import heat

    vertices, faces = mesh(name)
    sources = np.array([3], dtype=np.int32)
    got = heat.heat_geodesic(vertices, faces, sources)
    expected = ho.geodesic(vertices, faces, sources).distance
    worst, _ = ho.relative_error(got, expected)
    assert worst <= 1e-5, worst


def test_graded_saddle_falls_back_to_the_factorization(cuda):
    from ordito_amd import heat

    vertices, faces = mesh("saddle_graded")
    sources = np.array([0], dtype=np.int32)
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    got = heat.heat_geodesic(trimesh, _tensor(cuda, sources), method="meshlab").cpu().numpy()
    expected = ho.geodesic(vertices, faces, sources).distance
    worst, _ = ho.relative_error(got, expected)
    assert worst <= 1e-5, worst
    assert solver.ill_conditioned is (solver.last_stats["backward_error_failed"] == 1)


@pytest.mark.parametrize("name", ["icosphere4", "torus"])
def test_mixed_precision_matches_oracle(cuda, name):
    from ordito_amd import heat

    vertices, faces = mesh(name)
    sets = [np.array([s], dtype=np.int32) for s in (0, 7, 100)]
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab", precision="mixed")
    assert solver.factor()
    got = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
    expected = ho.geodesic_columns(vertices, faces, sets)
    for q in range(len(sets)):
        worst, _ = ho.relative_error(got[q], expected[q])
        assert worst <= 1e-4, (q, worst)


@pytest.mark.parametrize("name", ["icosphere3", "torus", "grid", "saddle_graded"])
def test_divergence_matches_oracle(cuda, name):
    """``hg_unit_gradient_divergence`` against the oracle's face divergence: 1e-12 of the column's
    largest entry (the kernel normalizes with an exact power-of-two scale and a reciprocal square
    root, the oracle with divisions, so the two agree to a few ulps per term). The direction is
    scale-free, so a column scaled by 2^-600 (every product stays a normal number) must give the
    same bits, a negated column the negated bits, and a zero column zeros."""
    import math

    from ordito_amd import _lib, heat

    vertices, faces = mesh(name)
    n = len(vertices)
    ref = ho.assemble(vertices, faces, None)
    p = vertices.astype(np.float64)
    u = np.exp(-np.sum((p - p[0]) ** 2, axis=1) / (0.5 * np.ptp(p, axis=0).max()) ** 2)
    rng = np.random.default_rng(3)
    noisy = rng.random(n)
    columns = np.stack([u, noisy, u * 2.0**-600, -noisy, np.zeros(n)])
    handle = heat._Handle(_tensor(cuda, vertices), _tensor(cuda, faces))
    try:
        _lib.call("hg_operators", handle.ptr, math.nan, 0, 1)
        d_heat = _tensor(cuda, columns.reshape(-1))
        d_out = cuda.zeros_like(d_heat)
        _lib.call("hg_unit_gradient_divergence", handle.ptr, d_heat.data_ptr(), d_out.data_ptr(),
                  len(columns))
        got = d_out.cpu().numpy().reshape(len(columns), n)
    finally:
        handle.close()
    for q in (0, 1):
        expected = ho.face_unit_divergence(vertices, faces, ref.terms, columns[q])
        err = np.abs(got[q] - expected).max()
        assert err <= 1e-12 * np.abs(expected).max(), (q, err)
    assert np.array_equal(got[2], got[0])
    assert np.array_equal(got[3], -got[1])
    assert not got[4].any()


def test_rent_or_buy_factors_on_the_second_call(cuda):
    from ordito_amd import heat

    vertices, faces = mesh("icosphere3")
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    expected = ho.geodesic(vertices, faces, np.array([5])).distance
    results = []
    for call in range(3):
        results.append(heat.heat_geodesic(trimesh, _tensor(cuda, np.array([5], np.int32)), method="meshlab"))
        if call == 0:
            assert solver.nbytes == 0
        else:
            assert solver.nbytes > 0
    for r in results:
        worst, _ = ho.relative_error(r.cpu().numpy(), expected)
        assert worst <= 1e-5
    solver.release()
    assert solver.nbytes == 0


def test_out_of_range_source_is_a_value_error(cuda):
    from ordito_amd import heat

    vertices, faces = mesh("icosphere3")
    with pytest.raises(ValueError):
        heat.heat_geodesic(vertices, faces, np.array([len(vertices)], dtype=np.int32))
    with pytest.raises(ValueError):
        heat.heat_geodesic(vertices, faces, np.array([-1], dtype=np.int32))


def test_empty_inputs_give_zeros(cuda):
    from ordito_amd import heat

    vertices, faces = mesh("icosphere3")
    out = heat.heat_geodesic(vertices, faces, np.zeros(0, dtype=np.int32))
    assert out.shape == (len(vertices),) and not out.any()
    out = heat.heat_geodesic(vertices, np.zeros(0, dtype=np.int32), np.array([0], np.int32))
    assert out.shape == (len(vertices),) and not out.any()


def test_reference_known_answers(cuda):
    """ordito's docstring answers: icosphere(3) pole to pole 3.09, source 0.0 (heat.py:529-534);
    maximum 3.1 from vertex 0 (mesh.py:2249-2253)."""
    from ordito_amd import creation, heat

    vertices, faces = creation.icosphere(3)
    north, south = int(np.argmax(vertices[:, 2])), int(np.argmin(vertices[:, 2]))
    d = heat.heat_geodesic(vertices, faces, np.array([north], np.int32))
    assert d[north] == 0.0
    assert f"{d[south]:.2f}" == "3.09"
    d0 = heat.heat_geodesic(vertices, faces, np.array([0], np.int32))
    assert f"{d0.max():.1f}" == "3.1"


# --- method="igl" --------------------------------------------------------------------------------


@pytest.mark.parametrize("name", ["grid", "saddle_graded", "icosphere3", "bumpy"])
def test_igl_operators_match_oracle(cuda, name):
    """The Voronoi mass: each corner's share is formed in float64 from the float32 corners with
    the oracle's operations in its order and rounded once to float32, so the two agree to one
    float32 ulp of a share (2^-23 of a vertex's mass) whatever the float64 library rounds; the
    Laplacian is bit-exact and the heat system's off-diagonal within the time's rounding."""
    from ordito_amd import heat

    vertices, faces = mesh(name)
    ops = heat.heat_operators(_tensor(cuda, vertices), _tensor(cuda, faces), method="igl")
    heat_sys, _, laplacian, _, _, _, _, _ = ops
    ref = ho.assemble(vertices, faces, None, method="igl")
    assert np.array_equal(laplacian.values.cpu().numpy(), ref.laplacian)
    n = len(vertices)
    got = heat_sys.values.cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(ref.pattern.indptr))
    diagonal = ref.pattern.indices == rows
    assert np.allclose(got[~diagonal], ref.heat[~diagonal], rtol=1e-14, atol=0.0)
    assert np.allclose(got[diagonal], ref.heat[diagonal], rtol=2.0**-23, atol=0.0)


@pytest.mark.parametrize("name", ["grid", "saddle_graded", "icosphere4", "two_bodies"])
def test_igl_geodesic_matches_oracle(cuda, name):
    """heat_geodesic(method="igl"), from the middle and from a rim vertex, iterated on the first
    call (1e-5 of the range) and through the factors on the later ones (1e-9)."""
    from ordito_amd import heat

    vertices, faces = mesh(name)
    n = len(vertices)
    rim = ho.rim_vertices(ho.operator_pattern(n, faces))
    sources = [n // 2 + 17] + ([int(np.flatnonzero(rim)[0])] if rim.any() else [])
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="igl")
    assert solver.rim_size == int(rim.sum())
    for s in sources:
        expected = ho.geodesic(vertices, faces, np.array([s]), method="igl").distance
        once = heat.heat_geodesic(vertices, faces, np.array([s], np.int32), method="igl")
        worst, _ = ho.relative_error(once, expected)
        assert worst <= 1e-5, (s, worst)
        for call in range(2):
            got = heat.heat_geodesic(trimesh, _tensor(cuda, np.array([s], np.int32)), method="igl")
            worst, _ = ho.relative_error(got.cpu().numpy(), expected)
            assert worst <= (1e-5 if call == 0 and s == sources[0] else 1e-9), (s, call, worst)
    if rim.any():
        assert solver.last_stats["pinned_solves"] >= 1
        assert solver._factored() & 4
        meshlab = ho.geodesic(vertices, faces, np.array([sources[0]])).distance
        assert np.abs(got.cpu().numpy() - meshlab).max() > 1e-3 * np.ptp(meshlab)


# --- method="potpourri3d" ------------------------------------------------------------------------


@pytest.mark.parametrize("name", ["grid", "icosphere4", "two_bodies", "saddle_graded"])
def test_potpourri3d_geodesic_matches_oracle(cuda, name):
    """heat_geodesic(method="potpourri3d"): the unique-edge time and the Poisson step regularized
    with 1e-6 I (no null space, nothing pinned), iterated on the first call (1e-5 of the range)
    and through the heat and regularized factors on the later ones (1e-9); the source reads 0 and
    the answer differs from MeshLab's."""
    from ordito_amd import heat

    vertices, faces = mesh(name)
    n = len(vertices)
    s = n // 2 + 17
    expected = ho.geodesic(vertices, faces, np.array([s]), method="potpourri3d").distance
    once = heat.heat_geodesic(vertices, faces, np.array([s], np.int32), method="potpourri3d")
    worst, _ = ho.relative_error(once, expected)
    assert worst <= 1e-5, worst
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="potpourri3d")
    for call in range(3):
        got = heat.heat_geodesic(
            trimesh, _tensor(cuda, np.array([s], np.int32)), method="potpourri3d"
        ).cpu().numpy()
        worst, _ = ho.relative_error(got, expected)
        assert worst <= (1e-5 if call == 0 else 1e-9), (call, worst)
    assert got[s] == 0.0
    assert solver._factored() & 8 and not solver._factored() & 2
    meshlab = ho.geodesic(vertices, faces, np.array([s])).distance
    assert np.abs(got - meshlab).max() > 1e-9 * np.ptp(meshlab)


def test_diffuse_pinned_holds_the_rim_at_zero(cuda):
    """HeatSolver.diffuse_pinned (the reference's check, tests/test_heat.py:295-329): zero on the
    rim and the interior system's direct solve elsewhere, iterated on the first call (rtol 1e-6,
    the reference's bar) and factored on the second (1e-10); a closed mesh reports no rim and
    leaves the solution alone."""
    from ordito_amd import heat

    vertices, faces = mesh("saddle_graded")
    n = len(vertices)
    t, _ = ho.resolve_time(vertices, faces, None)  # the solver's time rule
    ref = ho.assemble(vertices, faces, t, method="meshlab")
    rim = ho.rim_vertices(ref.pattern)
    rhs = np.zeros(n)
    rhs[n // 2 + 17] = 1.0
    expected = ho.diffuse_pinned(ref, rhs)
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    for call, rtol in enumerate((1e-6, 1e-10)):
        solution = cuda.zeros(n, dtype=cuda.float64, device="cuda")
        assert solver.diffuse_pinned(_tensor(cuda, rhs), solution)
        got = solution.cpu().numpy()
        assert np.array_equal(got[rim], np.zeros(int(rim.sum()))), call
        assert np.allclose(got, expected, rtol=rtol, atol=1e-12 * expected.max()), call
    assert solver._factored() & 4

    closed_v, closed_f = mesh("icosphere3")
    closed_mesh = heat.Trimesh(_tensor(cuda, closed_v), _tensor(cuda, closed_f))
    closed = closed_mesh.heat_solver()
    untouched = cuda.full((len(closed_v),), 7.0, dtype=cuda.float64, device="cuda")
    assert not closed.diffuse_pinned(cuda.zeros_like(untouched), untouched)
    assert bool((untouched == 7.0).all())


@pytest.mark.parametrize("width", [1, 7, 16, 33, 64, 70])
def test_batch_widths_give_each_query_its_own_answer(cuda, width):
    """One call of `width` source sets through the factors: every column equals the same source
    set solved alone to 1e-12 of the range (the columns never mix: per-column sums in fixed
    order, whatever the kernels' lane packing at that width), and the oracle to 1e-9."""
    from ordito_amd import heat

    vertices, faces = mesh("bumpy")
    n = len(vertices)
    rng = np.random.default_rng(width)
    sets = [rng.integers(0, n, size=1 + q % 3).astype(np.int32) for q in range(width)]
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    assert solver.factor()
    together = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
    oracle = ho.FactoredOracle(vertices, faces)
    for q in (0, width // 2, width - 1):
        alone = solver.geodesic([_tensor(cuda, sets[q])])[0].cpu().numpy()
        span = np.ptp(alone)
        assert np.abs(together[q] - alone).max() <= 1e-12 * span, q
        worst, _ = ho.relative_error(together[q], oracle.distance(sets[q]))
        assert worst <= 1e-9, (q, worst)


def test_memory_budget_refuses_and_the_solves_march(cuda):
    """A factorization whose peak exceeds the memory budget is refused, not an error
    (OperatorFactorization.factor returns False, linalg.py:2393-2420): factor() is False, nothing
    is held, and every later heat solve marches band by band (linalg.py:2677-2682) and every
    Poisson solve runs under the multigrid V-cycle, to the iterated bar; lifting the budget factors
    again. On the graded saddle, whose settled iterate fails its backward-error check, the probed
    refusal marches before any iteration: the mesh is never found ill-conditioned."""
    from ordito_amd import heat

    vertices, faces = mesh("torus")
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    solver.set_memory_budget(1)
    assert not solver.factor()
    assert solver.nbytes == 0 and solver.unfactorable
    expected = ho.geodesic(vertices, faces, np.array([3])).distance
    source = _tensor(cuda, np.array([3], np.int32))
    for _ in range(3):
        got = heat.heat_geodesic(trimesh, source, method="meshlab").cpu().numpy()
        worst, _ = ho.relative_error(got, expected)
        assert worst <= 1e-5, worst
        stats = solver.last_stats
        assert stats["heat_marched"] == 1 and stats["poisson_multigrid"] == 1, stats
        assert stats["heat_rounds"] > 0 and solver.nbytes == 0
    solver.set_memory_budget(0)
    assert solver.factor() and not solver.unfactorable
    got = heat.heat_geodesic(trimesh, source, method="meshlab").cpu().numpy()
    worst, _ = ho.relative_error(got, expected)
    assert worst <= 1e-9, worst

    sv, sf = mesh("saddle_graded")
    saddle = heat.Trimesh(_tensor(cuda, sv), _tensor(cuda, sf))
    graded = saddle.heat_solver(method="meshlab")
    graded.set_memory_budget(1)
    d = heat.heat_geodesic(saddle, _tensor(cuda, np.array([0], np.int32)), method="meshlab")
    got = d.cpu().numpy()
    worst, _ = ho.relative_error(got, ho.geodesic(sv, sf, np.array([0])).distance)
    assert worst <= 1e-5, worst
    assert graded.last_stats["heat_marched"] == 1, graded.last_stats
    assert not graded.ill_conditioned and graded.nbytes == 0


def test_a_failed_speculated_round_solves_again(cuda):
    """Through FP64 factors the first refinement round runs on without waiting for its test, whose
    verdict is read at the call's end. A heat tolerance no first round meets (1e-15 componentwise)
    makes the call solve again, refining round by round: the answer still matches the oracle and
    the statistics count the extra rounds."""
    from ordito_amd import heat

    vertices, faces = mesh("bumpy")
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    assert solver.factor()
    sets = [np.array([s], dtype=np.int32) for s in (0, 17, 300)]
    expected = ho.geodesic_columns(vertices, faces, sets)
    solver.set_refinement(heat=1e-15)
    got = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
    assert solver.last_stats["heat_refinements"] > 1
    for q in range(len(sets)):
        worst, _ = ho.relative_error(got[q], expected[q])
        assert worst <= 1e-9, (q, worst)
    solver.set_refinement(heat=1e-12)
    again = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
    assert solver.last_stats["heat_refinements"] == 1
    assert np.abs(again - got).max() <= 1e-9 * np.ptp(expected)


def test_geodesic_flat_equals_geodesic_and_checks_its_offsets(cuda):
    from ordito_amd import heat

    vertices, faces = mesh("icosphere3")
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method="meshlab")
    assert solver.factor()  # both calls through the factors
    sets = [np.array([0, 5], np.int32), np.array([], np.int32), np.array([17], np.int32)]
    listed = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
    flat = _tensor(cuda, np.concatenate(sets))
    got = solver.geodesic_flat(flat, [0, 2, 2, 3]).cpu().numpy()
    assert np.array_equal(got, listed)
    with pytest.raises(ValueError):
        solver.geodesic_flat(flat, [0, 2, 4])  # past the end of the sources
    with pytest.raises(ValueError):
        solver.geodesic_flat(flat, [0, 3, 2, 3])  # decreasing
    with pytest.raises(ValueError):
        solver.geodesic_flat(flat.long(), [0, 3])  # not int32


@pytest.mark.parametrize("name", ["torus", "grid"])
def test_graph_replays_equal_the_eager_path(cuda, name):
    """The factored FP64 path replayed as a HIP graph gives the eager path's bits, call after call:
    new sources and a new output buffer each replay, several column counts (one graph each),
    eager calls and synchronizations in between, and the igl path with its rim-pinned solve."""
    from ordito_amd import heat

    vertices, faces = mesh(name)
    n = len(vertices)
    method = "igl" if name == "grid" else "meshlab"
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method=method)
    assert solver.factor()
    solver.geodesic([_tensor(cuda, np.array([1], np.int32))])  # factors the pinned system at igl
    rng = np.random.default_rng(5)
    for k in (1, 5, 64, 5, 1, 64):
        sets = [rng.integers(0, n, size=1 + q % 4).astype(np.int32) for q in range(k)]
        solver.set_graphs(False)
        eager = solver.geodesic([_tensor(cuda, s) for s in sets]).cpu().numpy()
        solver.set_graphs(True)
        for _ in range(2):
            replay = solver.geodesic([_tensor(cuda, s) for s in sets])
            cuda.cuda.synchronize()
            assert np.array_equal(replay.cpu().numpy(), eager), k
        cached, broken = solver.graph_state()
        assert cached >= 1 and not broken
    expected = ho.geodesic(vertices, faces, sets[0], method=method).distance
    worst, _ = ho.relative_error(eager[0], expected)
    assert worst <= 1e-9, worst


@pytest.mark.parametrize(
    "name, method",
    [("icosphere4", "meshlab"), ("bumpy", "meshlab"), ("pinched_grid", "meshlab"),
     ("pinched_grid", "igl")],
)
def test_use_robust_matches_the_oracle(cuda, name, method):
    """use_robust (heat.py:229-243): the half-cotangent table from the mollified side lengths bit
    for bit, the constant itself, and the factored and one-shot distances against the oracle built
    on the same weights (1e-9 and 1e-5 of the range). The pinched grid's needle, 6e6 by the
    positions, weighs about 1e2 mollified; a clean mesh needs no constant."""
    from ordito_amd import heat

    vertices, faces = mesh(name)
    lengths, delta = ho.mollify_intrinsic(vertices, faces)
    assert (delta > 0.0) == (name == "pinched_grid")
    print(f"use_robust: {name} {method}", flush=True)
    trimesh = heat.Trimesh(_tensor(cuda, vertices), _tensor(cuda, faces))
    solver = trimesh.heat_solver(method=method, use_robust=True)
    cot = solver.operators[5].cpu().numpy()
    assert solver.mollify_delta == delta
    assert np.array_equal(cot, ho.intrinsic_half_cotangents(lengths))
    assert solver.factor()
    sets = [np.array([0], np.int32), np.array([len(vertices) // 2], np.int32),
            np.array([3, len(vertices) - 5], np.int32)]
    got = solver.geodesic([_tensor(cuda, q) for q in sets]).cpu().numpy()
    oracle = ho.FactoredOracle(vertices, faces, method=method, use_robust=True)
    for q, sources in enumerate(sets):
        worst, _ = ho.relative_error(got[q], oracle.distance(sources))
        assert worst <= 1e-9, (q, worst)
    one_shot = heat.heat_geodesic(vertices, faces, sets[1], method=method, use_robust=True)
    worst, _ = ho.relative_error(one_shot, oracle.distance(sets[1]))
    assert worst <= 1e-5, worst


def test_use_robust_at_potpourri3d_is_refused(cuda):
    """The reference's use_robust at "potpourri3d" is geometry-central's intrinsic Delaunay
    method (heat.py:425-430), not built here: refused before any work, not silently replaced."""
    from ordito_amd import heat

    vertices, faces = mesh("icosphere3")
    with pytest.raises(NotImplementedError):
        heat.heat_geodesic(vertices, faces, np.array([0], np.int32), method="potpourri3d",
                           use_robust=True)
