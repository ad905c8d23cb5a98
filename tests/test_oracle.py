"""
Pins of the FP64 oracle (CPU). Each check is one a plausible mistake in the restatement would fail:
closed forms, identities the discretization satisfies exactly, brute force on tiny meshes, the
reference's own known answers and the bounds its tests hold the method to.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import heat_oracle as ho
from tests.conftest import mesh


def _tri(points):
    return np.asarray(points, dtype=np.float32), np.array([0, 1, 2], dtype=np.int32)


# --- face terms ---------------------------------------------------------------------------------


def test_equilateral_half_cotangents_are_one_over_two_root_three():
    v, f = _tri([[0, 0, 0], [1, 0, 0], [0.5, math.sqrt(3) / 2, 0]])
    terms = ho.face_terms(v, f)
    assert np.allclose(terms.cot, 1.0 / (2.0 * math.sqrt(3.0)), rtol=1e-6)
    assert np.allclose(terms.normals, [[0, 0, 1]])
    assert np.isclose(terms.areas[0], math.sqrt(3) / 4, rtol=1e-6)


def test_right_angle_half_cotangent_is_zero_and_others_are_half_cotangents():
    v, f = _tri([[0, 0, 0], [2, 0, 0], [0, 1, 0]])
    c = ho.face_terms(v, f).cot[0]
    assert c[0] == 0.0  # right angle at corner 0
    assert np.isclose(c[1], 0.5 * (2.0 / 1.0), rtol=1e-6)  # angle at (2,0): cot = 2/1
    assert np.isclose(c[2], 0.5 * (1.0 / 2.0), rtol=1e-6)


def test_degenerate_face_has_zero_terms():
    v, f = _tri([[0, 0, 0], [1, 0, 0], [2, 0, 0]])
    terms = ho.face_terms(v, f)
    assert not terms.cot.any() and not terms.normals.any() and terms.areas[0] == 0.0


@pytest.mark.parametrize("name", ["bumpy", "saddle_graded", "torus"])
def test_face_terms_match_float64_recomputation(name):
    v, f = mesh(name)
    terms = ho.face_terms(v, f)
    p = v.astype(np.float64)
    tri = f.reshape(-1, 3)
    a, b, c = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    cross = np.cross(b - a, c - a)
    norm = np.linalg.norm(cross, axis=1)
    assert np.allclose(terms.areas, 0.5 * norm, rtol=1e-5)
    assert np.allclose(terms.normals, cross / norm[:, None], atol=2e-4)
    cot0 = np.einsum("ij,ij->i", b - a, c - a) / (2 * norm)
    assert np.allclose(terms.cot[:, 0], cot0, rtol=1e-3, atol=1e-5)


# --- operators -----------------------------------------------------------------------------------


def _dense_cotmatrix(v, f):
    """Brute force: L_ij += 1/2 cot(angle opposite ij) per face, float64, diagonal = -row sum."""
    p = v.astype(np.float64)
    n = p.shape[0]
    L = np.zeros((n, n))
    for a, b, c in f.reshape(-1, 3):
        for i, j, k in ((a, b, c), (b, c, a), (c, a, b)):
            e1, e2 = p[i] - p[k], p[j] - p[k]
            cot = np.dot(e1, e2) / np.linalg.norm(np.cross(e1, e2))
            L[i, j] += 0.5 * cot
            L[j, i] += 0.5 * cot
    L -= np.diag(L.sum(axis=1))
    return L


def test_laplacian_matches_brute_force_on_the_icosahedron():
    from ordito_amd import creation

    v, f = creation.icosphere(0)
    ops = ho.assemble(v, f, 0.1)
    assert np.allclose(ops.matrix(ops.laplacian).toarray(), _dense_cotmatrix(v, f), atol=1e-6)


@pytest.mark.parametrize("name", ["icosphere3", "bumpy", "grid", "saddle_graded"])
def test_operator_identities(name):
    v, f = mesh(name)
    ops = ho.assemble(v, f, None)
    L = ops.matrix(ops.laplacian)
    assert (abs(L - L.T) > 0).nnz == 0  # symmetric, bit for bit
    assert np.max(np.abs(L @ np.ones(L.shape[0]))) <= 1e-12 * np.max(np.abs(L.diagonal()))
    assert np.isclose(ops.mass.sum(), ops.terms.areas.astype(np.float64).sum(), rtol=1e-12)
    H = ops.matrix(ops.heat)
    ones = np.ones(H.shape[0])
    assert np.allclose(H @ ones, ops.mass, rtol=1e-9, atol=1e-15)  # (M - tL) 1 = M 1
    assert np.array_equal(ops.matrix(ops.poisson).data, -L.data)
    # heat = M - tL for one t: read t off an off-diagonal entry (tests/test_heat.py:122-188).
    rows = np.repeat(np.arange(L.shape[0]), np.diff(L.indptr))
    k = np.flatnonzero((L.indices != rows) & (L.data != 0))[0]
    assert np.isclose(-ops.heat[k] / ops.laplacian[k], ops.t, rtol=1e-12)


def test_default_time_is_the_squared_face_edge_mean_and_unique_mean_on_a_closed_mesh():
    v, f = mesh("icosphere3")
    ops = ho.assemble(v, f, None)
    h = ho.mean_face_edge_length(v, f)
    assert np.isclose(ops.t, h * h, rtol=1e-12)
    # Closed manifold: every edge has two faces, so the face-edge mean is the unique-edge mean.
    tri = f.reshape(-1, 3)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    unique = np.unique(edges, axis=0)
    p = v.astype(np.float64)
    mean_unique = np.linalg.norm(p[unique[:, 1]] - p[unique[:, 0]], axis=1).mean()
    assert np.isclose(h, mean_unique, rtol=1e-6)


def _flat_grid():
    from ordito_amd import creation

    # Jitter small enough that no triangle inverts (linear precision needs a consistent winding).
    return creation.planar_grid(30, 24, 1.0, 0.05, 4)


def test_linear_precision_at_interior_vertices_of_a_planar_mesh():
    v, f = _flat_grid()
    ops = ho.assemble(v, f, None)
    L = ops.matrix(ops.laplacian)
    interior = np.ones(v.shape[0], dtype=bool)
    tri = f.reshape(-1, 3)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    uniq, count = np.unique(edges, axis=0, return_counts=True)
    interior[uniq[count == 1].ravel()] = False
    for axis in (0, 1):
        lx = L @ v[:, axis].astype(np.float64)
        scale = np.abs(L).max() * np.abs(v[:, axis]).max()
        assert np.max(np.abs(lx[interior])) <= 1e-5 * scale


def test_a_transposed_cotangent_column_breaks_linear_precision():
    """Planted in the oracle: pairing each edge with a neighbouring corner's weight fails the check
    above by orders of magnitude, so the check can see such a mistake."""
    v, f = _flat_grid()
    ops = ho.assemble(v, f, None)
    terms = ops.terms
    rolled = np.roll(terms.cot, 1, axis=1)
    bad = ho.FaceTerms(rolled, terms.normals, terms.areas)
    cot_flat = bad.cot.reshape(-1).astype(np.float64)
    pat = ops.pattern
    vals = np.zeros(pat.indices.size)
    counts = np.diff(pat.corner_ptr)
    for k in range(int(counts.max())):
        has = counts > k
        vals[has] += cot_flat[pat.corner_slots[pat.corner_ptr[:-1][has] + k]]
    rows = np.repeat(np.arange(pat.n), np.diff(pat.indptr))
    diag = pat.indices == rows
    L = sp.csr_matrix((vals, pat.indices, pat.indptr), shape=(pat.n, pat.n)).tolil()
    L.setdiag(0)
    L = L.tocsr()
    L = L - sp.diags(np.asarray(L.sum(axis=1)).ravel())
    lx = L @ v[:, 0].astype(np.float64)
    assert np.max(np.abs(lx)) > 1e-2 * np.abs(L).max() * np.abs(v[:, 0]).max()
    del diag


# --- gradient, divergence, shift ---------------------------------------------------------------


def test_gradient_is_exact_for_linear_fields_in_the_plane():
    v, f = _flat_grid()
    terms = ho.face_terms(v, f)
    a = np.array([0.3, -1.7, 0.0])
    u = v.astype(np.float64) @ a
    grad = ho.face_gradients(v, f, terms, u)
    signs = np.sign(terms.normals[:, 2:3]).astype(np.float64)
    assert np.allclose(grad * signs * signs, a[None, :], atol=1e-5)


@pytest.mark.parametrize("name", ["icosphere3", "bumpy", "torus"])
def test_divergence_sums_to_zero_on_a_closed_mesh(name):
    v, f = mesh(name)
    ops = ho.assemble(v, f, None)
    heat = ho.solve_direct(ops.matrix(ops.heat), ho.seed_indicator(v.shape[0], [0]))
    div = ho.face_unit_divergence(v, f, ops.terms, heat)
    assert abs(div.sum()) <= 1e-10 * np.abs(div).sum()


def test_shift_zeroes_a_single_source_and_orients_positive():
    rng = np.random.default_rng(0)
    phi = rng.normal(size=50)
    out = ho.shift_and_orient(phi, np.array([7]))
    assert out[7] == 0.0 and out.mean() >= 0.0
    out2 = ho.shift_and_orient(-phi, np.array([7]))
    assert np.allclose(out, out2)


def test_out_of_range_sources_raise():
    with pytest.raises(ValueError):
        ho.seed_indicator(10, [10])
    with pytest.raises(ValueError):
        ho.seed_indicator(10, [-1])


# --- whole pipeline ----------------------------------------------------------------------------


def test_reference_known_answers():
    """icosphere(3): pole to pole 3.09 and the source 0.0 (heat.py:529-534); max 3.1 from vertex 0
    (mesh.py:2249-2253)."""
    from ordito_amd import creation

    v, f = creation.icosphere(3)
    north, south = int(np.argmax(v[:, 2])), int(np.argmin(v[:, 2]))
    d = ho.geodesic(v, f, [north]).distance
    assert d[north] == 0.0
    assert f"{d[south]:.2f}" == "3.09"
    assert f"{ho.geodesic(v, f, [0]).distance.max():.1f}" == "3.1"


def test_great_circle_error_is_within_the_reference_bound():
    """On a sphere the heat method stays within 7 % of the exact great-circle distance, with
    correlation above 0.9993 (tests/test_heat.py:501-566)."""
    v, f = mesh("icosphere4")
    d = ho.geodesic(v, f, [0]).distance
    p = v.astype(np.float64)
    p /= np.linalg.norm(p, axis=1)[:, None]
    exact = np.arccos(np.clip(p @ p[0], -1.0, 1.0))
    assert np.max(np.abs(d - exact)) / np.ptp(exact) < 0.07
    assert np.corrcoef(d, exact)[0, 1] > 0.9993


def test_long_strip_far_end_within_three_percent():
    """The far end of a long strip is within 3 % of its length (tests/test_heat.py:3104-3131)."""
    from ordito_amd import creation

    v, f = creation.planar_grid(301, 4, extent=30.0)
    d = ho.geodesic(v, f, [0]).distance
    far = np.argmax(v[:, 0])
    assert abs(d[far] - 30.0) / 30.0 < 0.03


def test_pipeline_matches_dense_solves_on_a_tiny_mesh():
    from ordito_amd import creation

    v, f = creation.jittered_icosphere(1, 0.02, 3)
    ops = ho.assemble(v, f, None)
    u0 = ho.seed_indicator(v.shape[0], [0])
    heat = np.linalg.solve(ops.matrix(ops.heat).toarray(), u0)
    div = ho.face_unit_divergence(v, f, ops.terms, heat)
    P = ops.matrix(ops.poisson).toarray()
    phi = np.linalg.lstsq(P, div, rcond=None)[0]
    expected = ho.shift_and_orient(phi, np.array([0]))
    got = ho.geodesic(v, f, [0]).distance
    assert np.allclose(got, expected, atol=1e-10 * np.ptp(expected))


def test_factored_oracle_equals_the_one_shot_oracle():
    v, f = mesh("bumpy")
    oracle = ho.FactoredOracle(v, f)
    for s in (0, 17):
        assert np.allclose(oracle.distance(np.array([s])), ho.geodesic(v, f, [s]).distance, atol=1e-12)


def test_settled_cg_reaches_the_direct_solve_within_the_settle_rule():
    v, f = mesh("icosphere3")
    ops = ho.assemble(v, f, None)
    A = ops.matrix(ops.heat)
    b = ho.seed_indicator(v.shape[0], [0])
    x, rounds = ho.settled_cg(A, b)
    exact = ho.solve_direct(A, b)
    assert rounds > 0
    assert np.max(np.abs(x - exact) / np.abs(exact)) < 1e-4
    assert not ho.backward_error_exceeds(A, x, b, 1e-6)
    assert ho.backward_error_exceeds(A, x * (1 + 1e-3), b, 1e-6)


# --- the lumped masses' corner scatter ----------------------------------------------------------
#
# Sum(m) = Sum(area) and (M - tL) 1 = M 1 hold whichever vertex an area lands on, so the scatter is
# pinned by brute force instead: a Python loop over every face's three corners, each vertex's list
# of contributions summed with math.fsum (kernels/heat.py:125-131, 423-425; heat.py:2169).


def _corner_loop(n_vertices, faces, value_of):
    """Per vertex, the exactly rounded sum of value_of(face, corner) over the corners it holds."""
    lists = [[] for _ in range(n_vertices)]
    for f, corners in enumerate(np.asarray(faces).reshape(-1, 3)):
        for c, v in enumerate(corners):
            lists[int(v)].append(value_of(f, c))
    return np.array([math.fsum(values) for values in lists])


def _float64_areas(vertices, faces):
    """Face areas from an independent float64 cross product of the float32 corners."""
    p = np.asarray(vertices, dtype=np.float32).astype(np.float64)
    tri = np.asarray(faces).reshape(-1, 3)
    cross = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    return 0.5 * np.linalg.norm(cross, axis=1)


def _barycentric_mass_error(vertices, faces, mass_fn):
    """Largest relative gap between mass_fn's barycentric mass and the two brute-force sums: one
    over the oracle's float32 areas (to 1e-14), one over float64 areas (to float32 rounding)."""
    n = vertices.shape[0]
    areas32 = ho.face_terms(vertices, faces).areas
    got = mass_fn(n, faces, areas32)
    exact32 = _corner_loop(n, faces, lambda f, c: float(areas32[f])) / 3.0
    exact64 = _corner_loop(n, faces, lambda f, c, a=_float64_areas(vertices, faces): a[f]) / 3.0
    return max(
        np.max(np.abs(got - exact32) / exact32) / 1e-14,
        np.max(np.abs(got - exact64) / exact64) / 1e-5,
    )


@pytest.mark.parametrize("name", ["saddle_graded", "bumpy", "icosphere3", "grid"])
def test_barycentric_mass_equals_a_per_corner_loop(name):
    v, f = mesh(name)
    assert _barycentric_mass_error(v, f, ho.lumped_mass) <= 1.0


def test_barycentric_mass_pin_catches_misassigned_areas():
    """Two slips that keep Sum(m) = Sum(area): areas laid out face-major instead of corner-major
    (np.tile for np.repeat), and every face given its neighbour's area. Each fails the pin."""

    def tiled(n, faces, areas):
        tri = np.asarray(faces, dtype=np.int64).reshape(-1)
        return np.bincount(tri, weights=np.tile(areas.astype(np.float64), 3), minlength=n) / 3.0

    def rolled(n, faces, areas):
        return ho.lumped_mass(n, faces, np.roll(areas, 1))

    for name in ("saddle_graded", "bumpy", "icosphere3"):
        v, f = mesh(name)
        assert _barycentric_mass_error(v, f, tiled) > 1.0, name
        assert _barycentric_mass_error(v, f, rolled) > 1.0, name


def _voronoi_mass_error(vertices, faces, mass_fn):
    n = vertices.shape[0]
    shares = ho.voronoi_corner_areas(vertices, faces)
    got = mass_fn(n, faces, shares)
    exact = _corner_loop(n, faces, lambda f, c: float(shares[f, c]))
    return np.max(np.abs(got - exact) / exact) / 1e-14


@pytest.mark.parametrize("name", ["saddle_graded", "bumpy", "icosphere3", "grid"])
def test_voronoi_mass_equals_a_per_corner_loop(name):
    v, f = mesh(name)
    assert _voronoi_mass_error(v, f, ho.voronoi_mass) <= 1.0


def test_voronoi_mass_pin_catches_misassigned_shares():
    """Shares read face-major ((F, 3) transposed) or with the corners rotated keep Sum(m); each
    fails the pin on an irregular mesh."""

    def transposed(n, faces, shares):
        return ho.voronoi_mass(n, faces, np.ascontiguousarray(shares.T).reshape(-1, 3))

    def rotated(n, faces, shares):
        return ho.voronoi_mass(n, faces, np.roll(shares, 1, axis=1))

    for name in ("saddle_graded", "bumpy"):
        v, f = mesh(name)
        assert _voronoi_mass_error(v, f, transposed) > 1.0, name
        assert _voronoi_mass_error(v, f, rotated) > 1.0, name


# --- the diffusion-time floor --------------------------------------------------------------------


def _floor_closed_form(extent):
    """(|bounding-box diagonal| / 400)^2 for a box of the given edge lengths (heat.py:5228-5231)."""
    return (math.sqrt(sum(e * e for e in extent)) / 400.0) ** 2


def test_diffusion_floor_matches_closed_form_bounding_boxes():
    """An unjittered 1024 x 512 torus (R = 1, r = 0.3) has grid angles at every quarter turn, so its
    box is exactly 2.6 x 2.6 x 0.6; a 40 x 30 grid of step 0.25 spans 9.75 x 7.25 x 0; a box's
    eight corners span it. A floor from the wrong diagonal (one axis dropped, a half box, the
    decay applied once) misses each by far more than float32 rounding."""
    from ordito_amd import creation

    v, f = creation.torus(1024, 512, 1.0, 0.3)
    assert math.isclose(ho.diffusion_floor(v), _floor_closed_form((2.6, 2.6, 0.6)), rel_tol=1e-6)
    xs, ys = np.meshgrid(np.arange(40) * 0.25, np.arange(30) * 0.25, indexing="ij")
    grid = np.column_stack([xs.ravel(), ys.ravel(), np.zeros(xs.size)]).astype(np.float32)
    assert math.isclose(ho.diffusion_floor(grid), _floor_closed_form((9.75, 7.25, 0.0)), rel_tol=1e-7)
    corners = np.array([[x, y, z] for x in (-1.5, 2.0) for y in (0.5, 3.0) for z in (-4.0, 4.0)],
                       dtype=np.float32)
    assert math.isclose(ho.diffusion_floor(corners), _floor_closed_form((3.5, 2.5, 8.0)), rel_tol=1e-7)
    # The floor decides the time on the torus (h^2 ~ 3.2e-5 < 8.7e-5) and the window follows it.
    t, window = ho.resolve_time(v, f, None)
    assert t == ho.diffusion_floor(v)
    assert math.isclose(t, 8.6755e-5, rel_tol=1e-3)


def test_time_rule_raises_the_floor_on_fine_meshes():
    """(diag/400)^2 replaces h^2 when longer, and the settle window grows as sqrt(t/h^2)
    (heat.py:3714-3736, 5234-5243)."""
    from ordito_amd import creation

    v, f = creation.torus(1024, 512, 1.0, 0.3)  # about 700 mean edges across its bounding box
    t, window = ho.resolve_time(v, f, None)
    h = ho.mean_face_edge_length(v, f)
    assert t is not None and t > h * h
    assert window == math.ceil(192 * math.sqrt(t / (h * h)))
    v, f = creation.icosphere(3)
    assert ho.resolve_time(v, f, None) == (None, 192)


# --- method="igl": Voronoi mass, rim-pinned heat -----------------------------------------------


def _meyer_shares(p: np.ndarray) -> np.ndarray:
    """Meyer et al.'s Voronoi area of each corner of an acute triangle, by their cotangent formula
    ``(|e_ij|^2 cot C_k + |e_ik|^2 cot C_j) / 8`` (a different route from the oracle's cosines
    and circumcentric partials)."""

    def cot(at, a, b):
        u, v = a - at, b - at
        return np.dot(u, v) / np.linalg.norm(np.cross(u, v))

    out = np.zeros(3)
    for i in range(3):
        j, k = (i + 1) % 3, (i + 2) % 3
        e_ij, e_ik = p[j] - p[i], p[k] - p[i]
        out[i] = (np.dot(e_ij, e_ij) * cot(p[k], p[i], p[j]) + np.dot(e_ik, e_ik) * cot(p[j], p[i], p[k])) / 8
    return out


def test_voronoi_shares_follow_meyers_cotangent_formula_on_acute_faces():
    rng = np.random.default_rng(11)
    checked = 0
    for _ in range(200):
        p = rng.normal(size=(3, 3)).astype(np.float32)
        q = p.astype(np.float64)
        cosines = []
        for i in range(3):
            u, v = q[(i + 1) % 3] - q[i], q[(i + 2) % 3] - q[i]
            cosines.append(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))
        if min(cosines) < 0.05:
            continue  # the formula holds on acute faces only
        shares = ho.voronoi_corner_areas(p, np.array([0, 1, 2]))[0]
        assert np.allclose(shares, _meyer_shares(q), rtol=1e-6)
        checked += 1
    assert checked > 20


def test_voronoi_right_and_obtuse_corners_take_half_and_degenerate_faces_split_evenly():
    right = ho.voronoi_corner_areas(np.array([[0, 0, 0], [2, 0, 0], [0, 1, 0]], np.float32), np.arange(3))[0]
    assert np.allclose(right, [0.5, 0.25, 0.25], rtol=1e-7)  # circumcentre on the hypotenuse
    obtuse = ho.voronoi_corner_areas(np.array([[3, 0, 0], [0, 0, 0], [1, 0.2, 0]], np.float32), np.arange(3))[0]
    area = 0.5 * 3 * 0.2
    assert np.allclose(obtuse, [area / 4, area / 4, area / 2], rtol=1e-6)  # obtuse at corner 2
    flat = ho.voronoi_corner_areas(np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0]], np.float32), np.arange(3))[0]
    assert not flat.any()
    equilateral = ho.voronoi_corner_areas(
        np.array([[0, 0, 0], [1, 0, 0], [0.5, math.sqrt(3) / 2, 0]], np.float32), np.arange(3)
    )[0]
    assert np.allclose(equilateral, math.sqrt(3) / 12, rtol=1e-6)


@pytest.mark.parametrize("name", ["bumpy", "grid", "saddle_graded"])
def test_voronoi_mass_sums_to_the_area_and_differs_from_the_barycentric(name):
    v, f = mesh(name)
    igl = ho.assemble(v, f, None, method="igl")
    meshlab = ho.assemble(v, f, None, method="meshlab")
    p = v.astype(np.float64)
    tri = f.reshape(-1, 3)
    area = 0.5 * np.linalg.norm(np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1).sum()
    assert np.isclose(igl.mass.sum(), area, rtol=1e-6)
    assert np.abs(igl.mass - meshlab.mass).max() > 1e-3 * meshlab.mass.max()
    # Same Laplacian; the time is the unique-edge mean squared.
    assert np.array_equal(igl.laplacian, meshlab.laplacian)
    edges = np.unique(np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1), axis=0)
    h = np.linalg.norm(p[edges[:, 1]] - p[edges[:, 0]], axis=1).mean()
    assert np.isclose(igl.t, h * h, rtol=1e-6)


@pytest.mark.parametrize("name", ["grid", "saddle_graded", "icosphere3"])
def test_rim_is_the_vertices_of_edges_on_one_face(name):
    v, f = mesh(name)
    tri = f.reshape(-1, 3)
    edges = np.sort(np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]), axis=1)
    unique, counts = np.unique(edges, axis=0, return_counts=True)
    expected = np.zeros(v.shape[0], dtype=bool)
    expected[unique[counts == 1].reshape(-1)] = True
    assert np.array_equal(ho.rim_vertices(ho.operator_pattern(v.shape[0], f)), expected)


@pytest.mark.parametrize("name", ["grid", "saddle_graded"])
def test_pinned_solve_is_the_interior_system_padded_with_zeros(name):
    """The reference's own check (tests/test_heat.py:295-329): zero on the rim, the interior
    system's direct solve elsewhere."""
    v, f = mesh(name)
    n = v.shape[0]
    ops = ho.assemble(v, f, None, method="igl")
    rhs = np.zeros(n)
    rhs[n // 2 + 17] = 1.0
    got = ho.diffuse_pinned(ops, rhs)
    rim = ho.rim_vertices(ops.pattern)
    interior = np.flatnonzero(~rim)
    A = ops.matrix(ops.heat)
    expected = np.zeros(n)
    expected[interior] = np.linalg.solve(A[interior][:, interior].toarray(), rhs[interior])
    assert np.array_equal(got[rim], np.zeros(rim.sum()))
    assert np.allclose(got, expected, rtol=1e-10, atol=1e-14 * expected.max())


def test_a_closed_mesh_has_no_rim_and_no_pinned_solve():
    v, f = mesh("icosphere3")
    ops = ho.assemble(v, f, None, method="igl")
    assert ho.diffuse_pinned(ops, np.ones(v.shape[0])) is None


def test_igl_from_a_rim_source_is_the_voronoi_neumann_distance():
    """A source on the rim zeroes the pinned right-hand side, so the heat is the Neumann heat
    halved; the unit gradient does not see the factor 2, so the distance is the same bits."""
    v, f = mesh("grid")
    ops = ho.assemble(v, f, None, method="igl")
    rim = ho.rim_vertices(ops.pattern)
    source = int(np.flatnonzero(rim)[0])
    got = ho.geodesic(v, f, [source], method="igl").distance
    u0 = ho.seed_indicator(v.shape[0], [source])
    heat = ho.solve_direct(ops.matrix(ops.heat), u0)
    phi = ho.solve_poisson(ops, ho.face_unit_divergence(v, f, ops.terms, heat))
    assert np.array_equal(got, ho.shift_and_orient(phi, np.array([source])))


def test_igl_choices_matter_and_stay_within_the_great_circle_bound():
    v, f = mesh("icosphere4")
    d = ho.geodesic(v, f, [0], method="igl").distance
    p = v.astype(np.float64)
    p /= np.linalg.norm(p, axis=1)[:, None]
    exact = np.arccos(np.clip(p @ p[0], -1.0, 1.0))
    assert np.max(np.abs(d - exact)) / np.ptp(exact) < 0.07
    g, _ = mesh("grid")
    _, gf = mesh("grid")
    middle = g.shape[0] // 2 + 17
    igl = ho.geodesic(g, gf, [middle], method="igl").distance
    meshlab = ho.geodesic(g, gf, [middle]).distance
    assert np.abs(igl - meshlab).max() > 1e-3 * np.ptp(meshlab)
    factored = ho.FactoredOracle(g, gf, method="igl").distance(np.array([middle]))
    assert np.allclose(factored, igl, atol=1e-12 * np.ptp(igl))



# --- method="potpourri3d": geometry-central's regularized Poisson system ------------------------


@pytest.mark.parametrize("name", ["icosphere3", "grid", "two_bodies"])
def test_regularized_poisson_returns_ones_from_its_shift(name):
    """``(-L + 1e-6 I) 1 = 1e-6 1`` since ``L 1 = 0``: the right-hand side ``1e-6`` everywhere
    solves to ones, which fixes the regularization's sign and size (a ``+L``, a missing shift or a
    shift of the wrong size gives something else), on one and on two components."""
    v, f = mesh(name)
    ops = ho.assemble(v, f, None, "potpourri3d")
    phi = ho.solve_regularized_poisson(ops, np.full(len(v), ho.POTPOURRI_POISSON_SHIFT))
    assert np.max(np.abs(phi - 1.0)) < 1e-6
    wrong = ops.poisson.copy()
    rows = np.repeat(np.arange(ops.pattern.n), np.diff(ops.pattern.indptr))
    diagonal = ops.pattern.indices == rows
    wrong[diagonal] += 1e-5  # ten times the shift
    bad = sp.linalg.spsolve(ops.matrix(wrong).tocsc(), np.full(len(v), ho.POTPOURRI_POISSON_SHIFT))
    assert np.max(np.abs(bad - 1.0)) > 0.5


def test_potpourri3d_stays_within_the_great_circle_bound_and_differs_from_meshlab():
    v, f = mesh("icosphere4")
    d = ho.geodesic(v, f, [0], method="potpourri3d").distance
    p = v.astype(np.float64)
    p /= np.linalg.norm(p, axis=1)[:, None]
    exact = np.arccos(np.clip(p @ p[0], -1.0, 1.0))
    assert np.max(np.abs(d - exact)) / np.ptp(exact) < 0.07
    assert d[0] == 0.0 and d.min() >= 0.0
    meshlab = ho.geodesic(v, f, [0]).distance
    # The regularization and the unique-edge time move the answer, by less than the method's error.
    gap = np.abs(d - meshlab).max() / np.ptp(meshlab)
    assert 1e-9 < gap < 0.03
    factored = ho.FactoredOracle(v, f, method="potpourri3d").distance(np.array([0]))
    assert np.allclose(factored, d, atol=1e-12 * np.ptp(d))


# ---- use_robust: mollified side lengths ----------------------------------------------------------


def _box() -> tuple[np.ndarray, np.ndarray]:
    """The unit cube's 8 corners and 12 outward-wound triangles."""
    v = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], np.float32)
    quads = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tri = []
    for a, b, c, d in quads:
        tri += [(a, b, c), (a, c, d)]
    return v, np.array(tri, np.int32).reshape(-1)


def test_intrinsic_half_cotangents_reference_known_answers():
    """laplacian.py:319-322: the equilateral triangle weighs each side cot(60)/2, the 3-4-5
    triangle its hypotenuse nothing, then 0.375 and 2/3."""
    got = ho.intrinsic_half_cotangents(np.array([[1, 1, 1], [5, 4, 3]], np.float32))
    assert np.allclose(got[0], 1.0 / (2.0 * math.sqrt(3.0)), rtol=1e-6)
    assert got[1, 0] == 0.0
    assert np.allclose(got[1, 1:], [0.375, 2.0 / 3.0], rtol=1e-6)


def test_mollify_reference_known_answers():
    """laplacian.py:830-837: a cube needs no constant; the collinear triangle of sides 1, 2, 1
    needs 1e-5 of its mean side 4/3, added to every side."""
    v, f = _box()
    lengths, delta = ho.mollify_intrinsic(v, f)
    assert delta == 0.0 and np.array_equal(lengths, ho.face_edge_lengths(v, f))
    v = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0]], np.float32)
    lengths, delta = ho.mollify_intrinsic(v, np.array([0, 1, 2]))
    assert math.isclose(delta, 1e-5 * 4.0 / 3.0, rel_tol=1e-6)
    assert np.array_equal(lengths, np.array([[1, 2, 1]], np.float32) + np.float32(delta))


def test_mollified_margins_reach_the_tolerance_and_clean_meshes_keep_their_weights():
    """After mollification every triangle-inequality margin is at least epsilon times the mean
    side (the constant's definition, up to float32 rounding of the sum); on a clean mesh the
    law-of-cosines weights equal the position-based ones within float32 rounding -- two formulas
    for the same numbers, so a wrong side, sign or Heron ordering fails one of them."""
    from tests.meshes import PINCHED_FACE, pinched_grid

    v, f = pinched_grid()
    face = PINCHED_FACE
    lengths, delta = ho.mollify_intrinsic(v, f)
    base = ho.face_edge_lengths(v, f)
    assert delta > 0.0
    eps = 1e-5 * float(base.astype(np.float64).mean())
    a, b, c = (lengths[:, k].astype(np.float64) for k in range(3))
    assert min((a + b - c).min(), (b + c - a).min(), (c + a - b).min()) >= eps * (1 - 1e-3)
    # The needle's weights are finite and large after mollification, zero-area guarded before.
    cot = ho.intrinsic_half_cotangents(lengths)
    assert np.isfinite(cot).all() and np.abs(cot[face]).max() > 100.0
    for name in ("icosphere4", "bumpy", "saddle_graded"):
        vv, ff = mesh(name)
        lengths, delta = ho.mollify_intrinsic(vv, ff)
        assert delta == 0.0, name
        ext = ho.face_terms(vv, ff).cot
        got = ho.intrinsic_half_cotangents(lengths)
        scale = np.maximum(np.abs(ext), 1.0)
        # Thin faces (the hull's, the saddle's needles) lose a few float32 digits to the law of
        # cosines' cancellation; a wrong side or sign is off by O(1).
        assert (np.abs(got - ext) / scale).max() < (1e-5 if name == "icosphere4" else 1e-3), name


def test_heron_takes_the_largest_side_first():
    """kernels/predicates.py:197-210: Kahan's ordering keeps a needle's float32 double area within
    a few ulp of the float64 one; the smallest side first loses digits on the graded saddle's
    needles. The pin compares against float64 Heron of the same float32 sides."""
    vv, ff = mesh("saddle_graded")
    lengths = ho.face_edge_lengths(vv, ff)
    srt = np.sort(lengths.astype(np.float64), axis=1)
    c, b, a = srt[:, 0], srt[:, 1], srt[:, 2]
    exact = 0.5 * np.sqrt(np.maximum((a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c)), 0))

    def dbl_area(order_largest_first: bool) -> np.ndarray:
        s32 = np.sort(lengths, axis=1)
        if order_largest_first:
            c32, b32, a32 = s32[:, 0], s32[:, 1], s32[:, 2]
        else:
            a32, b32, c32 = s32[:, 0], s32[:, 1], s32[:, 2]
        arg = (a32 + (b32 + c32)) * (c32 - (a32 - b32)) * (c32 + (a32 - b32)) * (a32 + (b32 - c32))
        return (np.float32(0.5) * np.sqrt(np.maximum(arg, np.float32(0)))).astype(np.float64)

    right = np.abs(dbl_area(True) - exact) / exact
    wrong = np.abs(dbl_area(False) - exact) / exact
    assert right.max() < 1e-5
    assert wrong.max() > 10 * right.max()
    # And the oracle's weights are the right ordering's.
    cot = ho.intrinsic_half_cotangents(lengths).astype(np.float64)
    l2 = lengths.astype(np.float64) ** 2
    want = (l2[:, 1] + l2[:, 2] - l2[:, 0]) / (4 * exact)
    assert (np.abs(cot[:, 0] - want) / np.maximum(np.abs(want), 1.0)).max() < 1e-4
