"""
FP64 CPU oracle for the heat-method geodesic path. TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import this module, and only as the checker: no product path calls it, and the product path fails
loudly when the HIP library is missing instead of falling back to anything here.

A restatement in numpy/SciPy of ordito's ``heat_geodesic(method="meshlab")`` and
``method="igl"`` (ordito 0.5.2), each function citing the reference lines it follows. Precision follows the reference's split: the
per-face half-cotangents, unit normals and areas are formed in float32 from float32 positions,
every assembled operator, vector and solve is float64. Direct solves use SciPy's ``splu``, which
the reference's own Cholesky tests use as their oracle (``tests/test_cholesky.py:38-98``,
``tests/test_heat.py:3063-3101``).

Float32 expressions are spelled out operation by operation (left-to-right sums, no fused
multiply-add), so the HIP face kernel compiled with contraction off forms the same bits.

Pinning (what checks each function; see ``tests/test_oracle.py``):

- ``face_terms``: closed forms (equilateral half-cotangent ``1/(2*sqrt(3))``, right angle ``0``,
  unit normals, ``sum(area)`` = the sphere's area), float64 recomputation to 1e-6.
- ``operator_pattern`` / ``assemble``: symmetric, zero row sums, linear precision ``L x = 0`` at
  interior vertices of a planar mesh, ``sum(m) = sum(area)``, ``(M - tL) 1 = M 1``, ``t`` equal to
  the unique-edge mean squared on a closed mesh; brute-force dense cotangent matrix on tiny meshes.
- ``face_unit_divergence``: exact for linear fields (gradient), sums to zero on closed meshes,
  brute-force per-face loop on tiny meshes.
- ``geodesic``: the reference's docstring known answers (icosphere(3) pole to pole 3.09, maximum
  3.1, source exactly 0), great-circle error under 7 %, dense ``numpy.linalg`` solves on tiny meshes.
- ``settled_cg``: converges to the direct solve within the settle rule's accuracy.
- ``voronoi_corner_areas`` / ``voronoi_mass``: Meyer et al.'s cotangent formula
  ``(|e_ij|^2 cot C_k + |e_ik|^2 cot C_j) / 8`` on acute faces, the right and obtuse corners' half
  and quarters, even thirds for a degenerate face, ``sum(m) = sum(area)``.
- ``rim_vertices`` / ``diffuse_pinned``: the rim against the edges counted on one face; the pinned
  solve against a direct solve of the interior system padded with zeros; no rim on a closed mesh.
- ``geodesic(method="igl")``: a source on the rim leaves only the (halved) Neumann heat, so the
  distance is the Voronoi Neumann one bit for bit; the great-circle bound on a sphere; ``"igl"``
  and ``"meshlab"`` differ where the choices matter.
- ``mollify_intrinsic`` / ``intrinsic_half_cotangents`` (``use_robust``): the reference's
  docstring known answers (the equilateral triangle's ``cot(60)/2`` thrice, the 3-4-5 triangle's
  ``0, 0.375, 0.6667``; a closed box needs no constant, the collinear triangle of lengths 1, 2, 1
  needs ``1e-5`` of its mean ``4/3``), every margin of the mollified table at least the constant
  ``epsilon`` asked for, and on a clean mesh the law-of-cosines weights within float32 rounding of
  the position-based ``face_terms`` ones (a different formula for the same numbers).
- ``solve_regularized_poisson`` / ``geodesic(method="potpourri3d")``: ``(-L + 1e-6 I) 1 = 1e-6``
  (``L 1 = 0``), so the right-hand side ``1e-6`` gives ones exactly up to rounding and the
  regularization's sign and size are fixed; an unreferenced vertex stays zero; the great-circle
  bound on a sphere; the regularization moves the distance by well under the method's own error
  on a coarse mesh.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# heat.py:106-147: the settle rule of the heat solve's Jacobi-CG.
HEAT_CHECK_ROUNDS = 16
HEAT_CHANGE_TOLERANCE = 1e-6
HEAT_SETTLE_ROUNDS = 192
# heat.py:149-158: the default time is at least (bbox diagonal / 400)^2.
DIFFUSION_DECAY = 400.0
# linalg.py:171-181: backward-error bar of a settled iterate; componentwise refinement tolerance.
SETTLE_BACKWARD_ERROR = 1e-6
SETTLE_DIRECT_TOLERANCE = 1e-12
# heat.py:108: the Poisson solve's tolerance.
CG_TOLERANCE = 1e-8
# heat.py:190-193: geometry-central's identity multiple on the Poisson system ("potpourri3d").
POTPOURRI_POISSON_SHIFT = 1e-6
# constants.py:31: use_robust's triangle-inequality margin, relative to the mean edge length.
TOLERANCE_MOLLIFY = 1e-5


def _f32(x):
    return np.asarray(x, dtype=np.float32)


def _dot3(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    # Left-to-right float sum of products, each rounded (no fused multiply-add).
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]


def _cross3(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    return np.stack(
        [
            u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1],
            u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2],
            u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0],
        ],
        axis=1,
    )


def triangle_area_vector(a: np.ndarray, b: np.ndarray, c: np.ndarray) -> np.ndarray:
    """
    ``cross(b - a, c - a)`` formed at the corner opposite the longest edge.

    Follows ``kernels/predicates.py:106-130`` (``triangle_area_vector``): the cross product of the
    two edges at the largest angle, ties resolved to corner a, then b.
    """
    ab = b - a
    bc = c - b
    ca = a - c
    l_ab = _dot3(ab, ab)
    l_bc = _dot3(bc, bc)
    l_ca = _dot3(ca, ca)
    at_a = (l_bc >= l_ca) & (l_bc >= l_ab)
    at_b = ~at_a & (l_ca >= l_ab)
    out = _cross3(ca, -bc)
    out[at_b] = _cross3(bc[at_b], -ab[at_b])
    out[at_a] = _cross3(ab[at_a], -ca[at_a])
    return out


@dataclass
class FaceTerms:
    cot: np.ndarray  # (F, 3) float32 half-cotangents, column e at corner e
    normals: np.ndarray  # (F, 3) float32 unit normals
    areas: np.ndarray  # (F,) float32


def face_terms(vertices: np.ndarray, faces: np.ndarray) -> FaceTerms:
    """
    Per-face float32 half-cotangents, unit normals and areas.

    Follows ``kernels/heat.py:91-131`` (``heat_face_terms``), ``kernels/laplacian.py:93-106``
    (``half_cotangents_of``: ``dot / (2 |cross|)``, zero for a zero-area face) and
    ``kernels/triangles.py:332-358`` (``face_normals_and_area``: ``cross / |cross|``, area
    ``0.5 |cross|``).
    """
    p = _f32(vertices)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    v0, v1, v2 = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    cross = triangle_area_vector(v0, v1, v2)
    norm = np.sqrt(_dot3(cross, cross))
    positive = norm > np.float32(0.0)
    safe = np.where(positive, norm, np.float32(1.0))
    normals = np.where(positive[:, None], cross / safe[:, None], cross).astype(np.float32)
    areas = (np.float32(0.5) * norm).astype(np.float32)
    denominator = np.float32(2.0) * norm
    dpos = denominator > np.float32(0.0)
    dsafe = np.where(dpos, denominator, np.float32(1.0))
    c0 = _dot3(v1 - v0, v2 - v0) / dsafe
    c1 = _dot3(v2 - v1, v0 - v1) / dsafe
    c2 = _dot3(v0 - v2, v1 - v2) / dsafe
    cot = np.stack([c0, c1, c2], axis=1)
    cot[~dpos] = np.float32(0.0)
    return FaceTerms(cot.astype(np.float32), normals, areas)


def face_edge_lengths(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """
    ``(F, 3)`` float32 side lengths, column ``e`` opposite corner ``e``: ``|v1 - v2|``,
    ``|v2 - v0|``, ``|v0 - v1|``, each the float32 square root of a left-to-right float32 sum of
    squares (``kernels/triangles.py:204-221``, ``kernels/predicates.py:221-228``).
    """
    p = _f32(vertices)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    v0, v1, v2 = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    sides = []
    for a, b in ((v1, v2), (v2, v0), (v0, v1)):
        d = (a - b).astype(np.float32)
        sides.append(np.sqrt(_dot3(d, d)).astype(np.float32))
    return np.stack(sides, axis=1).astype(np.float32)


def mollify_intrinsic(
    vertices: np.ndarray, faces: np.ndarray, epsilon: float = TOLERANCE_MOLLIFY
) -> tuple[np.ndarray, float]:
    """
    The smallest constant added to every edge length that leaves every triangle's
    triangle-inequality margin ``a + b - c`` at least ``epsilon`` times the mean edge length, and
    the mollified ``(F, 3)`` float32 table (``laplacian.py:779-874``,
    ``kernels/laplacian.py:1585-1630``). Adding ``delta`` to all three sides grows each margin by
    exactly ``delta``, so ``delta = max(epsilon * mean - min margin, 0)``, formed in float32; the
    mean is the float64 mean of the float32 table. A clean mesh gets ``delta = 0`` and its
    lengths unchanged.
    """
    lengths = face_edge_lengths(vertices, faces)
    if lengths.shape[0] == 0:
        return lengths, 0.0
    a, b, c = lengths[:, 0], lengths[:, 1], lengths[:, 2]
    margin = np.minimum(np.minimum((a + b) - c, (b + c) - a), (c + a) - b).min()
    mean = float(lengths.astype(np.float64).sum()) / lengths.size
    eps_scaled = np.float32(epsilon * mean)
    delta = float(max(np.float32(eps_scaled - np.float32(margin)), np.float32(0.0)))
    if delta <= 0.0:
        return lengths, 0.0
    return (lengths + np.float32(delta)).astype(np.float32), delta


def intrinsic_half_cotangents(lengths: np.ndarray) -> np.ndarray:
    """
    ``(F, 3)`` float32 half-cotangents from side lengths (column ``e`` opposite corner ``e``), in
    float32: ``(l_{e+1}^2 + l_{e+2}^2 - l_e^2) / (4 * 2A)`` with twice the area from Kahan's
    ordering of Heron's formula, largest side first, ``0.5 * sqrt(max(arg, 0))``; a zero-area face
    gives zeros (``kernels/laplacian.py:59-90``, ``kernels/predicates.py:194-210``,
    ``laplacian.py:269-340``).
    """
    lengths = _f32(lengths)
    srt = np.sort(lengths, axis=1)
    c, b, a = srt[:, 0], srt[:, 1], srt[:, 2]
    arg = (a + (b + c)) * (c - (a - b)) * (c + (a - b)) * (a + (b - c))
    with np.errstate(invalid="ignore"):
        dbl = (np.float32(0.5) * np.sqrt(np.maximum(arg, np.float32(0.0)))).astype(np.float32)
    dbl = np.where(np.isnan(dbl), np.float32(0.0), dbl)
    den = np.float32(4.0) * dbl
    ok = den > np.float32(0.0)
    safe = np.where(ok, den, np.float32(1.0))
    l2 = lengths * lengths
    out = np.stack(
        [((l2[:, 1] + l2[:, 2]) - l2[:, 0]) / safe, ((l2[:, 2] + l2[:, 0]) - l2[:, 1]) / safe,
         ((l2[:, 0] + l2[:, 1]) - l2[:, 2]) / safe],
        axis=1,
    ).astype(np.float32)
    out[~ok] = np.float32(0.0)
    return out


@dataclass
class Pattern:
    """CSR sparsity of the cotangent operator: every face edge both ways plus referenced diagonals."""

    n: int
    indptr: np.ndarray  # (n+1,) int64
    indices: np.ndarray  # (nnz,) int64, sorted per row
    # For each stored entry, the contributing corner slots 3f+e (corner e opposite the edge),
    # ascending, as a CSR over entries; empty for diagonal entries.
    corner_ptr: np.ndarray
    corner_slots: np.ndarray


def operator_pattern(n_vertices: int, faces: np.ndarray) -> Pattern:
    """
    The operator pattern and each off-diagonal entry's run of corner slots.

    Follows ``laplacian.py:1071-1389`` / ``kernels/laplacian.py:707-760`` (``store_operator_keys``):
    a key ``i*n+j`` for both directions of each face edge with the opposite corner's slot as
    payload, plus a diagonal key for each referenced vertex; sorting the keys stably spells the
    CSR, and an entry's contributors are its run of equal keys in ascending corner order.
    """
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    n_faces = tri.shape[0]
    n = int(n_vertices)
    corner = np.arange(3 * n_faces, dtype=np.int64).reshape(-1, 3)
    # Corner e is opposite the edge (tri[e+1], tri[e+2]).
    rows, cols, slots = [], [], []
    for e in range(3):
        i = tri[:, (e + 1) % 3]
        j = tri[:, (e + 2) % 3]
        rows += [i, j]
        cols += [j, i]
        slots += [corner[:, e], corner[:, e]]
    referenced = np.unique(tri.reshape(-1)) if n_faces else np.zeros(0, dtype=np.int64)
    r = np.concatenate(rows + [referenced])
    c = np.concatenate(cols + [referenced])
    s = np.concatenate(slots + [np.full(referenced.size, -1, dtype=np.int64)])
    keys = r * n + c
    order = np.lexsort((s, keys))  # by key, then ascending corner slot
    keys, s = keys[order], s[order]
    unique, first = np.unique(keys, return_index=True)
    entry_rows = unique // n
    entry_cols = unique % n
    indptr = np.zeros(n + 1, dtype=np.int64)
    np.add.at(indptr, entry_rows + 1, 1)
    indptr = np.cumsum(indptr)
    # Runs of equal keys; the diagonal's run holds only the -1 marker.
    run_of = np.searchsorted(unique, keys)
    valid = s >= 0
    counts = np.bincount(run_of[valid], minlength=unique.size)
    corner_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    corner_slots = s[valid]
    return Pattern(n, indptr, entry_cols.astype(np.int64), corner_ptr, corner_slots)


@dataclass
class Operators:
    pattern: Pattern
    laplacian: np.ndarray  # (nnz,) float64
    heat: np.ndarray  # (nnz,) float64, M - t L
    poisson: np.ndarray  # (nnz,) float64, -L
    mass: np.ndarray  # (V,) float64
    t: float
    terms: FaceTerms

    def matrix(self, values: np.ndarray) -> sp.csr_matrix:
        n = self.pattern.n
        return sp.csr_matrix((values, self.pattern.indices, self.pattern.indptr), shape=(n, n))


def lumped_mass(n_vertices: int, faces: np.ndarray, areas: np.ndarray) -> np.ndarray:
    """
    Barycentric lumped mass ``m_i = (sum of float64 areas of faces at i) / 3``, summed over each
    vertex's corners in ascending corner order.

    Follows ``kernels/heat.py:125-131`` (area corner sums) and ``kernels/heat.py:423-425``
    (``m = area_sums / mass_divisor``, divisor 3 for the barycentric lumping).
    """
    tri = np.asarray(faces, dtype=np.int64).reshape(-1)
    weights = np.repeat(areas.astype(np.float64), 3)
    sums = np.bincount(tri, weights=weights, minlength=n_vertices) if tri.size else np.zeros(
        n_vertices
    )
    return sums / 3.0


def voronoi_corner_areas(vertices: np.ndarray, faces: np.ndarray) -> np.ndarray:
    """
    ``(F, 3)`` float32: each corner's share of its face under Meyer et al.'s mixed Voronoi
    lumping, as ``igl::massmatrix`` forms it at ``MASSMATRIX_TYPE_VORONOI``.

    Follows ``kernels/triangles.py:276-330`` (``face_voronoi_areas``), in float64 from the float32
    corners: edge lengths ``l_e = |p_{e+1} - p_{e+2}|``, area ``|cross(p1 - p0, p2 - p0)| / 2``,
    cosines by the law of cosines over the rounded lengths, circumcentric partials
    ``(cos_e l_e / sum) area``, each corner the mean of the two partials beside it; an obtuse
    corner (cosine < 0) takes half the area and the others a quarter; a zero-length edge or a zero
    cosine sum splits the area evenly. Rounded once to float32.
    """
    p = _f32(vertices).astype(np.float64)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    p0, p1, p2 = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    d0, d1, d2 = p1 - p2, p2 - p0, p0 - p1
    l0 = np.sqrt(_dot3(d0, d0))
    l1 = np.sqrt(_dot3(d1, d1))
    l2 = np.sqrt(_dot3(d2, d2))
    c = _cross3(p1 - p0, p2 - p0)
    area = 0.5 * np.sqrt(_dot3(c, c))
    even = area / 3.0
    q = np.stack([even, even, even], axis=1)
    sides = (l0 > 0.0) & (l1 > 0.0) & (l2 > 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        c0 = (l2 * l2 + l1 * l1 - l0 * l0) / (2.0 * l1 * l2)
        c1 = (l0 * l0 + l2 * l2 - l1 * l1) / (2.0 * l2 * l0)
        c2 = (l1 * l1 + l0 * l0 - l2 * l2) / (2.0 * l0 * l1)
        b0, b1, b2 = c0 * l0, c1 * l1, c2 * l2
        weight = b0 + b1 + b2
        circum = sides & (weight != 0.0)
        part0 = b0 / weight * area
        part1 = b1 / weight * area
        part2 = b2 / weight * area
    q[circum, 0] = (0.5 * (part1 + part2))[circum]
    q[circum, 1] = (0.5 * (part2 + part0))[circum]
    q[circum, 2] = (0.5 * (part0 + part1))[circum]
    half, quarter = 0.5 * area, 0.25 * area
    obtuse0 = sides & (c0 < 0.0)
    obtuse1 = sides & ~obtuse0 & (c1 < 0.0)
    obtuse2 = sides & ~obtuse0 & ~obtuse1 & (c2 < 0.0)
    for corner, mask in ((0, obtuse0), (1, obtuse1), (2, obtuse2)):
        q[mask] = quarter[mask, None]
        q[mask, corner] = half[mask]
    return q.astype(np.float32)


def voronoi_mass(n_vertices: int, faces: np.ndarray, shares: np.ndarray) -> np.ndarray:
    """
    Voronoi lumped mass ``m_i = sum of the float64-widened corner shares at i`` (no division).

    Follows ``kernels/heat.py:126-128`` (the shares' corner scatter) and ``heat.py:2169`` (divisor
    1 for the Voronoi lumping).
    """
    tri = np.asarray(faces, dtype=np.int64).reshape(-1)
    if not tri.size:
        return np.zeros(n_vertices)
    return np.bincount(tri, weights=shares.reshape(-1).astype(np.float64), minlength=n_vertices)


def unique_edge_sums(vertices: np.ndarray, pattern: Pattern) -> tuple[float, float]:
    """
    ``(sum of |v_j - v_i|, count)`` over the strict upper triangle: every unique edge once.

    Follows ``kernels/heat.py:143-235`` (``upper_edge_sums`` without ``face_edges``): float32
    lengths widened to float64.
    """
    p = _f32(vertices)
    rows = np.repeat(np.arange(pattern.n), np.diff(pattern.indptr))
    cols = pattern.indices
    upper = cols > rows
    d = p[cols[upper]] - p[rows[upper]]
    lengths = np.sqrt(_dot3(d, d)).astype(np.float64)
    return float(np.sum(lengths)), float(lengths.size)


def face_edge_sums(vertices: np.ndarray, pattern: Pattern) -> tuple[float, float]:
    """
    ``(sum of w * |v_j - v_i|, sum of w)`` over the strict upper triangle, ``w`` the faces holding
    the edge (MeshLab's mean over every face's three edges).

    Follows ``kernels/heat.py:143-235`` (``upper_edge_sums`` with ``face_edges``): float32 lengths
    widened to float64 before the weighted sum.
    """
    p = _f32(vertices)
    rows = np.repeat(np.arange(pattern.n), np.diff(pattern.indptr))
    cols = pattern.indices
    upper = cols > rows
    d = p[cols[upper]] - p[rows[upper]]
    lengths = np.sqrt(_dot3(d, d)).astype(np.float64)
    weights = np.diff(pattern.corner_ptr)[upper].astype(np.float64)
    return float(np.sum(weights * lengths)), float(np.sum(weights))


def mean_face_edge_length(vertices: np.ndarray, faces: np.ndarray) -> float:
    """Mean of the 3F per-face float32 edge lengths (``edges.py:899``, ``mesh.py:1179-1191``)."""
    p = _f32(vertices)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    if tri.shape[0] == 0:
        return 0.0
    lengths = []
    for a, b in ((0, 1), (1, 2), (2, 0)):
        d = p[tri[:, b]] - p[tri[:, a]]
        lengths.append(np.sqrt(_dot3(d, d)).astype(np.float64))
    return float(np.mean(np.concatenate(lengths)))


def diffusion_floor(vertices: np.ndarray) -> float:
    """``(bbox diagonal / 400)^2`` (``heat.py:5228-5231``)."""
    p = _f32(vertices).astype(np.float64)
    if p.shape[0] == 0:
        return 0.0
    low, high = p.min(axis=0), p.max(axis=0)
    return (math.dist(low, high) / DIFFUSION_DECAY) ** 2


def mean_unique_edge_length(vertices: np.ndarray, faces: np.ndarray) -> float:
    """Mean float32 length of the unique edges (``heat.py:2143-2146``), each edge once."""
    pattern = operator_pattern(_f32(vertices).shape[0], faces)
    total, count = unique_edge_sums(vertices, pattern)
    return total / count if count > 0.0 else 0.0


# heat.py:2001-2012: each method's (mass lumping, edges of the mean length h).
METHODS = {
    "meshlab": ("barycentric", "face_edges"),
    "igl": ("voronoi", "unique_edges"),
    "potpourri3d": ("barycentric", "unique_edges"),
}


def resolve_time(
    vertices: np.ndarray, faces: np.ndarray, t: float | None, method: str = "meshlab"
) -> tuple[float | None, int]:
    """
    The time the systems are built at (``None``: the device-formed ``h^2``) and the settle window.

    Follows ``HeatSolver._vertex_time`` (``heat.py:3714-3736``) and ``_settle_window``
    (``heat.py:5234-5243``), ``h`` over ``method``'s edges.
    """
    if METHODS[method][1] == "face_edges":
        mean = mean_face_edge_length(vertices, faces)
    else:
        mean = mean_unique_edge_length(vertices, faces)
    default = mean * mean
    if t is None:
        floor = diffusion_floor(vertices)
        t = floor if floor > default else None
    if t is None or not t > default > 0.0:
        window = HEAT_SETTLE_ROUNDS
    else:
        window = math.ceil(HEAT_SETTLE_ROUNDS * math.sqrt(t / default))
    return t, window


def assemble(
    vertices: np.ndarray,
    faces: np.ndarray,
    t: float | None = None,
    method: str = "meshlab",
    use_robust: bool = False,
) -> Operators:
    """
    ``L`` (igl sign: negative diagonal), ``M - tL`` and ``-L`` over one pattern, and ``m``
    (barycentric, or Voronoi at ``method="igl"``).

    Follows ``heat.py:2015-2206`` (``_heat_operators``) and ``kernels/heat.py:298-465``
    (``heat_system_rows``): ``L_ij = sum of the run's float64-widened half-cotangents`` in
    ascending corner order (``kernels/laplacian.py:970-989``), ``L_ii = -sum_j L_ij`` over the
    row's off-diagonal entries in column order, heat ``factor * L_ij`` (+ ``m_i`` on the diagonal,
    ``kernels/linalg.py:1612-1621``) with ``factor = -t`` and, for ``t=None``, ``-h^2`` from the
    face-edge sums (``kernels/linalg.py:1593-1609``); Poisson ``-L_ij``.
    ``t`` here is the time as resolved (explicit, or the floor), ``None`` for ``h^2``.
    ``use_robust``: the half-cotangents (the Laplacian's and the divergence's) from the mollified
    side lengths, everything else from the positions (``heat.py:2041-2052``).
    """
    vertices = _f32(vertices)
    n = vertices.shape[0]
    terms = face_terms(vertices, faces)
    if use_robust:
        lengths, _ = mollify_intrinsic(vertices, faces)
        terms = FaceTerms(intrinsic_half_cotangents(lengths), terms.normals, terms.areas)
    pattern = operator_pattern(n, faces)
    lumping, edges = METHODS[method]
    if lumping == "voronoi":
        mass = voronoi_mass(n, faces, voronoi_corner_areas(vertices, faces))
    else:
        mass = lumped_mass(n, faces, terms.areas)
    cot_flat = terms.cot.reshape(-1).astype(np.float64)
    nnz = pattern.indices.size
    # Off-diagonal values: sequential sum over each entry's corner run.
    counts = np.diff(pattern.corner_ptr)
    values = np.zeros(nnz)
    # Runs hold at most a handful of corners: add the k-th member of each run in turn, which is
    # the sequential order.
    for k in range(int(counts.max()) if counts.size else 0):
        has = counts > k
        values[has] = values[has] + cot_flat[pattern.corner_slots[pattern.corner_ptr[:-1][has] + k]]
    rows = np.repeat(np.arange(n), np.diff(pattern.indptr))
    diagonal = pattern.indices == rows
    if t is None:
        if edges == "face_edges":
            total, count = face_edge_sums(vertices, pattern)
        else:
            total, count = unique_edge_sums(vertices, pattern)
        h = total / count if count > 0.0 else 0.0
        factor = -1.0 * (h * h)
        t_value = h * h
    else:
        factor = -float(t)
        t_value = float(t)
    # Row totals of the off-diagonal entries in column order.
    lap = values.copy()
    totals = np.zeros(n)
    offd = np.where(~diagonal, values, 0.0)
    row_len = np.diff(pattern.indptr)
    for k in range(int(row_len.max()) if row_len.size else 0):
        has = row_len > k
        totals[has] = totals[has] + offd[pattern.indptr[:-1][has] + k]
    lap[diagonal] = -totals[rows[diagonal]]
    heat = factor * lap
    heat[diagonal] = heat[diagonal] + mass[rows[diagonal]]
    poisson = -lap
    return Operators(pattern, lap, heat, poisson, mass, t_value, terms)


def seed_indicator(n_vertices: int, sources: np.ndarray) -> np.ndarray:
    """``u0[s] = 1`` with duplicates allowed; ValueError on an out-of-range index.

    Follows ``kernels/heat.py:599-623`` and ``heat.py:5488-5503``.
    """
    sources = np.asarray(sources, dtype=np.int64).reshape(-1)
    if np.any(sources < 0) or np.any(sources >= n_vertices):
        raise ValueError(f"every entry of sources must be a vertex index in [0, {n_vertices})")
    u0 = np.zeros(n_vertices)
    u0[sources] = 1.0
    return u0


def stable_normalize(v: np.ndarray) -> np.ndarray:
    """Scale rows by their largest component, then normalize; zero rows stay zero.

    Follows ``kernels/predicates.py:718-724``.
    """
    largest = np.max(np.abs(v), axis=1)
    out = np.zeros_like(v)
    nz = largest != 0.0
    w = v[nz] / largest[nz][:, None]
    length = np.sqrt(w[:, 0] * w[:, 0] + w[:, 1] * w[:, 1] + w[:, 2] * w[:, 2])
    out[nz] = w / length[:, None]
    return out


def face_gradients(
    vertices: np.ndarray, faces: np.ndarray, terms: FaceTerms, u: np.ndarray
) -> np.ndarray:
    """
    Per-face gradient of a vertex field: ``sum_k u_k (N x e_k) / (2A)`` in float64 from the float32
    unit normal and area, ``e_k`` the edge opposite corner k; zero for a zero-area face.

    Follows ``kernels/triangles.py:834-864`` (``face_gradient_given``).
    """
    p = _f32(vertices).astype(np.float64)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    v0, v1, v2 = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    n = terms.normals.astype(np.float64)
    area = terms.areas.astype(np.float64)
    u0, u1, u2 = u[tri[:, 0]], u[tri[:, 1]], u[tri[:, 2]]
    grad = (
        u0[:, None] * _cross3(n, v2 - v1)
        + u1[:, None] * _cross3(n, v0 - v2)
        + u2[:, None] * _cross3(n, v1 - v0)
    ) / (2.0 * area[:, None] + (area[:, None] <= 0.0))
    grad[area <= 0.0] = 0.0
    return grad


def face_unit_divergence(
    vertices: np.ndarray, faces: np.ndarray, terms: FaceTerms, heat: np.ndarray
) -> np.ndarray:
    """
    ``-b``: the integrated divergence of the unit gradient direction ``x = grad u / |grad u|``.

    Follows ``kernels/heat.py:626-687`` (``accumulate_face_divergence``,
    ``unit_gradient_divergence``), ``kernels/triangles.py:834-894`` (``face_gradient_given``:
    ``sum_k u_k (N x e_k) / (2A)`` in float64 with float32 N and A), and
    ``heat.py:3112-3122`` / ``3218-3232`` (per-corner terms summed per vertex in ascending corner
    order). Linear in ``x``, so integrating ``+grad`` accumulates ``-b`` directly.
    """
    p = _f32(vertices).astype(np.float64)
    tri = np.asarray(faces, dtype=np.int64).reshape(-1, 3)
    v0, v1, v2 = p[tri[:, 0]], p[tri[:, 1]], p[tri[:, 2]]
    x = stable_normalize(face_gradients(vertices, faces, terms, heat))
    c = terms.cot.astype(np.float64)
    c0, c1, c2 = c[:, 0], c[:, 1], c[:, 2]
    d0 = c2 * _dot3(v1 - v0, x) + c1 * _dot3(v2 - v0, x)
    d1 = c0 * _dot3(v2 - v1, x) + c2 * _dot3(v0 - v1, x)
    d2 = c1 * _dot3(v0 - v2, x) + c0 * _dot3(v1 - v2, x)
    corners = np.stack([d0, d1, d2], axis=1).reshape(-1)
    return np.bincount(tri.reshape(-1), weights=corners, minlength=p.shape[0])


def shift_and_orient(phi: np.ndarray, sources: np.ndarray) -> np.ndarray:
    """
    ``phi - mean(phi[sources])``, negated if the shifted field's mean is negative.

    Follows ``kernels/heat.py:3383-3424`` (``source_and_global_sums``, ``shift_and_orient``).
    """
    sources = np.asarray(sources, dtype=np.int64).reshape(-1)
    offset = float(np.sum(phi[sources])) / sources.size
    mean_all = float(np.sum(phi)) / phi.size
    if mean_all - offset < 0.0:
        return offset - phi
    return phi - offset


def components(pattern: Pattern) -> tuple[int, np.ndarray]:
    from scipy.sparse.csgraph import connected_components

    n = pattern.n
    graph = sp.csr_matrix(
        (np.ones(pattern.indices.size), pattern.indices, pattern.indptr), shape=(n, n)
    )
    return connected_components(graph, directed=False)


def factorize(matrix: sp.csr_matrix, ordering: str = "MMD_AT_PLUS_A"):
    """FP64 sparse LU (``scipy.sparse.linalg.splu``) under a fill-reducing column ordering:
    SuperLU's symmetric minimum degree by default; ``"COLAMD"`` where minimum degree is slow (it
    takes minutes on subdivided icospheres of 10^5 vertices, where COLAMD takes seconds). The
    ordering changes only the rounding of the direct solve."""
    return spla.splu(matrix.tocsc(), permc_spec=ordering)


def solve_direct(matrix: sp.csr_matrix, rhs: np.ndarray) -> np.ndarray:
    """FP64 sparse LU solve, for ``rhs`` of shape (n,) or (n, k)."""
    return factorize(matrix).solve(rhs)


class FactoredOracle:
    """
    Both systems of one mesh factored once, for many source sets: the reference's amortized
    (``Trimesh``, factored) form (``heat.py:4394-4423``), solved directly in FP64.
    """

    def __init__(
        self,
        vertices: np.ndarray,
        faces: np.ndarray,
        t: float | None = None,
        ordering: str = "MMD_AT_PLUS_A",
        method: str = "meshlab",
        use_robust: bool = False,
    ) -> None:
        self.vertices = _f32(vertices)
        self.faces = np.asarray(faces, dtype=np.int64).reshape(-1)
        t_resolved, _ = resolve_time(self.vertices, self.faces, t, method)
        self.ops = assemble(self.vertices, self.faces, t_resolved, method, use_robust)
        self.heat_lu = factorize(self.ops.matrix(self.ops.heat), ordering)
        self.rim = rim_vertices(self.ops.pattern)
        self.pinned_lu = None
        if method == "igl" and self.rim.any():
            pinned = self.ops.matrix(pinned_heat_values(self.ops, self.rim))
            self.pinned_lu = factorize(pinned, ordering)
        n = self.ops.pattern.n
        n_comp, labels = components(self.ops.pattern)
        self.members = [np.flatnonzero(labels == c) for c in range(n_comp)]
        pinned = np.zeros(n, dtype=bool)
        for members in self.members:
            pinned[members[-1]] = True
        self.keep = ~pinned
        self.regularized = method == "potpourri3d"
        if self.regularized:
            self.keep = np.ones(n, dtype=bool)
            self.poisson_lu = factorize(regularized_poisson_matrix(self.ops), ordering)
        else:
            reduced = self.ops.matrix(self.ops.poisson)[self.keep][:, self.keep]
            self.poisson_lu = factorize(reduced, ordering)

    def distance(self, sources: np.ndarray) -> np.ndarray:
        n = self.ops.pattern.n
        u0 = seed_indicator(n, sources)
        heat = self.heat_lu.solve(u0)
        if self.pinned_lu is not None:
            heat = average_heat(heat, self.pinned_lu.solve(np.where(self.rim, 0.0, u0)))
        neg_div = face_unit_divergence(self.vertices, self.faces, self.ops.terms, heat)
        phi = np.zeros(n)
        if self.regularized:
            rhs = np.where(np.diff(self.ops.pattern.indptr) == 0, 0.0, neg_div)
            phi = self.poisson_lu.solve(rhs)
            return shift_and_orient(phi, sources)
        phi[self.keep] = self.poisson_lu.solve(neg_div[self.keep])
        for members in self.members:
            phi[members] -= phi[members].mean()
        return shift_and_orient(phi, sources)


def solve_poisson(ops: Operators, rhs: np.ndarray) -> np.ndarray:
    """
    ``-L phi = rhs`` with one row pinned per singular component, then each component's constant
    set to zero mean.

    Follows ``kernels/cholesky.py:114-214`` (one pin per component whose rows sum to zero, its
    equation dropped and its unknown held at zero) and ``kernels/cholesky.py:1467-1660`` (the
    constant set to the initial guess's mean, zero here). Which row is pinned moves the answer by
    the right-hand side's rounding inconsistency only.
    """
    pattern = ops.pattern
    n = pattern.n
    rhs = np.array(rhs, dtype=np.float64)
    n_comp, labels = components(pattern)
    pins = []
    pinned = np.zeros(n, dtype=bool)
    for c in range(n_comp):
        members = np.flatnonzero(labels == c)
        pin = int(members[-1])
        pins.append((pin, members))
        pinned[pin] = True
    keep = ~pinned
    reduced = ops.matrix(ops.poisson)[keep][:, keep]
    phi = np.zeros(n)
    if keep.any():
        phi[keep] = solve_direct(reduced, rhs[keep])
    for _, members in pins:
        phi[members] -= phi[members].mean()
    return phi


def regularized_poisson_matrix(ops: Operators) -> sp.csr_matrix:
    """
    geometry-central's Poisson system ``-L + 1e-6 I`` over the Poisson pattern
    (``HeatSolver._solve_regularized_poisson``, ``heat.py:4085-4110``: ``add_to_diagonal`` on the
    stored diagonal entries); a row with no stored diagonal (an unreferenced vertex) holds the
    identity's, so its unknown stays zero, as the factorization's pin keeps it.
    """
    pattern = ops.pattern
    rows = np.repeat(np.arange(pattern.n), np.diff(pattern.indptr))
    diagonal = pattern.indices == rows
    values = ops.poisson.copy()
    values[diagonal] = values[diagonal] + POTPOURRI_POISSON_SHIFT
    matrix = ops.matrix(values).tolil()
    for v in np.flatnonzero(np.diff(pattern.indptr) == 0):
        matrix[v, v] = 1.0
    return matrix.tocsr()


def solve_regularized_poisson(ops: Operators, rhs: np.ndarray) -> np.ndarray:
    """``(-L + 1e-6 I) phi = rhs``, nonsingular: no pin, no mean removed (``heat.py:3121-3128``);
    an unreferenced vertex's right-hand side is taken as zero."""
    rhs = np.array(rhs, dtype=np.float64)
    rhs[np.diff(ops.pattern.indptr) == 0] = 0.0
    return solve_direct(regularized_poisson_matrix(ops), rhs)


def rim_vertices(pattern: Pattern) -> np.ndarray:
    """
    ``(V,)`` bool: the vertices with an edge on a single face.

    Follows ``kernels/laplacian.py:1286-1321`` (``pattern_row_on_rim``): some off-diagonal entry of
    the row has a run of one corner slot.
    """
    rows = np.repeat(np.arange(pattern.n), np.diff(pattern.indptr))
    single = (np.diff(pattern.corner_ptr) == 1) & (pattern.indices != rows)
    rim = np.zeros(pattern.n, dtype=bool)
    rim[rows[single]] = True
    return rim


def pinned_heat_values(ops: Operators, rim: np.ndarray) -> np.ndarray:
    """
    The heat system's values with the solution pinned to zero on the rim: a rim row becomes the
    identity's, an interior row drops its rim columns (``kernels/heat.py:487-508``,
    ``pinned_heat_rows``).
    """
    pattern = ops.pattern
    rows = np.repeat(np.arange(pattern.n), np.diff(pattern.indptr))
    values = ops.heat.copy()
    values[rim[pattern.indices]] = 0.0
    on_rim = rim[rows]
    values[on_rim] = np.where(pattern.indices[on_rim] == rows[on_rim], 1.0, 0.0)
    return values


def diffuse_pinned(ops: Operators, rhs: np.ndarray) -> np.ndarray | None:
    """
    ``HeatSolver.diffuse_pinned`` (``heat.py:4153-4200``) solved directly: the pinned system
    against the right-hand side zeroed on the rim (``kernels/heat.py:511-514``, ``off_rim``);
    ``None`` for a mesh without a rim, where nothing is solved.
    """
    rim = rim_vertices(ops.pattern)
    if not rim.any():
        return None
    return solve_direct(ops.matrix(pinned_heat_values(ops, rim)), np.where(rim, 0.0, rhs))


def average_heat(neumann: np.ndarray, pinned: np.ndarray) -> np.ndarray:
    """``0.5 * (neumann + pinned)``: igl's boundary rule (``kernels/heat.py:517-520``)."""
    return 0.5 * (neumann + pinned)


@dataclass
class GeodesicResult:
    distance: np.ndarray
    heat: np.ndarray
    neg_divergence: np.ndarray
    ops: Operators


def geodesic(
    vertices: np.ndarray,
    faces: np.ndarray,
    sources: np.ndarray,
    t: float | None = None,
    method: str = "meshlab",
    use_robust: bool = False,
) -> GeodesicResult:
    """
    ``heat_geodesic(vertices, faces, sources, t, method=method)`` solved directly in FP64
    (``method`` "meshlab", "igl" or "potpourri3d"; at "igl" the Voronoi mass, the unique-edge time
    and the heat averaged with the rim-pinned heat, ``heat.py:594-597``; at "potpourri3d" the
    unique-edge time and the regularized Poisson system, ``heat.py:3121-3128``).

    Follows ``heat.py:552-609`` (``_heat_geodesic``): empty input gives zeros; the time rule of
    ``heat.py:3714-3736``; heat step ``(M - tL) u = u0`` (``heat.py:3890-3914``); divergence
    (``heat.py:3052-3131``); Poisson ``-L phi = -b`` (``heat.py:4062-4151``); shift and orient
    (``heat.py:3134-3162``). Both solves are direct, the reference's factored form (whose
    refinement reaches 1e-12 componentwise / 1e-8 normwise). ``use_robust``: the half-cotangents
    from mollified side lengths (``assemble``).
    """
    vertices = _f32(vertices)
    faces = np.asarray(faces, dtype=np.int64).reshape(-1)
    n = vertices.shape[0]
    sources = np.asarray(sources, dtype=np.int64).reshape(-1)
    if n == 0 or faces.size == 0 or sources.size == 0:
        z = np.zeros(n)
        return GeodesicResult(z, z, z, None)  # type: ignore[arg-type]
    u0 = seed_indicator(n, sources)
    t_resolved, _ = resolve_time(vertices, faces, t, method)
    ops = assemble(vertices, faces, t_resolved, method, use_robust)
    heat = solve_direct(ops.matrix(ops.heat), u0)
    if method == "igl":
        pinned = diffuse_pinned(ops, u0)
        if pinned is not None:
            heat = average_heat(heat, pinned)
    neg_div = face_unit_divergence(vertices, faces, ops.terms, heat)
    if method == "potpourri3d":
        phi = solve_regularized_poisson(ops, neg_div)
    else:
        phi = solve_poisson(ops, neg_div)
    distance = shift_and_orient(phi, sources)
    return GeodesicResult(distance, heat, neg_div, ops)


def geodesic_columns(
    vertices: np.ndarray,
    faces: np.ndarray,
    source_sets: list[np.ndarray],
    t: float | None = None,
    method: str = "meshlab",
) -> np.ndarray:
    """One distance field per source set, sharing one factorization of each system: (k, V)."""
    oracle = FactoredOracle(vertices, faces, t, method=method)
    return np.stack([oracle.distance(s) for s in source_sets]) if source_sets else np.zeros(
        (0, _f32(vertices).shape[0])
    )


def relative_error(d: np.ndarray, reference: np.ndarray) -> tuple[float, float]:
    """``(max, mean)`` of ``|d - d*| / ptp(d*)``, the reference's convention
    (``tests/test_heat.py:478``, ``700``)."""
    span = float(np.ptp(reference)) if reference.size else 0.0
    if span == 0.0:
        span = 1.0
    err = np.abs(np.asarray(d, dtype=np.float64) - reference) / span
    return float(err.max()) if err.size else 0.0, float(err.mean()) if err.size else 0.0


# ---------------------------------------------------------------------------------------------
# The settled Jacobi conjugate gradient (the reference's one-shot heat solve)
# ---------------------------------------------------------------------------------------------


def settled_cg(
    matrix: sp.csr_matrix,
    rhs: np.ndarray,
    *,
    check_rounds: int = HEAT_CHECK_ROUNDS,
    change_tolerance: float = HEAT_CHANGE_TOLERANCE,
    settle_rounds: int = HEAT_SETTLE_ROUNDS,
    narrow_values: bool = True,
    maxiter: int | None = None,
) -> tuple[np.ndarray, int]:
    """
    Jacobi-preconditioned Chronopoulos-Gear CG at zero tolerance, stopped by the settle rule.

    Follows ``linalg.py:2490-2727`` (``solve_spd_settled``), the round of
    ``kernels/algorithms/conjugate_gradient.py:381-866`` (``w = A u`` with float32-narrowed
    values widened, dots ``r.u, w.u, r.r``; ``beta = gamma/gamma_old``,
    ``alpha = gamma / (delta - beta gamma / alpha_old)``; ``p = u + beta p``, ``s = w + beta s``,
    ``x += alpha p``, ``r -= alpha s``, ``u = D^-1 r``; no step once ``r.r`` or the
    denominator is zero) and the settle monitor of ``conjugate_gradient.py:1491-1580`` (every
    ``check_rounds`` rounds: stop when the count of non-zero entries is unchanged and every entry
    moved by less than ``change_tolerance`` of itself, or ``settle_rounds`` rounds after the count
    last changed). Returns ``(x, rounds that stepped)``.
    """
    n = matrix.shape[0]
    a = matrix.copy()
    if narrow_values:
        a.data = a.data.astype(np.float32).astype(np.float64)
    diag = matrix.diagonal()
    inv_diag = np.where(diag != 0.0, 1.0 / np.where(diag != 0.0, diag, 1.0), 0.0)
    x = np.zeros(n)
    r = rhs.astype(np.float64).copy()
    u = inv_diag * r
    p = np.zeros(n)
    s = np.zeros(n)
    gamma_old, alpha_old = math.inf, 1.0
    maxiter = 10 * n if maxiter is None else maxiter
    previous = np.zeros(n)
    state_previous, reach_check, checks = 0.0, 0.0, 0.0
    stepped = 0
    rounds = 0
    running = True
    while running:
        for _ in range(check_rounds):
            rounds += 1
            w = a @ u
            gamma, delta, rr = float(r @ u), float(w @ u), float(r @ r)
            step = False
            if rr > 0.0 and rounds <= maxiter and gamma != 0.0:
                beta = gamma / gamma_old
                denominator = delta - beta * gamma / alpha_old
                if denominator != 0.0:
                    alpha = gamma / denominator
                    step = True
            if step:
                p = u + beta * p
                s = w + beta * s
                x = x + alpha * p
                r = r - alpha * s
                u = inv_diag * r
                gamma_old, alpha_old = gamma, alpha
                stepped += 1
            else:
                running = False
                break
        if not running:
            break
        # The settle check (cg_settle_change + cg_settle_verdict).
        nz = x != 0.0
        change = float(np.max(np.abs(x[nz] - previous[nz]) / np.abs(x[nz]))) if nz.any() else 0.0
        count = float(nz.sum()) + 1.0
        previous = x.copy()
        checks += 1.0
        if count != state_previous:
            state_previous = count
            reach_check = checks
            continue
        idle = (checks - reach_check) * check_rounds
        if change < change_tolerance or idle >= settle_rounds:
            running = False
        if rounds >= maxiter:
            running = False
    return x, stepped


def backward_error_exceeds(matrix: sp.csr_matrix, x: np.ndarray, b: np.ndarray, tol: float) -> bool:
    """Componentwise Oettli-Prager test ``|b - Ax|_i <= tol (|A||x| + |b|)_i`` everywhere
    (``kernels/linalg.py:1128-1230``)."""
    residual = b - matrix @ x
    scale = abs(matrix) @ np.abs(x) + np.abs(b)
    return bool(np.any(~(np.abs(residual) <= tol * scale)))
