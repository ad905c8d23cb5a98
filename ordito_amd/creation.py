"""
Seeded synthetic meshes for the heat-geodesic path: icospheres, tori, planar grids and the
spherical-harmonic displaced sphere.

Every generator returns ``(vertices, faces)`` as host numpy arrays: ``(V, 3) float32`` positions
and a flat ``(3F,) int32`` face buffer, the layout ``ordito.heat.heat_geodesic`` takes
(``vertices (V,) vec3 float32``, ``faces (3F,) int32``; ordito ``heat.py:370-400``).

The icosphere follows the geometry of ordito's ``creation.icosphere`` (``creation.py:641-712``):
recursive subdivision of the regular icosahedron with every new vertex the projected midpoint of
its edge, the 12 base corners first so that vertex 0 is the ``(-0.526, 0.851, 0)`` corner
(``creation.py:104-121``). The face connectivity here is built by this module's own midpoint
subdivision, so the vertex numbering past the 12 corners is this module's, not ordito's.
"""

from __future__ import annotations

import math

import numpy as np

# The regular icosahedron on the unit sphere: the cyclic permutations of (+-1, +-phi, 0) scaled by
# 1 / sqrt(1 + phi^2), in the corner order ordito's table uses (creation.py:104-121), so vertex 0
# is the corner (-1, phi, 0) / sqrt(1 + phi^2).
_PHI = (1.0 + math.sqrt(5.0)) / 2.0
_CORNERS = np.array(
    [
        [-1.0, _PHI, 0.0],
        [1.0, _PHI, 0.0],
        [-1.0, -_PHI, 0.0],
        [1.0, -_PHI, 0.0],
        [0.0, -1.0, _PHI],
        [0.0, 1.0, _PHI],
        [0.0, -1.0, -_PHI],
        [0.0, 1.0, -_PHI],
        [_PHI, 0.0, -1.0],
        [_PHI, 0.0, 1.0],
        [-_PHI, 0.0, -1.0],
        [-_PHI, 0.0, 1.0],
    ]
) / math.sqrt(1.0 + _PHI * _PHI)


def _icosahedron_faces(corners: np.ndarray) -> np.ndarray:
    """The 20 outward-wound faces: every triple of mutually adjacent corners (edge length 1.0515)."""
    distance = np.linalg.norm(corners[:, None, :] - corners[None, :, :], axis=2)
    edge = distance[0][distance[0] > 1e-9].min()
    adjacent = np.abs(distance - edge) < 1e-6
    faces = []
    n = corners.shape[0]
    for a in range(n):
        for b in range(a + 1, n):
            if not adjacent[a, b]:
                continue
            for c in range(b + 1, n):
                if adjacent[a, c] and adjacent[b, c]:
                    normal = np.cross(corners[b] - corners[a], corners[c] - corners[a])
                    if np.dot(normal, corners[a] + corners[b] + corners[c]) < 0.0:
                        faces.append((a, c, b))
                    else:
                        faces.append((a, b, c))
    result = np.array(faces, dtype=np.int64)
    assert result.shape == (20, 3)
    return result


def _subdivide_once(vertices: np.ndarray, faces: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Split every face into four at its edge midpoints, projected back to the unit sphere."""
    n = vertices.shape[0]
    a, b, c = faces[:, 0], faces[:, 1], faces[:, 2]
    # The three sides of every face, each as an undirected key.
    sides = np.stack([np.stack([a, b], 1), np.stack([b, c], 1), np.stack([c, a], 1)], 1)
    lo = sides.min(axis=2)
    hi = sides.max(axis=2)
    keys = (lo * n + hi).reshape(-1)
    unique, inverse = np.unique(keys, return_inverse=True)
    ends_lo = unique // n
    ends_hi = unique % n
    midpoints = vertices[ends_lo] + vertices[ends_hi]
    midpoints /= np.linalg.norm(midpoints, axis=1)[:, None]
    mid = (inverse + n).reshape(-1, 3)
    m_ab, m_bc, m_ca = mid[:, 0], mid[:, 1], mid[:, 2]
    new_faces = np.concatenate(
        [
            np.stack([a, m_ab, m_ca], 1),
            np.stack([b, m_bc, m_ab], 1),
            np.stack([c, m_ca, m_bc], 1),
            np.stack([m_ab, m_bc, m_ca], 1),
        ]
    )
    return np.concatenate([vertices, midpoints]), new_faces


def icosphere(subdivisions: int, radius: float = 1.0) -> tuple[np.ndarray, np.ndarray]:
    """
    Unit icosphere of ``10 * 4**k + 2`` vertices and ``20 * 4**k`` faces, scaled to ``radius``.

    The geometry is computed in float64 and rounded to float32 once at the end.
    """
    vertices = _CORNERS.copy()
    faces = _icosahedron_faces(vertices)
    for _ in range(max(0, int(subdivisions))):
        vertices, faces = _subdivide_once(vertices, faces)
    return (radius * vertices).astype(np.float32), faces.astype(np.int32).reshape(-1)


def jittered_icosphere(
    subdivisions: int, sigma: float = 0.01, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """Icosphere whose vertices are pushed radially to ``1 + N(0, sigma)``."""
    vertices, faces = icosphere(subdivisions)
    rng = np.random.default_rng(seed)
    radius = 1.0 + rng.normal(0.0, sigma, vertices.shape[0])
    unit = vertices.astype(np.float64)
    unit /= np.linalg.norm(unit, axis=1)[:, None]
    return (unit * radius[:, None]).astype(np.float32), faces


def torus(
    major_sections: int,
    minor_sections: int,
    major_radius: float = 1.0,
    minor_radius: float = 0.3,
    jitter: float = 0.0,
    seed: int = 0,
) -> tuple[np.ndarray, np.ndarray]:
    """
    Torus around Z from a periodic ``major x minor`` quad grid, each quad split into two triangles.

    ``jitter`` moves every vertex along its surface normal by ``N(0, jitter * h)``, ``h`` the
    unjittered mesh's mean edge length over every face's three edges.
    """
    nu, nv = int(major_sections), int(minor_sections)
    u = 2.0 * np.pi * np.arange(nu) / nu
    v = 2.0 * np.pi * np.arange(nv) / nv
    uu, vv = np.meshgrid(u, v, indexing="ij")
    ring = major_radius + minor_radius * np.cos(vv)
    vertices = np.stack([ring * np.cos(uu), ring * np.sin(uu), minor_radius * np.sin(vv)], -1)
    vertices = vertices.reshape(-1, 3)
    normals = np.stack([np.cos(vv) * np.cos(uu), np.cos(vv) * np.sin(uu), np.sin(vv)], -1)
    normals = normals.reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    i1, j1 = (i + 1) % nu, (j + 1) % nv
    p00 = (i * nv + j).ravel()
    p10 = (i1 * nv + j).ravel()
    p11 = (i1 * nv + j1).ravel()
    p01 = (i * nv + j1).ravel()
    faces = np.concatenate([np.stack([p00, p10, p11], 1), np.stack([p00, p11, p01], 1)])
    if jitter > 0.0:
        h = mean_face_edge_length(vertices, faces.reshape(-1))
        rng = np.random.default_rng(seed)
        vertices = vertices + normals * rng.normal(0.0, jitter * h, vertices.shape[0])[:, None]
    return vertices.astype(np.float32), faces.astype(np.int32).reshape(-1)


def planar_grid(
    nx: int, ny: int, extent: float = 1.0, jitter: float = 0.0, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """
    Open ``nx x ny``-vertex grid in the plane ``z = 0`` spanning ``[0, extent]`` along X, each cell
    split by alternating diagonals (a regular triangulation with no right-angle-only cells), with
    in-plane jitter of ``N(0, jitter * spacing)`` on interior vertices. Has a rim.
    """
    spacing = extent / (nx - 1)
    x = np.arange(nx) * spacing
    y = np.arange(ny) * spacing
    xx, yy = np.meshgrid(x, y, indexing="ij")
    vertices = np.stack([xx, yy, np.zeros_like(xx)], -1).reshape(-1, 3)
    if jitter > 0.0:
        rng = np.random.default_rng(seed)
        interior = ((xx > 0) & (xx < x[-1]) & (yy > 0) & (yy < y[-1])).reshape(-1)
        shift = rng.normal(0.0, jitter * spacing, (vertices.shape[0], 2))
        vertices[interior, :2] += shift[interior]
    i, j = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), indexing="ij")
    p00 = (i * ny + j).ravel()
    p10 = ((i + 1) * ny + j).ravel()
    p11 = ((i + 1) * ny + j + 1).ravel()
    p01 = (i * ny + j + 1).ravel()
    flip = ((i + j) % 2 == 1).ravel()
    first = np.where(flip[:, None], np.stack([p00, p10, p01], 1), np.stack([p00, p10, p11], 1))
    second = np.where(flip[:, None], np.stack([p10, p11, p01], 1), np.stack([p00, p11, p01], 1))
    faces = np.concatenate([first, second])
    return vertices.astype(np.float32), faces.astype(np.int32).reshape(-1)


def displaced_sphere(
    subdivisions: int, lmax: int = 16, amplitude: float = 0.05, seed: int = 0
) -> tuple[np.ndarray, np.ndarray]:
    """
    Icosphere with radius ``1 + sum_{l<=lmax, m} c_lm Y_lm``, ``c_lm ~ N(0, amplitude / l)``, the
    orthonormal real spherical harmonics (Condon-Shortley phase) of ``1 <= l <= lmax``,
    ``-l <= m <= l``, the coefficients drawn in that order.

    The associated Legendre functions come from the standard three-term recurrence in the degree
    at each order (``P_m^m = (-1)^m (2m-1)!! (1 - x^2)^(m/2)``, ``P_{m+1}^m = (2m+1) x P_m^m``,
    ``(l - m) P_l^m = (2l - 1) x P_{l-1}^m - (l + m - 1) P_{l-2}^m``), one order at a time, so only
    three functions of the vertices are held at once.
    """
    vertices, faces = icosphere(subdivisions)
    unit = vertices.astype(np.float64)
    unit /= np.linalg.norm(unit, axis=1)[:, None]
    x = np.clip(unit[:, 2], -1.0, 1.0)
    phi = np.arctan2(unit[:, 1], unit[:, 0])
    rng = np.random.default_rng(seed)
    coefficients = {}
    for degree in range(1, lmax + 1):
        for order in range(-degree, degree + 1):
            coefficients[(degree, order)] = rng.normal(0.0, amplitude / degree)
    radius = np.ones(unit.shape[0])
    sine = np.sqrt(np.maximum(1.0 - x * x, 0.0))
    p_mm = np.ones_like(x)  # P_m^m, built up order by order
    for k in range(0, lmax + 1):
        if k > 0:
            p_mm = -(2 * k - 1) * sine * p_mm
        cos_k, sin_k = (np.cos(k * phi), np.sin(k * phi)) if k > 0 else (None, None)
        previous, current = None, p_mm
        for degree in range(k, lmax + 1):
            if degree > k:
                if degree == k + 1:
                    nxt = (2 * k + 1) * x * current
                else:
                    nxt = ((2 * degree - 1) * x * current - (degree + k - 1) * previous) / (degree - k)
                previous, current = current, nxt
            if degree == 0:
                continue
            norm = math.sqrt(
                (2 * degree + 1) / (4.0 * math.pi) * math.factorial(degree - k)
                / math.factorial(degree + k)
            )
            if k == 0:
                radius += coefficients[(degree, 0)] * (norm * current)
            else:
                scaled = math.sqrt(2.0) * norm * current
                radius += coefficients[(degree, k)] * (scaled * cos_k)
                radius += coefficients[(degree, -k)] * (scaled * sin_k)
    return (unit * radius[:, None]).astype(np.float32), faces


def mesh_collection(count: int, low: float, high: float, rng) -> dict:
    """
    ``count`` seeded meshes keyed 0..count-1 with vertex counts log-uniform in ``[low, high]``, each
    a jittered icosphere at the nearest subdivision level, a jittered torus or a jittered planar
    grid (open, with a rim), chosen by ``rng`` (a ``numpy.random.Generator``).
    """
    meshes = {}
    for i in range(count):
        target = float(np.exp(rng.uniform(np.log(low), np.log(high))))
        kind = int(rng.integers(3))
        seed = int(rng.integers(1 << 30))
        if kind == 0:
            level = int(np.clip(round(np.log(max(target - 2, 10) / 10) / np.log(4)), 1, 9))
            meshes[i] = jittered_icosphere(level, 0.01, seed)
        elif kind == 1:
            minor = max(8, int(round(np.sqrt(target / 2))))
            major = max(8, int(round(target / minor)))
            meshes[i] = torus(major, minor, 1.0, 0.3, jitter=0.1, seed=seed)
        else:
            side = max(8, int(round(np.sqrt(target))))
            meshes[i] = planar_grid(side, side, 1.0, jitter=0.2, seed=seed)
    return meshes


def mean_face_edge_length(vertices: np.ndarray, faces: np.ndarray) -> float:
    """Mean length over every face's three edges (an interior edge twice), in float64."""
    tri = np.asarray(faces).reshape(-1, 3)
    p = np.asarray(vertices, dtype=np.float64)
    lengths = [
        np.linalg.norm(p[tri[:, 1]] - p[tri[:, 0]], axis=1),
        np.linalg.norm(p[tri[:, 2]] - p[tri[:, 1]], axis=1),
        np.linalg.norm(p[tri[:, 0]] - p[tri[:, 2]], axis=1),
    ]
    return float(np.mean(np.concatenate(lengths))) if tri.shape[0] else 0.0


def random_sources(n_vertices: int, count: int, seed: int) -> np.ndarray:
    """``count`` distinct vertex indices drawn uniformly from a seed, as int32."""
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(n_vertices, size=min(count, n_vertices), replace=False)).astype(
        np.int32
    )
