"""
Ill-shaped test surfaces, written from their geometric description.

Provenance: ordito's test suite names two meshes that stress a heat solver -- a saddle grid graded
towards x = 0 (``saddle_graded``, ordito ``tests/conftest.py:471-502``) and a bumpy star-shaped
hull of random directions (``sphere_irregular``, ``tests/conftest.py:595-639``). Only their
geometry is taken (the surface formulas, the grading exponent, the point count); the construction
below is this project's own, and the bumpy hull is not the reference's mesh (a different random
draw, order and placement).
"""

from __future__ import annotations

import numpy as np


def graded_saddle(side: int = 34, power: float = 3.4) -> tuple[np.ndarray, np.ndarray]:
    """
    A ``side x side`` vertex grid on the saddle ``z = 0.35 (x^2 - 0.6 y^2)`` over ``[-1, 1]^2``,
    its x spacing graded as ``sign(s) |s|^power`` of the even parameter ``s``: with an even side the
    two finest columns straddle ``x = 0``, ``2 / (side - 1)^power`` apart against a y spacing of
    ``2 / (side - 1)``, so the thinnest triangles reach an aspect ratio near
    ``(side - 1)^(power - 1)`` (about 4 800 at the defaults). Vertex ``(a, b)`` (x column a,
    y row b) has index ``a * side + b``; each cell is cut along the diagonal from its
    ``(a, b)`` corner to its ``(a + 1, b + 1)`` corner, both triangles wound counter-clockwise seen
    from +z. Returns float32 (V, 3) vertices and a flat int32 (3F,) face buffer.
    """
    n = int(side)
    s = np.linspace(-1.0, 1.0, n)
    xs = np.copysign(np.abs(s) ** power, s)
    ys = s
    vertices = np.empty((n * n, 3))
    for a in range(n):
        x = xs[a]
        vertices[a * n : (a + 1) * n, 0] = x
        vertices[a * n : (a + 1) * n, 1] = ys
        vertices[a * n : (a + 1) * n, 2] = 0.35 * (x * x - 0.6 * ys * ys)
    tris = []
    for a in range(n - 1):
        for b in range(n - 1):
            lo_lo, hi_lo = a * n + b, (a + 1) * n + b
            lo_hi, hi_hi = lo_lo + 1, hi_lo + 1
            tris.append((lo_lo, hi_lo, hi_hi))
            tris.append((lo_lo, hi_hi, lo_hi))
    return vertices.astype(np.float32), np.asarray(tris, dtype=np.int32).reshape(-1)


def bumpy_hull(points: int = 500, squeeze: float = 0.97, seed: int = 5) -> tuple[np.ndarray, np.ndarray]:
    """
    A closed, star-shaped, non-convex genus-0 surface with needle triangles: the convex hull of
    ``points`` seeded random unit directions (a Delaunay triangulation of the sphere, degrees 3-12),
    each direction then moved to longitude ``p - squeeze sin p`` (a ~33x compression around
    ``p = 0``, where the triangles become needles) and pushed out to the radius
    ``1 + 0.35 sin(3 t) cos(2 p) + 0.15 cos(5 t)`` of its colatitude t and squeezed longitude p.
    Faces are wound outward (the hull's own facet normals decide), then the vertex order, the face
    order and each face's first corner are shuffled from the seed.
    """
    from scipy.spatial import ConvexHull

    rng = np.random.default_rng(seed)
    # Uniform directions by Archimedes' hat-box theorem: z uniform in [-1, 1], longitude uniform.
    z = rng.uniform(-1.0, 1.0, points)
    a = rng.uniform(-np.pi, np.pi, points)
    ring = np.sqrt(1.0 - z * z)
    d = np.stack([ring * np.cos(a), ring * np.sin(a), z], axis=1)
    hull = ConvexHull(d)
    tri = np.array(hull.simplices, dtype=np.int64)
    # Outward winding: a facet whose corner order disagrees with its outward normal is flipped.
    p0, p1, p2 = d[tri[:, 0]], d[tri[:, 1]], d[tri[:, 2]]
    facing = np.sum(np.cross(p1 - p0, p2 - p0) * hull.equations[:, :3], axis=1)
    flip = facing < 0.0
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    colat = np.arccos(np.clip(d[:, 2], -1.0, 1.0))
    lon = np.arctan2(d[:, 1], d[:, 0])
    lon = lon - squeeze * np.sin(lon)
    radius = 1.0 + 0.35 * np.sin(3.0 * colat) * np.cos(2.0 * lon) + 0.15 * np.cos(5.0 * colat)
    sin_c = np.sin(colat)
    surface = np.stack(
        [radius * sin_c * np.cos(lon), radius * sin_c * np.sin(lon), radius * np.cos(colat)], axis=1
    )
    # Shuffle: new vertex i is old vertex order[i]; faces follow, then their order and start.
    order = rng.permutation(points)
    new_index = np.argsort(order)
    tri = new_index[tri]
    tri = tri[rng.permutation(len(tri))]
    start = rng.integers(0, 3, len(tri))
    tri = np.stack([tri[np.arange(len(tri)), (start + c) % 3] for c in range(3)], axis=1)
    return surface[order].astype(np.float32), tri.astype(np.int32).reshape(-1)


def hub_disk(spokes: int = 400, rings: int = 20, seed: int = 7) -> tuple[np.ndarray, np.ndarray]:
    """
    A polar grid on the paraboloid ``z = 0.3 r^2`` over the unit disk: a centre vertex joined to
    every vertex of the first of ``rings`` rings of ``spokes`` vertices (a hub of degree
    ``spokes``, far past what a row of the divergence's face groups holds), the rings joined as
    quads split along alternating diagonals, angles jittered by a tenth of a spoke. Open, with a
    rim (the last ring).
    """
    rng = np.random.default_rng(seed)
    step = 2.0 * np.pi / spokes
    theta = np.arange(spokes) * step + rng.uniform(-0.1, 0.1, (rings, spokes)) * step
    radius = (np.arange(1, rings + 1) / rings)[:, None]
    x, y = radius * np.cos(theta), radius * np.sin(theta)
    ring_xyz = np.stack([x, y, 0.3 * (x * x + y * y)], axis=-1).reshape(-1, 3)
    vertices = np.concatenate([np.zeros((1, 3)), ring_xyz])

    def at(i, j):  # vertex of ring i (0-based), spoke j
        return 1 + i * spokes + (j % spokes)

    j = np.arange(spokes)
    faces = [np.stack([np.zeros(spokes, np.int64), at(0, j), at(0, j + 1)], axis=1)]
    for i in range(rings - 1):
        a, b, c, d = at(i, j), at(i, j + 1), at(i + 1, j + 1), at(i + 1, j)
        even = (i + j) % 2 == 0
        faces.append(np.where(even[:, None], np.stack([a, c, d], 1), np.stack([a, b, d], 1)))
        faces.append(np.where(even[:, None], np.stack([a, b, c], 1), np.stack([b, c, d], 1)))
    tri = np.concatenate(faces)
    # The surface is a graph over the plane: every face wound with its normal upward.
    p0, p1, p2 = vertices[tri[:, 0]], vertices[tri[:, 1]], vertices[tri[:, 2]]
    down = np.cross(p1 - p0, p2 - p0)[:, 2] < 0.0
    tri[down, 1], tri[down, 2] = tri[down, 2].copy(), tri[down, 1].copy()
    return vertices.astype(np.float32), tri.astype(np.int32).reshape(-1)


# The needle of pinched_grid: an interior face of the 30 x 24 grid.
PINCHED_FACE = 2 * 29 * 23 // 2 + 7


def pinched_grid() -> tuple[np.ndarray, np.ndarray]:
    """
    The jittered 30 x 24 planar grid with one interior triangle's third corner moved onto the
    midpoint of its opposite side: a needle of (nearly) zero area, the degenerate face that
    ``use_robust``'s mollified lengths exist for.
    """
    from ordito_amd import creation

    v, f = creation.planar_grid(30, 24, 1.0, 0.2, 4)
    v, tri = np.array(v, np.float32), np.asarray(f).reshape(-1, 3)
    a, b, c = tri[PINCHED_FACE]
    v[c] = (v[a] + v[b]) * np.float32(0.5)
    return v, tri.reshape(-1).astype(np.int32)
