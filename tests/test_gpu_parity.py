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
