"""
The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py): the
oracle still reproduces them (CPU), and the HIP path matches them in every mode (GPU).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from oracle import heat_oracle as ho

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def _load(name):
    data = np.load(GOLDEN / f"{name}.npz")
    lengths = data["source_lengths"]
    flat = data["sources"]
    sets = np.split(flat, np.cumsum(lengths)[:-1])
    return data["vertices"], data["faces"], sets, data["distance"]


def test_fixtures_exist():
    assert len(CASES) >= 4


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden(name):
    vertices, faces, sets, expected = _load(name)
    for q, s in enumerate(sets):
        got = ho.geodesic(vertices, faces, s).distance
        assert np.allclose(got, expected[q], rtol=0, atol=1e-12 * np.ptp(expected[q]))


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_matches_golden(cuda, name, precision):
    """Iterated on the first call, factored from the second (rent or buy), all columns at once on
    the third: each within 1e-5 of the range (mixed: 1e-4, the served accuracy bar)."""
    from ordito_amd import heat

    vertices, faces, sets, expected = _load(name)
    mesh = heat.Trimesh(cuda.from_numpy(vertices).cuda(), cuda.from_numpy(faces).cuda())
    solver = mesh.heat_solver(method="meshlab", precision=precision)
    bar = 1e-5 if precision == "fp64" else 1e-4
    for call in range(2):
        for q, s in enumerate(sets):
            got = solver.geodesic([cuda.from_numpy(s).cuda()])[0].cpu().numpy()
            worst, _ = ho.relative_error(got, expected[q])
            assert worst <= bar, (call, q, worst)
    batch = solver.geodesic([cuda.from_numpy(s).cuda() for s in sets]).cpu().numpy()
    for q in range(len(sets)):
        worst, _ = ho.relative_error(batch[q], expected[q])
        assert worst <= bar, (q, worst)
