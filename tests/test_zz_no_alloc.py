"""
hg_shift_to_sources_noalloc against hg_shift_to_sources through the C ABI: the same bits on two
calls in a row, the second on a smaller mesh with fewer columns (the shared device array keeps the
first call's contents), and sizes past the array's limits refused before any device work.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

from tests.conftest import mesh

pytestmark = pytest.mark.gpu


def _tensor(cuda, array):
    return cuda.from_numpy(np.ascontiguousarray(array)).cuda()


def _shift_both(cuda, name, columns, seed):
    """Shift the same random fields with both entry points; returns (allocating, no-alloc)."""
    import math

    from ordito_amd import _lib, heat

    vertices, faces = mesh(name)
    n = len(vertices)
    rng = np.random.default_rng(seed)
    fields = rng.standard_normal((columns, n)) + rng.standard_normal((columns, 1))
    counts = rng.integers(0, 5, size=columns)
    counts[0] = 1
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    sources = rng.integers(0, n, size=int(offsets[-1])).astype(np.int32)
    # A leading pad before the first query's sources: the offsets need not start at zero.
    padded = np.concatenate([np.array([7, 7], np.int32), sources])
    host_offsets = _lib.offsets_array(offsets + 2)
    handle = heat._Handle(_tensor(cuda, vertices), _tensor(cuda, faces))
    try:
        _lib.call("hg_operators", handle.ptr, math.nan, 0, 1)
        d_sources = _tensor(cuda, padded)
        a = _tensor(cuda, fields.reshape(-1))
        b = a.clone()
        _lib.call("hg_shift_to_sources", handle.ptr, a.data_ptr(), d_sources.data_ptr(),
                  host_offsets, columns)
        print(f"no-alloc shift: {name}, {columns} columns", flush=True)
        _lib.call("hg_shift_to_sources_noalloc", handle.ptr, b.data_ptr(), d_sources.data_ptr(),
                  host_offsets, columns)
        return a.cpu().numpy(), b.cpu().numpy(), fields, offsets, sources
    finally:
        handle.close()


def test_noalloc_shift_equals_the_allocating_shift_twice(cuda):
    for name, columns, seed in (("torus", 40, 1), ("icosphere3", 5, 2)):
        a, b, fields, offsets, sources = _shift_both(cuda, name, columns, seed)
        assert np.array_equal(a, b), name
        # And the shift itself: the sources' mean is zero wherever a column has sources.
        got = b.reshape(columns, -1)
        for q in range(columns):
            mine = sources[offsets[q] : offsets[q + 1]]
            if len(mine):
                assert abs(got[q, mine].mean()) <= 1e-12 * np.abs(fields[q]).max() * 4
            else:
                assert np.array_equal(got[q], fields[q]) or np.array_equal(got[q], -fields[q])


def test_noalloc_shift_refuses_sizes_past_its_array(cuda):
    import math

    from ordito_amd import _lib, heat

    vertices, faces = mesh("torus")
    handle = heat._Handle(_tensor(cuda, vertices), _tensor(cuda, faces))
    try:
        _lib.call("hg_operators", handle.ptr, math.nan, 0, 1)
        too_many = 4097
        offsets = _lib.offsets_array(np.zeros(too_many + 1, np.int64))
        with pytest.raises(ValueError, match="at most 4096 columns"):
            _lib.call("hg_shift_to_sources_noalloc", handle.ptr, ctypes.c_void_p(0),
                      ctypes.c_void_p(0), offsets, too_many)
        bad = _lib.offsets_array([0, 3, 1])
        with pytest.raises(ValueError, match="non-decreasing"):
            _lib.call("hg_shift_to_sources_noalloc", handle.ptr, ctypes.c_void_p(0),
                      ctypes.c_void_p(0), bad, 2)
    finally:
        handle.close()
