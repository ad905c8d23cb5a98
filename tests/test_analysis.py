"""
The Cholesky analysis on the CPU: plan invariants, every index the device kernels form checked
against its buffer, and a numpy emulation of the device algorithm (scatter, extend-add, partial
factorization, partitioned-inverse solve blocks, multifrontal forward sweep, backward sweep) through
the same maps, against SciPy's direct solve -- the check ordito's own Cholesky tests make
(``tests/test_cholesky.py:38-98``, 1e-12 after refinement; here 1e-10 unrefined on
well-conditioned heat systems).
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import heat_oracle as ho
from tests.conftest import mesh


def _plan_for(name, leaf_rows=0, top_rows=0):
    from ordito_amd import analysis

    vertices, faces = mesh(name)
    ops = ho.assemble(vertices, faces, None)
    plan = analysis.analyse(
        ops.pattern.indptr, ops.pattern.indices, vertices, leaf_rows=leaf_rows, top_rows=top_rows
    )
    return vertices, faces, ops, plan


def _front_size(plan, s):
    return int(plan.ncol[s] + plan.nrow[s])


def check_plan_invariants(plan, n):
    perm = plan.perm
    assert np.array_equal(np.sort(perm), np.arange(n))
    S = plan.n_super
    # Columns: contiguous, in postorder, covering 0..n-1.
    assert plan.col_begin[0] == 0
    assert np.array_equal(plan.col_begin[1:], plan.col_begin[:-1] + plan.ncol[:-1])
    assert plan.col_begin[-1] + plan.ncol[-1] == n
    assert np.all(plan.ncol > 0)
    for s in range(S):
        rows = plan.rows[plan.row_ptr[s] : plan.row_ptr[s + 1]]
        last = plan.col_begin[s] + plan.ncol[s] - 1
        assert np.all(np.diff(rows) > 0) and (rows.size == 0 or rows[0] > last)
        p = plan.parent[s]
        if p < 0:
            assert rows.size == 0
            continue
        assert p > s and plan.level[p] > plan.level[s]
        rel = plan.rel[plan.row_ptr[s] : plan.row_ptr[s + 1]]
        mp = _front_size(plan, p)
        assert np.all((rel >= 0) & (rel < mp)) and np.all(np.diff(rel) > 0)
        # Each relative position names the same global row in the parent's front.
        parent_index = np.concatenate(
            [
                np.arange(plan.col_begin[p], plan.col_begin[p] + plan.ncol[p]),
                plan.rows[plan.row_ptr[p] : plan.row_ptr[p + 1]],
            ]
        )
        assert np.array_equal(parent_index[rel], rows)
    # Buffers: fronts, solve blocks and vectors tile their arenas exactly.
    m = plan.ncol + plan.nrow
    assert np.array_equal(plan.front_off, np.concatenate([[0], np.cumsum(m * m)[:-1]]))
    assert plan.arena_size == int(np.sum(m * m))
    assert np.array_equal(plan.w_off, np.concatenate([[0], np.cumsum(m * plan.ncol)[:-1]]))
    assert plan.w_size == int(np.sum(m * plan.ncol))
    assert plan.vec_size == int(np.sum(m))
    assert np.all((plan.entry_dst >= 0) & (plan.entry_dst < plan.arena_size))
    # Levels: every supernode once, children strictly below their parent.
    assert np.array_equal(np.sort(plan.level_nodes), np.arange(S))


def emulate_factor(plan, values, pinned=None):
    arena = np.zeros(plan.arena_size)
    v = values[plan.entry_src].copy()
    if pinned is not None:
        hit = pinned[plan.entry_row] | pinned[plan.entry_col]
        v[hit] = np.where(plan.entry_row[hit] == plan.entry_col[hit], 1.0, 0.0)
    arena[plan.entry_dst] = v
    w = np.zeros(plan.w_size)

    def front(s):
        m = _front_size(plan, s)
        base = plan.front_off[s]
        return arena[base : base + m * m].reshape(m, m).T  # column-major view: F[i, j]

    for level in range(plan.n_levels):
        for s in plan.level_nodes[plan.level_ptr[level] : plan.level_ptr[level + 1]]:
            F = front(s)
            for c in plan.children[plan.child_ptr[s] : plan.child_ptr[s + 1]]:
                nc = plan.ncol[c]
                rel = plan.rel[plan.row_ptr[c] : plan.row_ptr[c + 1]]
                U = front(c)[nc:, nc:]
                low = np.tril(U)
                F[np.ix_(rel, rel)] += low
            nc = plan.ncol[s]
            m = F.shape[0]
            for j in range(nc):
                d = F[j, j]
                assert d > 0.0
                F[j, j] = np.sqrt(d)
                F[j + 1 :, j] /= F[j, j]
                for k in range(j + 1, m):
                    F[k:, k] -= F[k:, j] * F[k, j]
            L11 = np.tril(F[:nc, :nc])
            W11 = np.linalg.inv(L11)
            W21 = F[nc:, :nc] @ W11
            block = np.vstack([W11, W21])
            base = plan.w_off[s]
            w[base : base + m * nc] = block.T.reshape(-1)  # column-major, leading dimension m
    return w


def emulate_solve(plan, w, b_perm):
    n = plan.n
    y = np.zeros(n)
    x = np.zeros(n)
    f = np.zeros(plan.vec_size)
    for level in range(plan.n_levels):
        for s in plan.level_nodes[plan.level_ptr[level] : plan.level_ptr[level + 1]]:
            nc, nr, c0 = plan.ncol[s], plan.nrow[s], plan.col_begin[s]
            m = nc + nr
            fs = f[plan.vec_off[s] : plan.vec_off[s] + m]
            fs[:nc] = b_perm[c0 : c0 + nc]
            fs[nc:] = 0.0
            for c in plan.children[plan.child_ptr[s] : plan.child_ptr[s + 1]]:
                rel = plan.rel[plan.row_ptr[c] : plan.row_ptr[c + 1]]
                fc = f[plan.vec_off[c] + plan.ncol[c] : plan.vec_off[c] + plan.ncol[c] + plan.nrow[c]]
                fs[rel] += fc
            W = w[plan.w_off[s] : plan.w_off[s] + m * nc].reshape(nc, m).T
            head = fs[:nc].copy()
            y[c0 : c0 + nc] = W[:nc] @ head
            fs[nc:] -= W[nc:] @ head
    for level in range(plan.n_levels - 1, -1, -1):
        for s in plan.level_nodes[plan.level_ptr[level] : plan.level_ptr[level + 1]]:
            nc, nr, c0 = plan.ncol[s], plan.nrow[s], plan.col_begin[s]
            m = nc + nr
            W = w[plan.w_off[s] : plan.w_off[s] + m * nc].reshape(nc, m).T
            rows = plan.rows[plan.row_ptr[s] : plan.row_ptr[s + 1]]
            x[c0 : c0 + nc] = W[:nc].T @ y[c0 : c0 + nc] - W[nc:].T @ x[rows]
    return x


@pytest.mark.parametrize("name", ["icosphere3", "torus", "grid", "bumpy", "saddle_graded"])
def test_plan_invariants(name):
    vertices, _, _, plan = _plan_for(name)
    check_plan_invariants(plan, vertices.shape[0])


@pytest.mark.parametrize("leaf_rows", [8, 32, 64])
def test_plan_invariants_across_leaf_sizes(leaf_rows):
    vertices, _, _, plan = _plan_for("torus", leaf_rows=leaf_rows, top_rows=64)
    check_plan_invariants(plan, vertices.shape[0])


@pytest.mark.parametrize("name", ["icosphere3", "torus", "grid", "bumpy"])
def test_emulated_heat_solve_matches_scipy(name):
    vertices, _, ops, plan = _plan_for(name)
    A = ops.matrix(ops.heat)
    w = emulate_factor(plan, ops.heat)
    rng = np.random.default_rng(0)
    b = rng.random(plan.n)
    x_perm = emulate_solve(plan, w, b[plan.perm])
    x = np.empty(plan.n)
    x[plan.perm] = x_perm
    expected = spla.spsolve(A.tocsc(), b)
    assert np.max(np.abs(x - expected)) <= 1e-10 * np.max(np.abs(expected))


def test_emulated_pinned_poisson_solve_matches_scipy():
    vertices, faces, ops, plan = _plan_for("torus")
    n = plan.n
    pin_old = int(plan.perm[-1])  # the last-eliminated row
    pinned_new = np.zeros(n, dtype=bool)
    pinned_new[n - 1] = True
    w = emulate_factor(plan, ops.poisson, pinned=pinned_new)
    rng = np.random.default_rng(1)
    b = rng.random(n)
    b -= b.mean()
    rp = b[plan.perm].copy()
    rp[n - 1] = 0.0
    x_perm = emulate_solve(plan, w, rp)
    x = np.empty(n)
    x[plan.perm] = x_perm
    keep = np.ones(n, dtype=bool)
    keep[pin_old] = False
    A = ops.matrix(ops.poisson)
    reduced = A[keep][:, keep]
    expected = np.zeros(n)
    expected[keep] = spla.spsolve(reduced.tocsc(), b[keep])
    assert abs(x[pin_old]) == 0.0
    assert np.max(np.abs(x - expected)) <= 1e-8 * np.max(np.abs(expected))


@pytest.mark.parametrize("name", ["icosphere3", "torus", "grid", "bumpy", "saddle_graded"])
def test_sweep_tasks_stay_in_bounds(name):
    """The library's own check of every device sweep index (task ranges, partial slots, gather
    sources) against its buffers, run on the host."""
    _plan_for_validated(name)


def test_sweep_tasks_stay_in_bounds_at_benchmark_scale():
    from ordito_amd import analysis, creation

    vertices, faces = creation.torus(1024, 512, 1.0, 0.3, 0.1, 0)
    pattern = ho.operator_pattern(len(vertices), faces)
    analysis.analyse(pattern.indptr, pattern.indices, vertices, validate=True)


def _plan_for_validated(name):
    from ordito_amd import analysis

    vertices, faces = mesh(name)
    pattern = ho.operator_pattern(len(vertices), faces)
    for leaf in (8, 48):
        analysis.analyse(pattern.indptr, pattern.indices, vertices, leaf_rows=leaf, validate=True)
