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
            block = np.vstack([W11, -W21])  # the inverse factor Z
            base = plan.w_off[s]
            w[base : base + m * nc] = block.T.reshape(-1)  # column-major, leading dimension m
    return w


TILE = 64


class _LaunchRecord:
    """Reads and writes of one launch's tasks, for the check that no task reads or writes an entry
    another task of the same launch writes (the device runs them concurrently)."""

    def __init__(self):
        self.writes = {"arena": [], "z": []}
        self.reads = {"arena": [], "z": []}

    def task(self, reads, writes):
        for space in ("arena", "z"):
            self.reads[space].append(np.concatenate(reads.get(space, [np.zeros(0, np.int64)])))
            self.writes[space].append(np.concatenate(writes.get(space, [np.zeros(0, np.int64)])))

    def check(self, sizes):
        for space, size in sizes.items():
            writes = self.writes[space]
            if not writes:
                continue
            allw = np.concatenate(writes)
            assert np.unique(allw).size == allw.size, f"two tasks write one {space} entry"
            owner = np.full(size, -1, dtype=np.int64)
            for t, wt in enumerate(writes):
                owner[wt] = t
            for t, rt in enumerate(self.reads[space]):
                o = owner[rt]
                assert np.all((o == -1) | (o == t)), f"a task reads a {space} entry another writes"


def _idx(off, m, rows, cols):
    """Flat column-major indices (leading dimension m) of rows x cols at off."""
    return (off + np.asarray(cols, np.int64)[None, :] * m + np.asarray(rows, np.int64)[:, None])


def eliminate_block(D):
    """The panel kernel's in-LDS elimination: returns L_D^-1 of the lower triangle of D, built by
    applying each pivot's row operations to the identity."""
    b = D.shape[0]
    D = np.tril(D).copy()
    Z = np.eye(b)
    for k in range(b):
        d = D[k, k]
        assert d > 0.0
        lik = D[k + 1 :, k] / d
        for i in range(k + 1, b):
            D[i, k + 1 : i + 1] -= lik[i - k - 1] * D[k + 1 : i + 1, k]
        Z[k + 1 :, : k + 1] -= np.outer(lik, Z[k, : k + 1])
        Z[k, : k + 1] /= np.sqrt(d)
    return Z


def emulate_factor_tasks(plan, values, pinned=None, check_races=False):
    """
    Runs the library's factorization launches (``hg_plan_factor_tasks``) tile by tile with the
    device kernels' arithmetic, every task of a launch against the state before the launch, and
    returns the inverse factors Z. With ``check_races`` each launch is also checked for two tasks
    touching one entry that either writes.
    """
    arena = np.zeros(plan.arena_size)
    v = values[plan.entry_src].copy()
    if pinned is not None:
        hit = pinned[plan.entry_row] | pinned[plan.entry_col]
        v[hit] = np.where(plan.entry_row[hit] == plan.entry_col[hit], 1.0, 0.0)
    arena[plan.entry_dst] = v
    z = np.zeros(plan.w_size)
    sizes = {"arena": arena.size, "z": z.size}
    for kind, level, first, count in plan.factor_launches:
        record = _LaunchRecord()
        updates = []  # (space, flat indices, new values), applied after every task ran
        for s, f1, f2, f3 in plan.factor_tasks[first : first + count]:
            assert plan.level[s] == level
            nc = int(plan.ncol[s])
            m = nc + int(plan.nrow[s])
            fo, zo = int(plan.front_off[s]), int(plan.w_off[s])
            if kind == 0:  # extend-add of child f1's tile (f2, f3)
                c = int(f1)
                ncc, rc = int(plan.ncol[c]), int(plan.nrow[c])
                mc = ncc + rc
                mp = plan.rel[plan.row_ptr[c] : plan.row_ptr[c] + rc]
                rows = np.arange(f2, min(f2 + TILE, rc))
                cols = np.arange(f3, min(f3 + TILE, rc))
                ii, kk = np.meshgrid(rows, cols, indexing="ij")
                keep = ii >= kk
                ii, kk = ii[keep], kk[keep]
                src = int(plan.front_off[c]) + (ncc + kk) * mc + ncc + ii
                dst = fo + mp[kk].astype(np.int64) * m + mp[ii]
                updates.append(("arena", dst, arena[dst] + arena[src]))
                record.task({"arena": [src, dst]}, {"arena": [dst]})
                continue
            j0 = int(f1)
            b = min(TILE, nc - j0)
            J = np.arange(j0, j0 + b)
            if kind == 1:  # panel
                dblock = _idx(fo, m, J, J)
                Zd = eliminate_block(arena[dblock])
                low = np.tril(np.ones((b, b), dtype=bool))
                if f2 == 2:
                    dst = _idx(zo, m, J, J)[low]
                    updates.append(("z", dst, Zd[low]))
                    record.task({"arena": [dblock.ravel()]}, {"z": [dst]})
                elif f2 == 0:
                    rows = np.arange(f3, min(f3 + TILE, m))
                    at = _idx(fo, m, rows, J)
                    updates.append(("arena", at.ravel(), (arena[at] @ Zd.T).ravel()))
                    record.task({"arena": [dblock.ravel(), at.ravel()]}, {"arena": [at.ravel()]})
                else:
                    cols = np.arange(f3, f3 + TILE)
                    at = _idx(zo, m, J, cols)
                    updates.append(("z", at.ravel(), (Zd @ z[at]).ravel()))
                    record.task({"arena": [dblock.ravel()], "z": [at.ravel()]}, {"z": [at.ravel()]})
            else:  # update
                rows = np.arange(f2, min(f2 + TILE, m))
                P = _idx(fo, m, rows, J)
                if f3 >= j0 + b:
                    cols = np.arange(f3, min(f3 + TILE, m))
                    K = _idx(fo, m, cols, J)
                    at = _idx(fo, m, rows, cols)
                    low = rows[:, None] >= cols[None, :]
                    new = arena[at] - arena[P] @ arena[K].T
                    updates.append(("arena", at[low], new[low]))
                    record.task({"arena": [P.ravel(), K.ravel(), at[low]]}, {"arena": [at[low]]})
                else:
                    cols = np.arange(f3, min(f3 + TILE, j0 + b))
                    B = _idx(zo, m, J, cols)
                    at = _idx(zo, m, rows, cols)
                    new = z[at] - arena[P] @ z[B]
                    updates.append(("z", at.ravel(), new.ravel()))
                    record.task({"arena": [P.ravel()], "z": [B.ravel(), at.ravel()]}, {"z": [at.ravel()]})
        if check_races:
            record.check(sizes)
        for space, dst, val in updates:
            (arena if space == "arena" else z)[dst] = val
    return z


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
            Z = w[plan.w_off[s] : plan.w_off[s] + m * nc].reshape(nc, m).T
            head = fs[:nc].copy()
            y[c0 : c0 + nc] = Z[:nc] @ head
            fs[nc:] += Z[nc:] @ head
    for level in range(plan.n_levels - 1, -1, -1):
        for s in plan.level_nodes[plan.level_ptr[level] : plan.level_ptr[level + 1]]:
            nc, nr, c0 = plan.ncol[s], plan.nrow[s], plan.col_begin[s]
            m = nc + nr
            Z = w[plan.w_off[s] : plan.w_off[s] + m * nc].reshape(nc, m).T
            rows = plan.rows[plan.row_ptr[s] : plan.row_ptr[s + 1]]
            x[c0 : c0 + nc] = Z[:nc].T @ y[c0 : c0 + nc] + Z[nc:].T @ x[rows]
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


@pytest.mark.parametrize(
    "name,leaf,top",
    [("icosphere3", 0, 0), ("icosphere4", 0, 0), ("torus", 0, 0), ("torus", 8, 16), ("grid", 0, 0),
     ("bumpy", 8, 0), ("saddle_graded", 0, 0)],
)
def test_factor_tasks_give_the_inverse_factor(name, leaf, top):
    """The blocked factorization's launches, run task by task (each launch against the state before
    it, and checked for two tasks of one launch touching an entry either writes), give the inverse
    factors the column-by-column factorization gives, and a solve through them matches SciPy."""
    vertices, _, ops, plan = _plan_for(name, leaf_rows=leaf, top_rows=top)
    z = emulate_factor_tasks(plan, ops.heat, check_races=True)
    z_ref = emulate_factor(plan, ops.heat)
    assert np.max(np.abs(z - z_ref)) <= 1e-12 * np.max(np.abs(z_ref))
    b = np.random.default_rng(2).random(plan.n)
    x = np.empty(plan.n)
    x[plan.perm] = emulate_solve(plan, z, b[plan.perm])
    expected = spla.spsolve(ops.matrix(ops.heat).tocsc(), b)
    assert np.max(np.abs(x - expected)) <= 1e-10 * np.max(np.abs(expected))


def test_factor_tasks_cover_multi_panel_fronts():
    """The fronts the tests above factor include ones wider than a panel (several steps, a last
    panel narrower than 64), so every task kind runs."""
    _, _, _, plan = _plan_for("icosphere4")
    kinds = {int(k) for k in plan.factor_launches[:, 0]}
    assert kinds == {0, 1, 2}
    assert np.any(plan.ncol > TILE) and np.any((plan.ncol > TILE) & (plan.ncol % TILE != 0))
    panel = plan.factor_tasks[
        np.concatenate([np.arange(f, f + c) for k, _, f, c in plan.factor_launches if k == 1])
    ]
    assert {0, 1, 2} <= set(panel[:, 2].tolist())


def test_factor_tasks_pinned_poisson_matches_scipy():
    vertices, faces, ops, plan = _plan_for("torus")
    n = plan.n
    pinned_new = np.zeros(n, dtype=bool)
    pinned_new[n - 1] = True
    z = emulate_factor_tasks(plan, ops.poisson, pinned=pinned_new)
    b = np.random.default_rng(3).random(n)
    b -= b.mean()
    rp = b[plan.perm].copy()
    rp[n - 1] = 0.0
    x = np.empty(n)
    x[plan.perm] = emulate_solve(plan, z, rp)
    pin_old = int(plan.perm[-1])
    keep = np.ones(n, dtype=bool)
    keep[pin_old] = False
    expected = np.zeros(n)
    expected[keep] = spla.spsolve(ops.matrix(ops.poisson)[keep][:, keep].tocsc(), b[keep])
    assert np.max(np.abs(x - expected)) <= 1e-8 * np.max(np.abs(expected))


def test_eliminate_block_is_the_inverse_cholesky_factor():
    rng = np.random.default_rng(4)
    for b in (1, 2, 5, 64):
        a = rng.standard_normal((b, b))
        spd = a @ a.T + b * np.eye(b)
        L = np.linalg.cholesky(spd)
        assert np.allclose(eliminate_block(spd), np.linalg.inv(L), rtol=0, atol=1e-12)


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


def test_threaded_analysis_is_deterministic():
    """The dissection, symbolic rows and entry maps run on threads (a split's halves, separate
    components, a level's supernodes, ranges of rows): the plan does not depend on the order the
    threads finish in. A 98k-vertex torus (splits large enough to go to threads) and forty packed
    tori (one component a thread), each analysed three times, give the same arrays every time."""
    from ordito_amd import analysis, creation

    cases = [creation.torus(384, 256, 1.0, 0.3, 0.1, 1)]
    vs, fs, off = [], [], 0
    for i in range(40):
        v, f = creation.torus(24 + i, 12 + i // 2, 1.0, 0.3, 0.1, i)
        vs.append(v + np.array([3.0 * i, 0.0, 0.0], dtype=np.float32))
        fs.append(np.asarray(f).reshape(-1, 3) + off)
        off += len(v)
    cases.append((np.concatenate(vs), np.concatenate(fs)))
    fields = ("perm", "ncol", "nrow", "parent", "level", "rows", "rel", "entry_src", "entry_dst")
    for vertices, faces in cases:
        pattern = ho.operator_pattern(len(vertices), np.asarray(faces).reshape(-1, 3))
        plans = [
            analysis.analyse(pattern.indptr, pattern.indices, vertices, leaf_rows=64)
            for _ in range(3)
        ]
        for name in fields:
            first = getattr(plans[0], name)
            assert all(np.array_equal(first, getattr(p, name)) for p in plans[1:]), name
