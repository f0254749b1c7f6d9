"""Pins the oracle's approx_nearest_one (oracle/kiddo_oracle.cpp or_approx_nearest_one)
with what the mathematics fixes, computed independently in numpy:

  * the straight descent of the reference's default `get_leaf_idx`
    (src/traits/stem_strategy.rs:182-200, called by straight_query,
    src/kd_tree/orchestrator/mod.rs:141-157): from Eytzinger index 1, at level l the
    dimension is l mod K and the query goes right iff q[dim] >= pivot (a +inf sentinel
    pivot therefore sends every finite query left), child = 2 idx + right, until the
    level passes max_stem_level; the leaf is idx - 2^(max_stem_level + 1);
  * then the minimum over that leaf alone (src/kd_tree/query/approx_nearest_one.rs:49-71:
    best starts at (A::max_value(), T::default()) and is replaced by a strictly smaller
    distance; the canonical rule takes the lower item among equal distances);
  * distances in Kiddo's scalar order, without FMA (tests/engines.py scalar_order_sqdist).

Queries sit exactly on pivots, inside and outside the data, on all-equal and grid data,
for f32 / f64 and K = 2, 3, 4, one-leaf and empty trees. Planted mistakes (`>` for `>=`,
a shifted dimension cycle, the nearest leaf entry chosen by item first) give answers that
differ from the oracle's, so each would fail this test.
"""
import numpy as np
import pytest

from engines import assert_same_bits, scalar_order_sqdist


def _descend(T, q, op=">=", dim_shift=0):
    """Leaf index per query by the straight descent (vectorised over queries)."""
    m, K = q.shape
    D = T.max_stem_level + 1
    if D <= 0:
        return np.zeros(m, np.int64)
    stems = T.stems()
    idx = np.ones(m, np.int64)
    rows = np.arange(m)
    for level in range(D):
        pivot = stems[idx]
        dim = (level + dim_shift) % K
        qd = q[rows, dim]
        right = (qd >= pivot) if op == ">=" else (qd > pivot)
        idx = 2 * idx + right.astype(np.int64)
    return idx - (1 << D)


def approx_expected(T, q, out_dtype, op=">=", dim_shift=0, item_first=False):
    m = q.shape[0]
    items = np.zeros(m, np.uint32)
    dists = np.full(m, np.inf, out_dtype)
    if T.size == 0:
        return items, dists
    leaves = T.leaves()
    leaf = _descend(T, q, op, dim_shift)
    for i in range(m):
        pts, its = leaves[int(leaf[i])]
        if len(its) == 0:
            continue
        d = scalar_order_sqdist(pts, q[i][None, :], out_dtype)
        order = np.lexsort((d, its)) if item_first else np.lexsort((its, d))
        j = order[0]
        items[i] = its[j]
        dists[i] = d[j]
    return items, dists


def _queries(pts, T, m, rng):
    """Uniform queries over a box 10% larger than the data, a quarter of them copies of
    data points, and every fourth one moved onto a pivot along that level's dimension."""
    K = pts.shape[1]
    lo, hi = pts.min(0).astype(np.float64), pts.max(0).astype(np.float64)
    span = np.maximum(hi - lo, 1.0)
    q = (lo - 0.05 * span + rng.random((m, K)) * 1.1 * span).astype(pts.dtype)
    q[: m // 4] = pts[rng.integers(0, len(pts), m // 4)]
    D = T.max_stem_level + 1
    if D > 0:
        stems = T.stems()
        for i in range(0, m, 4):
            # walk a random path and put the query exactly on the pivot it meets at a
            # random level (the `>=` decision there is the one a `>` would flip)
            lvl = int(rng.integers(0, D))
            idx = 1
            for level in range(lvl):
                idx = 2 * idx + int(q[i, level % K] >= stems[idx])
            pv = stems[idx]
            if np.isfinite(pv):
                q[i, lvl % K] = pv
    return q


def _points(kind, n, K, dtype, rng):
    if kind == "uniform":
        p = rng.random((n, K))
    elif kind == "grid":
        p = rng.integers(0, 6, (n, K)).astype(np.float64)
    elif kind == "clusters":
        c = rng.random((6, K))
        p = c[rng.integers(0, 6, n)] + rng.normal(0, 0.003, (n, K))
    elif kind == "all_same":
        p = np.full((n, K), 0.625)
    else:
        raise ValueError(kind)
    return p.astype(dtype)


CASES = [
    ("uniform", 20_000, 3, np.float64, 32),
    ("uniform", 20_000, 3, np.float32, 32),
    ("grid", 8_000, 3, np.float64, 32),
    ("grid", 5_000, 2, np.float32, 8),
    ("clusters", 12_000, 4, np.float64, 16),
    ("uniform", 9_000, 4, np.float32, 32),
    ("uniform", 6_000, 2, np.float64, 4),
    ("all_same", 500, 3, np.float64, 32),
    ("uniform", 20, 3, np.float64, 32),  # one leaf (no stems)
    ("uniform", 1, 2, np.float32, 32),
]


@pytest.mark.parametrize("kind,n,K,dtype,B", CASES,
                         ids=[f"{c[0]}-{c[1]}-K{c[2]}-{np.dtype(c[3]).name}-B{c[4]}" for c in CASES])
def test_approx_matches_straight_descent(oracle_mod, kind, n, K, dtype, B):
    rng = np.random.default_rng(n * 10 + K)
    pts = _points(kind, n, K, dtype, rng)
    T = oracle_mod.Tree.build(pts, B)
    q = _queries(pts, T, 3000, rng)
    oi, od = T.approx_nearest_one(q)
    ei, ed = approx_expected(T, q, np.dtype(dtype).type)
    assert_same_bits(oi, ei, "approx items")
    assert_same_bits(od, ed, "approx distances")
    # the same result on 4 threads
    ti, td = T.approx_nearest_one(q, threads=4)
    assert_same_bits(ti, oi, "threaded approx items")
    assert_same_bits(td, od, "threaded approx distances")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_approx_planted_mistakes_differ(oracle_mod, dtype):
    """Each planted mistake gives answers that differ from the oracle's on these inputs,
    so the straight-descent check above would catch it in the oracle."""
    rng = np.random.default_rng(5)
    pts = _points("grid", 8_000, 3, dtype, rng)
    T = oracle_mod.Tree.build(pts, 32)
    q = _queries(pts, T, 3000, rng)
    oi, od = T.approx_nearest_one(q)
    out = np.dtype(dtype).type
    for what, kw in (("'>' for '>='", {"op": ">"}), ("dimension cycle shifted", {"dim_shift": 1}),
                     ("item before distance", {"item_first": True})):
        pi, pd = approx_expected(T, q, out, **kw)
        differs = (pi != oi) | (pd.view(np.uint32 if out is np.float32 else np.uint64)
                                != od.view(np.uint32 if out is np.float32 else np.uint64))
        assert differs.any(), f"planted mistake ({what}) is not caught"


def test_approx_empty_tree(oracle_mod):
    T = oracle_mod.Tree.build(np.zeros((0, 3)), 32)
    i, d = T.approx_nearest_one(np.array([[0.5, 0.5, 0.5], [2.0, -1.0, 0.0]]))
    assert i.tolist() == [0, 0] and np.all(np.isinf(d))


def test_approx_is_nearest_of_its_leaf_and_at_least_exact(oracle_mod):
    """Two properties that need no descent: the approx distance is never below the exact
    nearest distance, and it equals it whenever the exact nearest entry lies in the
    landing leaf."""
    rng = np.random.default_rng(9)
    pts = rng.random((30_000, 3))
    T = oracle_mod.Tree.build(pts, 32)
    q = rng.random((4000, 3))
    ai, ad = T.approx_nearest_one(q)
    ni, nd = T.nearest_one(q)
    assert np.all(ad >= nd)
    leaf = _descend(T, q)
    leaves = T.leaves()
    same_leaf = np.array([ni[i] in set(leaves[int(leaf[i])][1].tolist()) for i in range(len(q))])
    assert same_leaf.mean() > 0.5
    assert_same_bits(ad[same_leaf], nd[same_leaf], "approx == exact when the nearest is in the leaf")
    assert_same_bits(ai[same_leaf], ni[same_leaf], "approx item == exact item then")
