"""Pins the CPU oracle (oracle/kiddo_oracle.cpp) before it is trusted as the checker:
  * every known-answer case the reference's own tests and docs hold,
  * brute force (a linear scan with the canonical (distance, item) rule) on random,
    duplicate-heavy, grid and degenerate inputs, for every family and dtype,
  * an independent numpy evaluation of the scalar-order distance,
  * the layout invariants the builder must satisfy (left < stem <= right, 2^d leaves),
  * Kiddo-faithful vs canonical collectors: same distances, only tie choices differ.
"""
import math

import numpy as np
import pytest

from engines import (OracleEngine, assert_same_bits, load_kats, run_kat_case,
                     scalar_order_sqdist)

KATS = load_kats()


@pytest.mark.parametrize("case", KATS["cases"], ids=lambda c: c["name"])
def test_reference_kat(oracle_mod, case):
    run_kat_case(OracleEngine(oracle_mod), case)


def test_layout_kats(oracle_mod):
    O = oracle_mod
    lc = {c["name"]: c for c in KATS["layout_cases"]}
    # update_pivot (construction.rs:1187-1213)
    c = lc["update_pivot_shifts_right"]
    pts = np.array(c["points"], np.float32)
    si = np.arange(len(pts), dtype=np.uint64)
    piv = O.update_pivot(pts, si, c["dim"], c["init_pivot"])
    assert piv == c["expect_pivot"]
    assert pts[si[piv - 1], 0] == c["expect_left_value"]
    assert pts[si[piv], 0] == c["expect_pivot_value"]
    # single-leaf arena extents and byte length
    c = lc["arena_extents_single_leaf"]
    t = O.Tree.build(np.array(c["points"]), 32, np.array(c["items"], np.uint32))
    assert t.extents().tolist() == c["expect_extents"]
    assert t.arena_len == c["expect_arena_len"]
    assert t.max_stem_level == -1
    # 47-entry leaf: tiles 32, 8, 4, 2, 1 and a lossless decode
    c = lc["arena_tile_lengths_47"]
    n = c["leaf_len"]
    pts = np.stack([np.arange(n), 100 + np.arange(n), 200 + np.arange(n)], 1).astype(np.float64)
    t = O.Tree.build(pts, 64, np.arange(1000, 1000 + n, dtype=np.uint32))
    (lp, li), = t.leaves()
    assert li.tolist() == list(range(1000, 1000 + n))
    assert lp[:, 0].tolist() == list(range(n))
    # tile walk: the byte size of the arena equals the sum over tiles 32+8+4+2+1
    assert t.arena_len == n * (3 * 8 + 4)
    for name in ("random_65536_f32_depth", "eytzinger_100_points_4_leaves"):
        c = lc[name]
        rng = np.random.default_rng(42)
        pts = rng.random((c["n"], c["dims"])).astype(np.float32)
        t = O.Tree.build(pts, c["bucket"])
        assert t.leaf_count == c["expect_leaf_count"]
        assert t.max_stem_level == c["expect_max_stem_level"]


def _check_layout_invariants(t, pts):
    """Every entry lies on the correct side of every stem on its leaf's path."""
    stems = t.stems()
    d = t.max_stem_level + 1
    leaves = t.leaves()
    assert len(leaves) == (1 << d if d > 0 else (1 if t.size else 0))
    assert sum(len(li) for _, li in leaves) == t.size
    all_items = np.concatenate([li for _, li in leaves]) if leaves else np.zeros(0)
    assert np.array_equal(np.sort(all_items), np.arange(t.size))
    K = t.K
    for leaf, (lp, li) in enumerate(leaves):
        assert np.array_equal(lp, pts[li])
        idx = 1
        for level in range(d):
            right = (leaf >> (d - 1 - level)) & 1
            stem = stems[idx]
            dim = level % K
            if np.isinf(stem) if stem.dtype.kind == "f" else stem == np.iinfo(stem.dtype).max:
                assert right == 0 or len(li) == 0
            elif right:
                assert np.all(lp[:, dim] >= stem)
            else:
                assert np.all(lp[:, dim] < stem)
            idx = idx * 2 + right


def _inputs(kind, n, K, dtype, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        pts = rng.random((n, K))
    elif kind == "dups":
        base = rng.random((max(n // 4, 1), K))
        pts = base[rng.integers(0, len(base), n)]
    elif kind == "grid":
        pts = rng.integers(0, 6, (n, K)).astype(np.float64)
    elif kind == "axis_degenerate":
        pts = rng.random((n, K))
        pts[:, 0] = 0.5
    elif kind == "all_same":
        pts = np.full((n, K), 0.25)
    else:
        raise ValueError(kind)
    if np.dtype(dtype).kind == "u":
        scale = 60000 if kind in ("uniform", "dups", "axis_degenerate", "all_same") else 1
        pts = (pts * scale).astype(dtype)
    else:
        pts = pts.astype(dtype)
    return pts


CASES = [
    ("uniform", 3000, 3, np.float64, 32),
    ("uniform", 3000, 3, np.float32, 32),
    ("uniform", 2000, 2, np.float64, 4),
    ("uniform", 2000, 4, np.float32, 16),
    ("dups", 3000, 3, np.float64, 32),
    ("grid", 2000, 3, np.float64, 32),
    ("grid", 2000, 2, np.float32, 8),
    ("axis_degenerate", 1000, 3, np.float64, 32),
    ("all_same", 500, 3, np.float64, 32),
    ("uniform", 3000, 4, np.uint16, 32),
    ("grid", 2000, 3, np.uint16, 16),
    ("uniform", 2000, 3, np.uint32, 32),
    ("uniform", 31, 3, np.float64, 32),
    ("uniform", 33, 3, np.float64, 32),
    ("uniform", 1, 3, np.float64, 32),
]


def _queries(pts, m, seed):
    rng = np.random.default_rng(seed + 1000)
    if pts.dtype.kind == "u":
        q = rng.integers(0, int(pts.max()) + 2, (m, pts.shape[1])).astype(pts.dtype)
    else:
        q = (rng.random((m, pts.shape[1])) * 1.1 - 0.05).astype(pts.dtype)
    # queries exactly on entries exercise zero distances and ties
    q[: m // 4] = pts[rng.integers(0, len(pts), m // 4)]
    return q


def _radius(pts):
    if pts.dtype.kind == "u":
        return 3.0e7 if pts.max() > 100 else 2.0
    return 0.01 if pts.max() <= 1.0 else 2.0


@pytest.mark.parametrize("kind,n,K,dtype,B", CASES,
                         ids=[f"{c[0]}-{c[1]}-K{c[2]}-{np.dtype(c[3]).name}-B{c[4]}" for c in CASES])
def test_oracle_vs_brute_force(oracle_mod, kind, n, K, dtype, B):
    O = oracle_mod
    pts = _inputs(kind, n, K, dtype, seed=n + K)
    t = O.Tree.build(pts, B)
    _check_layout_invariants(t, pts)
    q = _queries(pts, 200, seed=n)
    r = _radius(pts)
    # nearest_one == brute k=1
    it, d = t.nearest_one(q)
    bi, bd, bc = t.brute_nearest_n(q, 1)
    assert_same_bits(it, bi[:, 0], "nearest_one items")
    assert_same_bits(d, bd[:, 0], "nearest_one dists")
    for k in (1, 5, 16):
        it, d, c = t.nearest_n(q, k)
        bi, bd, bc = t.brute_nearest_n(q, k)
        assert_same_bits(c, bc, f"nearest_n({k}) counts")
        assert_same_bits(it, bi, f"nearest_n({k}) items")
        assert_same_bits(d, bd, f"nearest_n({k}) dists")
    for excl in (False, True):
        it, d, c = t.nearest_n(q, 8, r, excl)
        bi, bd, bc = t.brute_nearest_n(q, 8, r, excl)
        assert_same_bits(it, bi, f"nearest_n_within excl={excl}")
        assert_same_bits(d, bd, f"nearest_n_within dists excl={excl}")
        off, wi, wd = t.within(q, r, excl)
        boff, bwi, bwd = t.brute_within(q, r, excl)
        assert_same_bits(off, boff, "within offsets")
        assert_same_bits(wi, bwi, "within items")
        assert_same_bits(wd, bwd, "within dists")
        # unsorted: same multiset per query
        uoff, ui, ud = t.within(q, r, excl, sorted=False)
        assert_same_bits(uoff, boff, "within unsorted offsets")
        for i in range(len(q)):
            s, e = int(uoff[i]), int(uoff[i + 1])
            assert sorted(zip(ud[s:e].tolist(), ui[s:e].tolist())) == list(zip(bwd[s:e].tolist(), bwi[s:e].tolist()))
        it, d, c = t.best_n_within(q, r, 5, excl)
        bi, bd, bc = t.brute_best_n_within(q, r, 5, excl)
        assert_same_bits(it, bi, f"best_n_within excl={excl}")
        assert_same_bits(d, bd, f"best_n_within dists excl={excl}")


def test_distance_matches_independent_numpy(oracle_mod):
    """The oracle's distances equal a numpy evaluation in Kiddo's scalar order bit for bit
    (a dropped term, a fused multiply-add or a different summation order would not)."""
    O = oracle_mod
    for dtype in (np.float64, np.float32):
        pts = _inputs("uniform", 4000, 3, dtype, 7)
        t = O.Tree.build(pts, 32)
        q = _queries(pts, 300, 7)
        it, d, c = t.nearest_n(q, 16)
        ref = scalar_order_sqdist(pts[it], q[:, None, :], dtype)
        assert_same_bits(d, ref, f"numpy scalar order {np.dtype(dtype).name}")
        # a fused or reassociated evaluation differs somewhere on this input
        alt = ((pts[it].astype(np.float64) - q[:, None, :]) ** 2).sum(-1)[..., ::-1]
        if dtype == np.float64:
            p = pts[it]
            rev = np.zeros(p.shape[:-1])
            for dd in (2, 1, 0):
                df = p[..., dd] - q[:, None, dd]
                rev = rev + df * df
            assert np.any(rev.view(np.uint64) != d.view(np.uint64)), "reversed order should differ"
        del alt
    # u16 trees: f64 output equals exact integer arithmetic
    pts = _inputs("uniform", 3000, 4, np.uint16, 3)
    t = O.Tree.build(pts, 32)
    q = _queries(pts, 200, 3)
    it, d, c = t.nearest_n(q, 5)
    exact = ((pts[it].astype(np.int64) - q[:, None, :].astype(np.int64)) ** 2).sum(-1)
    assert np.array_equal(d, exact.astype(np.float64))


def test_faithful_and_canonical_agree_on_distances(oracle_mod):
    """Kiddo's own collectors (mode FAITHFUL) return the same distance sequence as the
    canonical ones; only the item chosen inside a tie group may differ."""
    O = oracle_mod
    pts = _inputs("dups", 4000, 3, np.float64, 11)
    t = O.Tree.build(pts, 32)
    q = _queries(pts, 300, 11)
    ci, cd, cc = t.nearest_n(q, 16)
    fi, fd, fc = t.nearest_n(q, 16, mode=O.FAITHFUL)
    assert_same_bits(cd, fd, "k-NN distances")
    assert_same_bits(cc, fc, "k-NN counts")
    # tie groups hold the same items once the boundary group is set aside
    n1c, d1c = t.nearest_one(q)
    n1f, d1f = t.nearest_one(q, mode=O.FAITHFUL)
    assert_same_bits(d1c, d1f, "nearest_one distances")
    ties = n1c != n1f
    assert np.all(scalar_order_sqdist(pts[n1f[ties]], q[ties], np.float64) == d1c[ties])
    # within sorted: Kiddo's stable distance sort vs canonical differ only inside ties
    off, wi, wd = t.within(q, 0.002)
    foff, fwi, fwd = t.within(q, 0.002, mode=O.FAITHFUL)
    assert_same_bits(off, foff)
    assert_same_bits(wd, fwd)


def test_stats_count_leaf_visits(oracle_mod):
    O = oracle_mod
    pts = _inputs("uniform", 20000, 3, np.float64, 5)
    t = O.Tree.build(pts, 32)
    q = _queries(pts, 1000, 5)
    it, d, c, st = t.nearest_n(q, 16, stats=True)
    leaves, points, stems = (int(x) for x in st)
    assert leaves >= len(q) and points >= 16 * len(q)
    assert stems >= len(q) * (t.max_stem_level + 1)
    # threads do not change results
    it2, d2, c2 = t.nearest_n(q, 16, threads=4)
    assert_same_bits(it, it2)
    assert_same_bits(d, d2)


def test_empty_tree(oracle_mod):
    O = oracle_mod
    t = O.Tree.build(np.zeros((0, 3)), 32)
    q = np.array([[0.1, 0.2, 0.3]])
    it, d = t.nearest_one(q)
    assert it[0] == 0 and math.isinf(d[0])
    it, d, c = t.nearest_n(q, 4)
    assert c[0] == 0
    off, wi, wd = t.within(q, 1.0)
    assert off.tolist() == [0, 0]
