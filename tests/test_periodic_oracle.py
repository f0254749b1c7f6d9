"""Pins the oracle's periodic-boundary queries (query/periodic.rs) before they check the GPU:
  * the reference's periodic known answers (query/builder.rs tests, in reference_kats.json),
  * brute force under the minimum-image distance, evaluated independently in numpy in the
    order Kiddo's periodic leaf scan uses, for nearest_n, nearest_n().within and within,
  * brute force for periodic nearest_one: the minimum over all 3^K shifted query images of
    the Cartesian scalar-order distance (Kiddo returns the image search's distance),
  * argument checks (box lengths must be > 0, float trees only).
"""
import itertools

import numpy as np
import pytest

from engines import min_image_sqdist, scalar_order_sqdist


def _inputs(n, K, dtype, box, seed, dup_frac=0.1, faces=True):
    rng = np.random.default_rng(seed)
    pts = (rng.random((n, K)) * box).astype(dtype)
    if dup_frac:
        nd = int(n * dup_frac)
        pts = np.concatenate([pts, pts[rng.integers(0, n, nd)]])
    q = (rng.random((300, K)) * box).astype(dtype)
    if faces:  # queries on and next to the faces and corners of the box
        q[:40] = (rng.integers(0, 2, (40, K)) * box * rng.choice([1e-4, 0.9999], (40, K))).astype(dtype)
        q[40:60] = pts[rng.integers(0, len(pts), 20)]
        q[60:70, 0] = 0
    return pts, q


def _canon_topk(d, k, radius, excl):
    out = []
    for row in d:
        ok = row < radius if excl else row <= radius
        idx = np.nonzero(ok)[0]
        order = np.lexsort((idx, row[idx]))[:k]
        out.append(idx[order])
    return out


BOXES = {2: np.array([1.0, 2.5]), 3: np.array([1.0, 2.5, 0.75]), 4: np.array([1.0, 1.0, 3.0, 0.5])}


@pytest.mark.parametrize("K,dtype,B", [(2, np.float64, 4), (3, np.float64, 32), (3, np.float32, 8),
                                       (4, np.float64, 16), (4, np.float32, 32)])
def test_periodic_nearest_n_and_within_vs_brute_force(oracle_mod, K, dtype, B):
    box = BOXES[K].astype(dtype)
    pts, q = _inputs(1500, K, dtype, box, seed=K * 10 + B)
    t = oracle_mod.Tree.build(pts, B)
    d = min_image_sqdist(pts[None, :, :], q[:, None, :], box, dtype)  # [m, n]
    for k, radius, excl in ((1, np.inf, False), (8, np.inf, False), (16, 0.01, False), (5, 0.004, True)):
        it, dd, cnt = t.periodic_nearest_n(q, box, k, radius, excl)
        want = _canon_topk(d, k, dtype(radius), excl)
        for i in range(len(q)):
            assert cnt[i] == len(want[i]), (k, radius, i)
            assert it[i, : cnt[i]].tolist() == want[i].tolist(), (k, radius, i)
            assert dd[i, : cnt[i]].tobytes() == d[i, want[i]].tobytes()
    for radius, excl in ((0.003, False), (0.02, True)):
        off, it, dd = t.periodic_within(q, box, radius, excl)
        want = _canon_topk(d, len(pts), dtype(radius), excl)
        for i in range(len(q)):
            s, e = int(off[i]), int(off[i + 1])
            assert it[s:e].tolist() == want[i].tolist(), (radius, i)
            assert dd[s:e].tobytes() == d[i, want[i]].tobytes()


@pytest.mark.parametrize("K,dtype", [(2, np.float64), (3, np.float64), (3, np.float32), (4, np.float32)])
def test_periodic_nearest_one_vs_all_images(oracle_mod, K, dtype):
    box = BOXES[K].astype(dtype)
    pts, q = _inputs(1200, K, dtype, box, seed=7 + K)
    t = oracle_mod.Tree.build(pts, 16)
    it, dd = t.periodic_nearest_one(q, box)
    best_d = np.full(len(q), np.inf, dtype)
    best_i = np.full(len(q), 0xFFFFFFFF, np.uint64)
    for shifts in itertools.product((-1, 0, 1), repeat=K):
        wq = q.copy()
        for a, s in enumerate(shifts):
            if s:
                wq[:, a] = (wq[:, a] + dtype(s) * box[a]).astype(dtype)
        d = scalar_order_sqdist(pts[None, :, :], wq[:, None, :], dtype)
        j = np.lexsort((np.arange(len(pts))[None, :].repeat(len(q), 0), d), axis=1)[:, 0]
        dj = d[np.arange(len(q)), j]
        better = (dj < best_d) | ((dj == best_d) & (j < best_i))
        best_d = np.where(better, dj, best_d)
        best_i = np.where(better, j, best_i)
    assert it.tolist() == best_i.tolist()
    assert dd.tobytes() == best_d.astype(dtype).tobytes()


def test_periodic_faithful_distances_match_canonical(oracle_mod):
    O = oracle_mod
    box = BOXES[3]
    pts, q = _inputs(3000, 3, np.float64, box, seed=3, dup_frac=0.3)
    t = O.Tree.build(pts, 32)
    _, dc, cc = t.periodic_nearest_n(q, box, 8, mode=O.CANONICAL)
    _, df, cf = t.periodic_nearest_n(q, box, 8, mode=O.FAITHFUL)
    assert_equal = np.testing.assert_array_equal
    assert_equal(cc, cf)
    assert_equal(dc.view(np.uint64), df.view(np.uint64))
    _, d1 = t.periodic_nearest_one(q, box, mode=O.CANONICAL)
    _, d2 = t.periodic_nearest_one(q, box, mode=O.FAITHFUL)
    assert_equal(d1.view(np.uint64), d2.view(np.uint64))


def test_periodic_argument_checks(oracle_mod):
    O = oracle_mod
    pts = np.random.default_rng(0).random((100, 3))
    t = O.Tree.build(pts, 8)
    q = pts[:3]
    for bad in ([1.0, 0.0, 1.0], [1.0, -1.0, 1.0], [np.nan, 1.0, 1.0]):
        with pytest.raises(RuntimeError):
            t.periodic_nearest_n(q, np.array(bad), 4)
        with pytest.raises(RuntimeError):
            t.periodic_nearest_one(q, np.array(bad))
    tu = O.Tree.build(np.random.default_rng(1).integers(0, 1000, (50, 3), dtype=np.uint16), 8)
    with pytest.raises(RuntimeError):
        tu.periodic_nearest_n(np.zeros((1, 3), np.uint16), np.array([10, 10, 10], np.uint16), 2)


def test_periodic_empty_tree(oracle_mod):
    t = oracle_mod.Tree.build(np.zeros((0, 3)), 32)
    box = np.ones(3)
    it, d = t.periodic_nearest_one(np.full((2, 3), 0.5), box)
    assert it.tolist() == [0, 0] and np.isinf(d).all()
    it, d, c = t.periodic_nearest_n(np.full((2, 3), 0.5), box, 3)
    assert c.tolist() == [0, 0]
    off, it, d = t.periodic_within(np.full((2, 3), 0.5), box, 1.0)
    assert off.tolist() == [0, 0, 0]
