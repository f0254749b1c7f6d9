"""Launch shapes for small batches, against the oracle bit for bit (items, order, distance
bits, counts):
  * best_n_within one thread per query and 4/8/16/32 lanes per query, and the automatic
    choice, on the parity cases (n from 1 to 20, inclusive and exclusive radius);
  * with repeated items, every shape keeps the same entries as one thread per query;
  * nearest_one's and nearest_n's automatic forms, which take the group kernels on small
    batches (below 2^16 queries, or 2^19 on trees of at least 2^22 entries), plain and
    FP32-screened, and nearest_one where every FP32 distance overflows to +inf.
"""
import math

import numpy as np
import pytest

from engines import GpuEngine, OracleEngine, assert_same_bits
from test_gpu_parity import CASES, IDS, _points, _queries, _radius

pytestmark = pytest.mark.gpu

NAMES = ("items", "dists", "counts")


def _announce(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("kind,n,K,dtype,B", CASES, ids=IDS)
def test_best_n_within_shapes_match_oracle(ka, oracle_mod, cuda, kind, n, K, dtype, B):
    _announce("best_n_within shapes", kind, n, K, np.dtype(dtype).name, B)
    pts = _points(kind, n, K, dtype, seed=n + K + 7)
    q = _queries(pts, 3000, seed=n + 1)
    r = _radius(pts)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, B)
    for flags in (0, ka.KD_MORTON):
        G = GpuEngine(ka, flags=flags)
        gt = G.build(pts, B)
        for nn in (1, 3, 5, 16, 20):
            for excl in (False, True):
                ref = O.best_n_within(ot, q, r, nn, excl)
                for shape in (1, 4, 8, 16, 32, 0):
                    gt.tune_best_n_within(shape)
                    res = G.best_n_within(gt, q, r, nn, excl)
                    for a, b, nm in zip(res, ref, NAMES):
                        assert_same_bits(a, b, f"best_n_within({nn}) {nm} excl={excl} shape={shape} flags={flags}")
        gt.tune_best_n_within(0)


def test_best_n_within_repeated_items_same_across_shapes(ka, cuda):
    """Items repeat (97 distinct values): every shape keeps the same entries, in the same
    order, as one thread per query (the first entry of an item visited stays first)."""
    _announce("best_n_within repeated items")
    rng = np.random.default_rng(11)
    pts = rng.random((60_000, 3))
    items = (np.arange(len(pts)) % 97).astype(np.uint32)
    q = rng.random((4000, 3))
    G = GpuEngine(ka, flags=ka.KD_MORTON)
    gt = G.build(pts, 32, items)
    for nn in (1, 5, 16):
        gt.tune_best_n_within(1)
        ref = G.best_n_within(gt, q, 0.01, nn)
        assert int(ref[2].sum()) > 0
        for shape in (4, 8, 16, 32, 0):
            gt.tune_best_n_within(shape)
            res = G.best_n_within(gt, q, 0.01, nn)
            for a, b, nm in zip(res, ref, NAMES):
                assert_same_bits(a, b, f"repeated items best_n_within({nn}) {nm} shape={shape}")
    gt.tune_best_n_within(0)


def test_tune_best_n_within_rejects_bad_shape(ka, cuda):
    t = ka.ImmutableKdTree.new_from_slice(np.random.default_rng(1).random((100, 3)), 32)
    for bad in (2, 3, 64, -1):
        with pytest.raises(Exception):
            t.tune_best_n_within(bad)


SMALL = [
    ("uniform", 100_000, 3, np.float64, 32),
    ("dups", 60_000, 3, np.float64, 32),
    ("grid", 30_000, 2, np.float32, 8),
    ("uniform", 65_536, 4, np.uint16, 32),
    ("all_same", 800, 3, np.float64, 32),
]


@pytest.mark.parametrize("kind,n,K,dtype,B", SMALL, ids=[f"{c[0]}-{c[1]}-K{c[2]}-{np.dtype(c[3]).name}" for c in SMALL])
def test_nearest_n_automatic_small_batch(ka, oracle_mod, cuda, kind, n, K, dtype, B):
    _announce("nearest_n automatic small batch", kind, n, K, np.dtype(dtype).name, B)
    pts = _points(kind, n, K, dtype, seed=n + K + 3)
    r = _radius(pts)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, B)
    engines = [GpuEngine(ka, flags=0), GpuEngine(ka, flags=ka.KD_MORTON)]
    if np.dtype(dtype) == np.float64:
        engines.append(GpuEngine(ka, flags=ka.KD_MORTON | ka.KD_SCREEN_F32, shadow=True))
    for G in engines:
        gt = G.build(pts, B)
        for m in (1, 100, 3000):
            q = _queries(pts, m, seed=m + n)
            ref1 = O.nearest_one(ot, q)
            for variant in (0, 1):  # automatic (the group form here) and one thread per query
                gt.tune(variant, 0, -1)
                res1 = G.nearest_one(gt, q)
                for a, b, nm in zip(res1, ref1, ("items", "dists")):
                    assert_same_bits(a, b, f"nearest_one {nm} m={m} variant={variant} flags={G.flags}")
            gt.tune(0, 0, -1)
            for k in (1, 4, 5, 8, 9, 16):
                for radius, excl in ((math.inf, False), (r, False), (r, True)):
                    ref = O.nearest_n(ot, q, k, radius, excl)
                    res = G.nearest_n(gt, q, k, radius, excl)
                    for a, b, nm in zip(res, ref, NAMES):
                        assert_same_bits(a, b, f"nearest_n({k}) {nm} m={m} r={radius} excl={excl} flags={G.flags}")


def test_nearest_n_automatic_large_tree(ka, oracle_mod, cuda):
    """A tree of at least 2^22 entries and a batch between 2^16 and 2^19 queries: the
    automatic form is the group form there; one thread per query and the oracle agree."""
    _announce("nearest_n automatic large tree")
    rng = np.random.default_rng(21)
    pts = rng.random(((1 << 22) + 1000, 3))
    q = rng.random((70_000, 3))
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, 32)
    G = GpuEngine(ka, flags=ka.KD_MORTON)
    gt = G.build(pts, 32)
    for k in (5, 16):
        ref = ot.nearest_n(q, k, math.inf, False, True, threads=16)
        gt.tune(0, 0, -1)
        res = G.nearest_n(gt, q, k)
        gt.tune(1, 0, -1)
        res_t = G.nearest_n(gt, q, k)
        gt.tune(0, 0, -1)
        for a, b, c, nm in zip(res, ref, res_t, NAMES):
            assert_same_bits(a, b, f"nearest_n({k}) {nm} automatic")
            assert_same_bits(c, b, f"nearest_n({k}) {nm} thread per query")


def test_nearest_one_infinite_distances(ka, oracle_mod, cuda):
    """f32 coordinates near 1e20 square to +inf: every distance is +inf, and the group and
    thread forms return what the oracle returns."""
    _announce("nearest_one infinite distances")
    rng = np.random.default_rng(5)
    pts = (rng.random((5000, 3)) * 1e20 + 1e20).astype(np.float32)
    q = (-rng.random((700, 3)) * 1e20 - 1e20).astype(np.float32)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, 32)
    ref = O.nearest_one(ot, q)
    assert np.all(np.isinf(ref[1]))
    G = GpuEngine(ka, flags=0)
    gt = G.build(pts, 32)
    for variant in (0, 1):
        gt.tune(variant, 0, -1)
        res = G.nearest_one(gt, q)
        for a, b, nm in zip(res, ref, ("items", "dists")):
            assert_same_bits(a, b, f"nearest_one inf {nm} variant={variant}")
