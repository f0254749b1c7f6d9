"""GPU parity for the other distance metrics (Manhattan, Chebyshev, Minkowski<P>): every
family through the C ABI, Cartesian and periodic, against the oracle bit for bit, plus the
traversal counters and the argument checks."""
import math

import numpy as np
import pytest

from engines import GpuEngine, assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


def _metrics(ka):
    return {"manhattan": ka.KD_METRIC_MANHATTAN, "chebyshev": ka.KD_METRIC_CHEBYSHEV,
            "minkowski3": ka.KD_METRIC_MINKOWSKI(3), "minkowski5": ka.KD_METRIC_MINKOWSKI(5)}


CASES = [(20000, 3, np.float64, 32), (20000, 3, np.float32, 16), (6000, 2, np.float64, 8),
         (6000, 4, np.float32, 32), (20000, 4, np.uint16, 32)]


def _data(n, K, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.uint16:
        pts = rng.integers(0, 1000, (n, K)).astype(np.uint16)
        q = rng.integers(0, 1000, (1500, K)).astype(np.uint16)
        q[:100] = pts[:100]
        return pts, q
    pts = rng.random((n, K))
    pts = np.concatenate([pts, pts[rng.integers(0, n, n // 20)]])
    q = rng.random((1500, K))
    q[:100] = pts[:100]
    return pts.astype(dtype), q.astype(dtype)


@pytest.mark.parametrize("metric", ["manhattan", "chebyshev", "minkowski3", "minkowski5"])
@pytest.mark.parametrize("n,K,dtype,B", CASES, ids=[f"{c[0]}-k{c[1]}-{np.dtype(c[2]).name}-b{c[3]}" for c in CASES])
def test_metric_families_match_oracle(ka, oracle_mod, cuda, metric, n, K, dtype, B):
    f = _metrics(ka)[metric]
    pts, q = _data(n, K, dtype, seed=n + K + B)
    ot = oracle_mod.Tree.build(pts, B).set_metric(f)
    t = ka.ImmutableKdTree.new_from_slice(pts, B)
    # a radius that holds a few tens of entries per query
    bi, bd, _ = ot.nearest_n(q[:64], 24)
    r = float(np.median(bd[:, -1]))
    for extra in (0, ka.KD_MORTON):
        g = GpuEngine(ka, flags=f | extra)
        _announce(metric, n, K, np.dtype(dtype).name, B, "morton" if extra else "input order")
        gi, gd = g.nearest_one(t, q)
        oi, od = ot.nearest_one(q)
        assert_same_bits(gi, oi, "nearest_one items")
        assert_same_bits(gd, od, "nearest_one dists")
        for k, radius, excl in ((1, math.inf, False), (8, math.inf, False), (16, r, False), (5, r, True)):
            gi, gd, gc = g.nearest_n(t, q, k, radius, excl)
            oi, od, oc = ot.nearest_n(q, k, radius, excl)
            assert_same_bits(gc, oc, f"k={k} counts")
            assert_same_bits(gi, oi, f"k={k} items")
            assert_same_bits(gd, od, f"k={k} dists")
        for excl in (False, True):
            go, gi, gd = g.within(t, q, r, excl, True)
            oo, oi, od = ot.within(q, r, excl, True)
            assert_same_bits(go, oo, "within offsets")
            assert_same_bits(gi, oi, "within items")
            assert_same_bits(gd, od, "within dists")
            gi, gd, gc = g.best_n_within(t, q, r, 5, excl)
            oi, od, oc = ot.best_n_within(q, r, 5, excl)
            assert_same_bits(gc, oc, "best counts")
            assert_same_bits(gi, oi, "best items")
            assert_same_bits(gd, od, "best dists")
        if dtype != np.uint16:
            gi, gd = g.approx_nearest_one(t, q)
            oi, od = ot.approx_nearest_one(q)
            assert_same_bits(gi, oi, "approx items")
            assert_same_bits(gd, od, "approx dists")


@pytest.mark.parametrize("metric", ["manhattan", "chebyshev", "minkowski3"])
def test_metric_periodic_matches_oracle(ka, oracle_mod, cuda, metric):
    f = _metrics(ka)[metric]
    rng = np.random.default_rng(17)
    box = np.array([1.0, 2.0, 0.5])
    pts = rng.random((15000, 3)) * box
    q = rng.random((1500, 3)) * box
    q[:100] = rng.integers(0, 2, (100, 3)) * box * rng.choice([1e-5, 0.99999], (100, 3))
    ot = oracle_mod.Tree.build(pts, 32).set_metric(f)
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    g = GpuEngine(ka, flags=f | ka.KD_MORTON)
    _announce("periodic", metric)
    gi, gd = g.periodic_nearest_one(t, q, box)
    oi, od = ot.periodic_nearest_one(q, box)
    assert_same_bits(gi, oi, "periodic nearest_one items")
    assert_same_bits(gd, od, "periodic nearest_one dists")
    _, bd, _ = ot.periodic_nearest_n(q[:64], box, 20)
    r = float(np.median(bd[:, -1]))
    for k, radius in ((4, math.inf), (12, r)):
        gi, gd, gc = g.periodic_nearest_n(t, q, box, k, radius)
        oi, od, oc = ot.periodic_nearest_n(q, box, k, radius)
        assert_same_bits(gc, oc, "periodic counts")
        assert_same_bits(gi, oi, "periodic items")
        assert_same_bits(gd, od, "periodic dists")
    go, gi, gd = g.periodic_within(t, q, box, r)
    oo, oi, od = ot.periodic_within(q, box, r)
    assert_same_bits(go, oo, "periodic within offsets")
    assert_same_bits(gi, oi, "periodic within items")
    assert_same_bits(gd, od, "periodic within dists")


def test_metric_traversal_counters_match_oracle(ka, oracle_mod, cuda):
    rng = np.random.default_rng(2)
    pts = rng.random((40000, 3))
    q = rng.random((3000, 3))
    for metric in ("manhattan", "chebyshev"):
        f = _metrics(ka)[metric]
        ot = oracle_mod.Tree.build(pts, 32).set_metric(f)
        t = ka.ImmutableKdTree.new_from_slice(pts, 32)
        _announce("counters", metric)
        *_, ost = ot.nearest_n(q, 8, stats=True)
        t.nearest_n_raw(q, 8, math.inf, f)
        st = t.query_stats()
        assert (st["leaves"], st["points"]) == (int(ost[0]), int(ost[1])), metric


def test_metric_argument_checks(ka, cuda):
    pts = np.random.default_rng(3).random((3000, 3))
    t = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    q = pts[:16]
    for bad in (ka.KD_METRIC_MINKOWSKI(1), ka.KD_METRIC_MINKOWSKI(2), 7 << 8, 1 << 12):
        with pytest.raises(ka.KiddoAmdError):
            t.nearest_n_raw(q, 4, math.inf, bad)
    with pytest.raises(ka.KiddoAmdError):
        t.nearest_n_raw(q, 4, math.inf, ka.KD_METRIC_MANHATTAN | ka.KD_SCREEN_F32)
    r = t.query(q).nearest_n(4, metric="chebyshev").execute()
    r2 = t.nearest_n_raw(q, 4, math.inf, ka.KD_METRIC_CHEBYSHEV)
    assert cuda.equal(r.items, r2.items)
