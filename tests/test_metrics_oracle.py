"""Pins the oracle's other distance metrics (src/dist/{manhattan,chebyshev,minkowski}):
  * the reference's metric tests (src/dist/mod.rs: distances 25, 7, 4, 91.125 and the
    Manhattan / Chebyshev rect-distance updates 9, 2, 6), from reference_kats.json;
  * query distances against an independent numpy evaluation in Kiddo's order;
  * every family (nearest_one, nearest_n, nearest_n().within, within, best_n_within)
    against brute force, for f32, f64 and u16 trees, K = 2, 3, 4, which pins the
    traversal's rect updates (a wrong update prunes a subtree that holds an answer);
    approx has no backtracking and is pinned separately (tests/test_oracle_approx.py);
  * periodic nearest_n / within under each metric against a minimum-image brute force.
"""
import numpy as np
import pytest

from engines import load_kats

KATS = load_kats()
MANHATTAN, CHEBYSHEV = 1 << 8, 2 << 8


def MINKOWSKI(p):
    return (3 << 8) | (p << 12)


METRICS = {"manhattan": MANHATTAN, "chebyshev": CHEBYSHEV, "minkowski3": MINKOWSKI(3),
           "minkowski4": MINKOWSKI(4)}


def _flags(m):
    if isinstance(m, list):
        return MINKOWSKI(m[1])
    return {"sqeuclidean": 0, "manhattan": MANHATTAN, "chebyshev": CHEBYSHEV}[m]


@pytest.mark.parametrize("case", KATS["metric_cases"], ids=lambda c: c["name"])
def test_metric_kats(oracle_mod, case):
    f = _flags(case["metric"])
    if "dist" in case:
        assert oracle_mod.metric_dist(f, case["a"], case["b"]) == case["dist"]
    else:
        assert oracle_mod.metric_rect(f, case["rd"], case["off"], case["dim"], case["new_off"]) == case["rect"]


def numpy_metric_dist(points, q, flags, out_dtype):
    """Kiddo's metric distance evaluated independently: per axis the difference, then
    |d| (Manhattan, Chebyshev) or |d|^P by binary exponentiation (Minkowski), combined
    from 0 in axis order by + (max for Chebyshev), every op rounded in out_dtype."""
    kind, p = (flags >> 8) & 0xF, (flags >> 12) & 0xFF
    P = points.astype(out_dtype)
    Q = q.astype(out_dtype)
    acc = np.zeros(np.broadcast_shapes(P.shape[:-1], Q.shape[:-1]), out_dtype)
    for d in range(P.shape[-1]):
        x = np.abs((P[..., d] - Q[..., d]).astype(out_dtype))
        if kind == 0:
            t = (x * x).astype(out_dtype)
        elif kind in (1, 2):
            t = x
        else:
            r = np.ones_like(x)
            a = x.copy()
            b = p
            while True:
                if b & 1:
                    r = (r * a).astype(out_dtype)
                b >>= 1
                if not b:
                    break
                a = (a * a).astype(out_dtype)
            t = r
        acc = np.maximum(acc, t) if kind == 2 else (acc + t).astype(out_dtype)
    return acc


def _data(kind, n, K, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.uint16:
        pts = rng.integers(0, 200, (n, K)).astype(np.uint16)
        q = rng.integers(0, 200, (200, K)).astype(np.uint16)
        return pts, q
    pts = rng.random((n, K))
    if kind == "dups":
        pts = pts[rng.integers(0, n // 3, n)]
    q = rng.random((200, K))
    q[:20] = pts[:20]
    return pts.astype(dtype), q.astype(dtype)


@pytest.mark.parametrize("metric", list(METRICS))
@pytest.mark.parametrize("kind,K,dtype,B", [("uniform", 3, np.float64, 32), ("dups", 2, np.float32, 8),
                                            ("uniform", 4, np.float64, 16), ("uniform", 3, np.uint16, 32)])
def test_metric_families_vs_brute_force(oracle_mod, metric, kind, K, dtype, B):
    f = METRICS[metric]
    pts, q = _data(kind, 3000, K, dtype, seed=K + B)
    t = oracle_mod.Tree.build(pts, B).set_metric(f)
    odt = np.float64 if dtype == np.uint16 else dtype
    d = numpy_metric_dist(pts[None, :, :], q[:, None, :], f, odt)
    # a radius around the 10th-smallest distance of a typical query
    r = float(np.sort(d, axis=1)[:, 10].mean())
    for k, rad, excl in ((1, np.inf, False), (9, np.inf, False), (6, r, False), (6, r, True)):
        it, dd, c = t.nearest_n(q, k, rad, excl)
        bi, bd, bc = t.brute_nearest_n(q, k, rad, excl)
        np.testing.assert_array_equal(c, bc)
        np.testing.assert_array_equal(it, bi)
        assert dd.tobytes() == bd.tobytes()
        for i in range(len(q)):  # the distances are Kiddo's, evaluated independently
            assert dd[i, : c[i]].tobytes() == d[i, it[i, : c[i]]].astype(odt).tobytes()
    oi, od = t.nearest_one(q)
    bi, bd, _ = t.brute_nearest_n(q, 1)
    np.testing.assert_array_equal(oi, bi[:, 0])
    for excl in (False, True):
        off, it, dd = t.within(q, r, excl)
        boff, bit, bdd = t.brute_within(q, r, excl)
        np.testing.assert_array_equal(off, boff)
        np.testing.assert_array_equal(it, bit)
        bi, bd, bc = t.best_n_within(q, r, 4, excl)
        ci, cd, cc = t.brute_best_n_within(q, r, 4, excl)
        np.testing.assert_array_equal(bi, ci)
        np.testing.assert_array_equal(bc, cc)


@pytest.mark.parametrize("metric", list(METRICS))
def test_metric_periodic_vs_brute_force(oracle_mod, metric):
    f = METRICS[metric]
    rng = np.random.default_rng(5)
    box = np.array([1.0, 2.0, 0.5])
    pts = rng.random((2500, 3)) * box
    q = rng.random((200, 3)) * box
    q[:30] = rng.integers(0, 2, (30, 3)) * box * rng.choice([1e-4, 0.9999], (30, 3))
    t = oracle_mod.Tree.build(pts, 16).set_metric(f)
    kind, p = (f >> 8) & 0xF, (f >> 12) & 0xFF
    delta = np.abs(q[:, None, :] - pts[None, :, :])
    mi = np.minimum(box - delta, delta)  # wrapped < delta ? wrapped : delta
    mi = np.where(box - delta < delta, box - delta, delta)
    d = numpy_metric_dist(mi, np.zeros(3), f, np.float64)
    for k, rad in ((1, np.inf), (8, np.inf), (5, float(np.sort(d, axis=1)[:, 6].mean()))):
        it, dd, c = t.periodic_nearest_n(q, box, k, rad)
        for i in range(len(q)):
            ok = np.nonzero(d[i] <= rad)[0]
            want = ok[np.lexsort((ok, d[i, ok]))][:k]
            assert it[i, : c[i]].tolist() == want.tolist(), (metric, k, i)
            assert dd[i, : c[i]].tobytes() == d[i, want].tobytes()
    r = float(np.sort(d, axis=1)[:, 12].mean())
    off, it, dd = t.periodic_within(q, box, r)
    for i in range(len(q)):
        ok = np.nonzero(d[i] <= r)[0]
        want = ok[np.lexsort((ok, d[i, ok]))]
        assert it[int(off[i]): int(off[i + 1])].tolist() == want.tolist()


def test_metric_argument_checks(oracle_mod):
    t = oracle_mod.Tree.build(np.random.default_rng(0).random((50, 3)), 8)
    for bad in (MINKOWSKI(1), MINKOWSKI(2), 5 << 8):
        with pytest.raises(RuntimeError):
            t.set_metric(bad)
