"""GPU parity for queries under periodic boundary conditions (query/periodic.rs): the HIP
kernels through the C ABI against the oracle on the same seeded inputs, bit for bit, plus
the traversal counters (leaves, entries, stems) of the periodic backtracking."""
import math

import numpy as np
import pytest

from engines import GpuEngine, assert_same_bits

pytestmark = pytest.mark.gpu

BOXES = {2: [1.0, 2.5], 3: [1.0, 2.5, 0.75], 4: [1.0, 1.0, 3.0, 0.5]}


def _announce(*what):
    print("case:", *what, flush=True)


def _inputs(n, K, dtype, seed):
    rng = np.random.default_rng(seed)
    box = np.array(BOXES[K], dtype)
    pts = (rng.random((n, K)) * box).astype(dtype)
    pts = np.concatenate([pts, pts[rng.integers(0, n, n // 10)]])  # exact duplicates
    q = (rng.random((2000, K)) * box).astype(dtype)
    q[:200] = (rng.integers(0, 2, (200, K)) * box * rng.choice([1e-5, 0.99999], (200, K))).astype(dtype)
    q[200:300] = pts[rng.integers(0, len(pts), 100)]  # on entries
    q[300:350, 0] = 0
    q[350:400] = (q[350:400] + box * rng.choice([-1.3, 1.2], (50, K))).astype(dtype)  # outside the box
    return pts, q, box


CASES = [(3000, 2, np.float64, 8), (20000, 3, np.float64, 32), (20000, 3, np.float32, 32),
         (8000, 4, np.float64, 16), (8000, 4, np.float32, 4)]


@pytest.mark.parametrize("n,K,dtype,B", CASES, ids=[f"{c[0]}-k{c[1]}-{np.dtype(c[2]).name}-b{c[3]}" for c in CASES])
def test_periodic_families_match_oracle(ka, oracle_mod, cuda, n, K, dtype, B):
    pts, q, box = _inputs(n, K, dtype, seed=n + K + B)
    ot = oracle_mod.Tree.build(pts, B)
    t = ka.ImmutableKdTree.new_from_slice(pts, B)
    for flags in (0, ka.KD_MORTON):
        g = GpuEngine(ka, flags=flags)
        _announce("periodic nearest_one", n, K, np.dtype(dtype).name, B, flags)
        gi, gd = g.periodic_nearest_one(t, q, box)
        oi, od = ot.periodic_nearest_one(q, box)
        assert_same_bits(gi, oi, "nearest_one items")
        assert_same_bits(gd, od, "nearest_one dists")
        for k, radius, excl in ((1, math.inf, False), (8, math.inf, False), (16, 0.002, False),
                                (5, 0.001, True), (40, math.inf, False)):
            _announce("periodic nearest_n", k, radius, excl, flags)
            gi, gd, gc = g.periodic_nearest_n(t, q, box, k, radius, excl)
            oi, od, oc = ot.periodic_nearest_n(q, box, k, radius, excl)
            assert_same_bits(gc, oc, f"k={k} counts")
            assert_same_bits(gi, oi, f"k={k} items")
            assert_same_bits(gd, od, f"k={k} dists")
        for radius, excl in ((0.001, False), (0.004, True)):
            _announce("periodic within", radius, excl, flags)
            go, gi, gd = g.periodic_within(t, q, box, radius, excl, True)
            oo, oi, od = ot.periodic_within(q, box, radius, excl, True)
            assert_same_bits(go, oo, "within offsets")
            assert_same_bits(gi, oi, "within items")
            assert_same_bits(gd, od, "within dists")
            # unsorted: the same multiset per query
            go2, gi2, gd2 = g.periodic_within(t, q, box, radius, excl, False)
            assert_same_bits(go2, oo, "unsorted offsets")
            for i in range(0, len(q), 97):
                s, e = int(oo[i]), int(oo[i + 1])
                a = sorted(zip(gd2[s:e].tolist(), gi2[s:e].tolist()))
                b = sorted(zip(od[s:e].tolist(), oi[s:e].tolist()))
                assert a == b, i


def test_periodic_traversal_counters_match_oracle(ka, oracle_mod, cuda):
    pts, q, box = _inputs(30000, 3, np.float64, seed=11)
    ot = oracle_mod.Tree.build(pts, 32)
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    _announce("periodic counters nearest_n")
    *_, ost = ot.periodic_nearest_n(q, box, 8, stats=True)
    t.periodic_nearest_n_raw(q, box, 8)
    st = t.query_stats()
    assert (st["leaves"], st["points"], st["stems"]) == tuple(int(x) for x in ost)
    _announce("periodic counters within")
    t2 = ka.ImmutableKdTree.new_from_slice(pts, 32)
    go, oi, od, ost = ot.periodic_within(q, box, 0.002, stats=True)
    lib = ka.lib()
    qq = t2._queries(q)
    offs = cuda.empty(len(q) + 1, dtype=cuda.int64, device="cuda")
    b = np.ascontiguousarray(box)
    ka._check(lib.kd_periodic_within_count(t2.handle, qq.data_ptr(), len(q), ka._np_ptr(b), 0.002, 0,
                                           offs.data_ptr(), None), "count")
    cuda.cuda.synchronize()
    st = t2.query_stats()
    assert (st["leaves"], st["points"], st["stems"]) == tuple(int(x) for x in ost)


def test_periodic_builder_surface_and_errors(ka, cuda):
    pts = np.random.default_rng(3).random((5000, 3))
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    q = pts[:64] + 0.25
    box = [1.0, 1.0, 1.0]
    r = t.query(q).periodic_boundary_condition(box).nearest_n(4).execute()
    r2 = t.periodic_nearest_n_raw(q, box, 4)
    assert cuda.equal(r.items, r2.items)
    w = t.query(q).periodic_boundary_condition(box).within(0.001).unsorted().execute()
    w2 = t.periodic_within_raw(q, box, 0.001, ka.KD_UNSORTED)
    assert cuda.equal(w.offsets, w2.offsets)
    o = t.query(q).periodic_boundary_condition(box).nearest_one().execute()
    assert cuda.equal(o.item, t.periodic_nearest_one_raw(q, box).item)
    with pytest.raises(ValueError):
        t.query(q).periodic_boundary_condition(box).best_n_within(0.1, 3)
    for bad in ([1.0, 0.0, 1.0], [1.0, -2.0, 1.0], [math.nan, 1.0, 1.0]):
        with pytest.raises(ka.KiddoAmdError):
            t.periodic_nearest_n_raw(q, bad, 4)
    tu = ka.ImmutableKdTree.new_from_slice(np.random.default_rng(4).integers(0, 999, (500, 3), dtype=np.uint16), 8)
    with pytest.raises(ka.KiddoAmdError):
        tu.periodic_nearest_one_raw(np.zeros((2, 3), np.uint16), [10, 10, 10])


def test_periodic_empty_tree(ka, cuda):
    t = ka.ImmutableKdTree.new_from_slice(np.zeros((0, 3)), 32)
    q = np.full((3, 3), 0.5)
    r = t.periodic_nearest_one_raw(q, [1.0, 1.0, 1.0])
    assert r.item.cpu().tolist() == [0, 0, 0]
    assert np.isinf(r.distance.cpu().numpy()).all()
    r = t.periodic_nearest_n_raw(q, [1.0, 1.0, 1.0], 3)
    assert r.counts.cpu().tolist() == [0, 0, 0]
    w = t.periodic_within_raw(q, [1.0, 1.0, 1.0], 0.5)
    assert w.offsets.cpu().tolist() == [0, 0, 0, 0]
