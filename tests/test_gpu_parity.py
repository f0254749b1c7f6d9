"""GPU parity: the HIP kernels (through the C ABI) against the oracle on the same seeded
inputs, bit for bit: items, result order (distance, then lower item) and distance bits.
"""
import math

import numpy as np
import pytest

from engines import (GpuEngine, OracleEngine, assert_same_bits, load_kats, run_kat_case,
                     scalar_order_sqdist)

pytestmark = pytest.mark.gpu

KATS = load_kats()


def _announce(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("case", KATS["cases"], ids=lambda c: c["name"])
def test_reference_kats_on_gpu(ka, cuda, case):
    _announce("kat", case["name"])
    run_kat_case(GpuEngine(ka), case)
    run_kat_case(GpuEngine(ka, flags=ka.KD_MORTON), case)


def _points(kind, n, K, dtype, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":
        p = rng.random((n, K))
    elif kind == "dups":
        base = rng.random((max(n // 4, 1), K))
        p = base[rng.integers(0, len(base), n)]
    elif kind == "grid":
        p = rng.integers(0, 7, (n, K)).astype(np.float64)
    elif kind == "mixture":
        from kiddo_amd.datasets import gaussian_mixture
        p = gaussian_mixture(n, 8, 0.02, K, seed=seed).astype(np.float64)
    elif kind == "all_same":
        p = np.full((n, K), 0.375)
    else:
        raise ValueError(kind)
    if np.dtype(dtype).kind == "u":
        if kind != "grid":
            p = p * 60000
        return p.astype(dtype)
    return p.astype(dtype)


def _queries(pts, m, seed):
    rng = np.random.default_rng(seed + 99)
    if pts.dtype.kind == "u":
        q = rng.integers(0, int(pts.max()) + 2, (m, pts.shape[1])).astype(pts.dtype)
    else:
        lo, hi = float(pts.min()), float(pts.max())
        span = max(hi - lo, 1.0)
        q = (lo - 0.05 * span + rng.random((m, pts.shape[1])) * 1.1 * span).astype(pts.dtype)
    if len(pts):
        q[: m // 4] = pts[rng.integers(0, len(pts), m // 4)]
    return q


def _radius(pts):
    if pts.dtype.kind == "u":
        return 2.0 if pts.max() < 100 else 2.5e7
    return 2.0 if pts.max() > 1.5 else 0.004


CASES = [
    ("uniform", 100_000, 3, np.float64, 32),
    ("uniform", 100_000, 3, np.float32, 32),
    ("dups", 60_000, 3, np.float64, 32),
    ("grid", 40_000, 3, np.float64, 32),
    ("grid", 30_000, 2, np.float32, 8),
    ("mixture", 80_000, 3, np.float32, 32),
    ("uniform", 50_000, 4, np.float64, 16),
    ("uniform", 65_536, 4, np.uint16, 32),
    ("grid", 20_000, 3, np.uint16, 32),
    ("uniform", 30_000, 3, np.uint32, 32),
    ("all_same", 800, 3, np.float64, 32),
    ("uniform", 20, 3, np.float64, 32),
    ("uniform", 1, 3, np.float32, 32),
]
IDS = [f"{c[0]}-{c[1]}-K{c[2]}-{np.dtype(c[3]).name}-B{c[4]}" for c in CASES]


@pytest.mark.parametrize("kind,n,K,dtype,B", CASES, ids=IDS)
def test_families_match_oracle(ka, oracle_mod, cuda, kind, n, K, dtype, B):
    _announce("families", kind, n, K, np.dtype(dtype).name, B)
    pts = _points(kind, n, K, dtype, seed=n + K)
    q = _queries(pts, 4000, seed=n)
    r = _radius(pts)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, B)
    for flags in (0, ka.KD_MORTON):
        G = GpuEngine(ka, flags=flags)
        gt = G.build(pts, B)
        tag = f"flags={flags}"
        gi, gd = G.nearest_one(gt, q)
        oi, od = O.nearest_one(ot, q)
        assert_same_bits(gi, oi, f"nearest_one items {tag}")
        assert_same_bits(gd, od, f"nearest_one dists {tag}")
        for k in (1, 5, 16):
            res_o = O.nearest_n(ot, q, k)
            # thread-per-query, group-per-query (several group sizes) and no LDS stems
            for variant, gsize, lds in ((1, 0, -1), (2, 0, -1), (2, 32, 3), (1, 0, 0)):
                gt.tune(variant, gsize, lds)
                res_g = G.nearest_n(gt, q, k)
                for a, b, nm in zip(res_g, res_o, ("items", "dists", "counts")):
                    assert_same_bits(a, b, f"nearest_n({k}) {nm} {tag} variant={variant}/{gsize}/{lds}")
            gt.tune(0, 0, -1)
        for excl in (False, True):
            res_g = G.nearest_n(gt, q, 8, r, excl)
            res_o = O.nearest_n(ot, q, 8, r, excl)
            for a, b, nm in zip(res_g, res_o, ("items", "dists", "counts")):
                assert_same_bits(a, b, f"nearest_n_within {nm} excl={excl} {tag}")
            go, gi, gd = G.within(gt, q, r, excl, sorted=True)
            oo, oi, od = O.within(ot, q, r, excl, sorted=True)
            assert_same_bits(go, oo, f"within offsets excl={excl} {tag}")
            assert_same_bits(gi, oi, f"within items excl={excl} {tag}")
            assert_same_bits(gd, od, f"within dists excl={excl} {tag}")
            uo, ui, ud = G.within(gt, q, r, excl, sorted=False)
            assert_same_bits(uo, oo, "within unsorted offsets")
            # unsorted: the same multiset per query
            key_g = np.lexsort((ui, ud, np.repeat(np.arange(len(q)), np.diff(uo.astype(np.int64)))))
            assert_same_bits(ui[key_g], oi, "within unsorted items (as multiset)")
            assert_same_bits(ud[key_g], od, "within unsorted dists (as multiset)")
            for nn in (1, 5):
                res_g = G.best_n_within(gt, q, r, nn, excl)
                res_o = O.best_n_within(ot, q, r, nn, excl)
                for a, b, nm in zip(res_g, res_o, ("items", "dists", "counts")):
                    assert_same_bits(a, b, f"best_n_within({nn}) {nm} excl={excl} {tag}")
        if pts.dtype.kind == "f":
            gi, gd = G.approx_nearest_one(gt, q)
            oi, od = O.approx_nearest_one(ot, q)
            assert_same_bits(gi, oi, f"approx items {tag}")
            assert_same_bits(gd, od, f"approx dists {tag}")


def test_large_k_and_k_above_n(ka, oracle_mod, cuda):
    _announce("large k")
    pts = _points("uniform", 3000, 3, np.float64, 1)
    q = _queries(pts, 500, 1)
    O, G = OracleEngine(oracle_mod), GpuEngine(ka)
    ot, gt = O.build(pts, 32), G.build(pts, 32)
    for k in (24, 32, 64):
        for a, b in zip(G.nearest_n(gt, q, k), O.nearest_n(ot, q, k)):
            assert_same_bits(a, b, f"k={k}")
    small = _points("uniform", 7, 3, np.float64, 2)
    ot, gt = O.build(small, 32), G.build(small, 32)
    for a, b in zip(G.nearest_n(gt, q, 16), O.nearest_n(ot, q, 16)):
        assert_same_bits(a, b, "k > n")


def test_empty_tree(ka, cuda):
    _announce("empty tree")
    G = GpuEngine(ka)
    t = G.build(np.zeros((0, 3)), 32)
    q = np.array([[0.1, 0.2, 0.3], [0.5, 0.5, 0.5]])
    it, d = G.nearest_one(t, q)
    assert it.tolist() == [0, 0] and np.all(np.isinf(d))
    it, d, c = G.nearest_n(t, q, 4)
    assert c.tolist() == [0, 0] and np.all(it == 0xFFFFFFFF)
    off, wi, wd = G.within(t, q, 1.0)
    assert off.tolist() == [0, 0, 0]


def test_screen_f32_equals_plain_f64(ka, oracle_mod, cuda):
    """FP32 screening + FP64 rescore returns exactly the plain FP64 kernel's results."""
    _announce("screen f32")
    from kiddo_amd.datasets import uniform_with_duplicates
    pts = uniform_with_duplicates(200_000, 0.05, seed=3)
    q = _queries(pts, 20000, 3)
    plain = ka.ImmutableKdTree.new_from_slice(pts, 32)
    shadow = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, 32)
    for k in (1, 8, 16):
        ref = O.nearest_n(ot, q, k)
        for variant, waves in ((1, 0), (2, 0), (1, 4), (1, 6)):
            shadow.tune(variant, 0, -1, waves)
            for flags in (ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON):
                r = shadow.nearest_n_raw(q, k, flags=flags)
                got = (r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy())
                for a, b in zip(got, ref):
                    assert_same_bits(a, b, f"screen k={k} flags={flags} variant={variant}/{waves}")
        screened, rescored = shadow.screen_stats()
        assert 0 < rescored < screened
        r = plain.nearest_n_raw(q, k)
        assert_same_bits(r.items.cpu().numpy(), ref[0])
    # radius-limited and exclusive forms
    for excl in (False, True):
        ref = O.nearest_n(ot, q, 16, 4e-6, excl)
        r = shadow.nearest_n_raw(q, 16, 4e-6, ka.KD_SCREEN_F32 | (ka.KD_EXCLUSIVE if excl else 0))
        assert_same_bits(r.items.cpu().numpy(), ref[0], "screen within items")
        assert_same_bits(r.distances.cpu().numpy(), ref[1], "screen within dists")


def test_distances_match_numpy_scalar_order(ka, cuda):
    _announce("numpy scalar order")
    rng = np.random.default_rng(4)
    for dtype in (np.float64, np.float32):
        pts = rng.random((50_000, 3)).astype(dtype)
        q = rng.random((2000, 3)).astype(dtype)
        t = ka.ImmutableKdTree.new_from_slice(pts, 32)
        r = t.query(q).nearest_n(16).execute()
        it = r.items.cpu().numpy()
        ref = scalar_order_sqdist(pts[it], q[:, None, :], dtype)
        assert_same_bits(r.distances.cpu().numpy(), ref, np.dtype(dtype).name)


def test_builder_surface(ka, cuda):
    _announce("builder surface")
    pts = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [0.5, 0.0]])
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    q = np.array([[0.0, 0.0]])
    r = t.query(q).nearest_n(8).within(1.0).exclusive_boundaries().execute()
    assert r.items.cpu().numpy()[0][: int(r.counts[0])].tolist() == [0, 3]
    r = t.query(q).within(1.0).unsorted().execute()
    assert sorted(r.items.cpu().numpy().tolist()) == [0, 1, 3]
    r = t.query(q).nearest_one().approx().execute()
    assert int(r.item[0]) == 0
    with pytest.raises(ValueError):
        t.query(q).within(1.0).approx()
    with pytest.raises(ka.KiddoAmdError):
        t.query(q).nearest_n(8).screen_f32().execute()  # no FP32 shadow on this tree


def test_traversal_visits_match_oracle(ka, oracle_mod, cuda):
    """The plain and group kernels visit exactly the leaves (and scan exactly the entries)
    that the oracle's restatement of Kiddo's DFS visits for the same queries. The
    thread-per-query FP32-screened kernel first scans the query's pre-scan block (the 4
    leaves under one stem here) and then runs the DFS from a certified bound on the k-th
    distance, so it visits at most 3 leaves per query more than Kiddo's DFS (its results
    are checked elsewhere)."""
    _announce("visit counters")
    rng = np.random.default_rng(12)
    pts = rng.random((100_000, 3))
    q = rng.random((5000, 3))
    t = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    ot = oracle_mod.Tree.build(pts, 32)
    _, _, _, ost = ot.nearest_n(q, 16, stats=True)
    for variant in (1, 2):
        t.tune(variant, 0, -1)
        for flags in (0, ka.KD_MORTON, ka.KD_SCREEN_F32 | ka.KD_MORTON):
            t.nearest_n_raw(q, 16, flags=flags)
            st = t.query_stats()
            if variant == 1 and flags & ka.KD_SCREEN_F32:
                assert 0 < st["leaves"] <= int(ost[0]) + 3 * len(q), (variant, flags, st, ost)
                assert 0 < st["points"] <= int(ost[1]) + 3 * 32 * len(q), (variant, flags, st, ost)
                continue
            assert st["leaves"] == int(ost[0]), (variant, flags, st, ost)
            assert st["points"] == int(ost[1]), (variant, flags, st, ost)
    t.tune(0, 0, -1)
    t.nearest_one_raw(q)
    st = t.query_stats()
    _, _, ost = ot.nearest_one(q, stats=True)
    assert (st["leaves"], st["points"]) == (int(ost[0]), int(ost[1]))
    t.set_timing(True)
    t.nearest_n_raw(q, 16)
    assert t.last_kernel_ms() > 0


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_within_dense_group_fill(ka, oracle_mod, cuda, dtype):
    """Dense within outputs (>= 32 results per query) take the group-per-query fill; its
    results equal the oracle's (sorted: bit for bit; unsorted: the same multiset per query,
    in the thread form's order)."""
    rng = np.random.default_rng(21)
    pts = rng.random((30_000, 3)).astype(dtype)
    q = rng.random((3000, 3)).astype(dtype)
    q[:50] = pts[:50]
    r2 = 0.01  # about 100 entries per query (r = 0.1)
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    ot = oracle_mod.Tree.build(pts, 32)
    for excl in (False, True):
        _announce("dense within", np.dtype(dtype).name, excl)
        oo, oi, od = ot.within(q, r2, excl, True)
        assert int(oo[-1]) >= 32 * len(q)
        for flags in (0, ka.KD_MORTON):
            for count, fill in ((0, 0), (1, 1), (4, 4), (8, 8), (16, 16), (32, 32), (8, 1), (1, 16)):
                t.tune_within(count, fill)  # every launch shape gives the same results
                f = flags | (ka.KD_EXCLUSIVE if excl else 0)
                r = t.within_raw(q, r2, f)
                assert_same_bits(r.offsets.cpu().numpy().astype(np.uint64), oo, "offsets")
                assert_same_bits(r.items.cpu().numpy(), oi, f"items shapes={count},{fill}")
                assert_same_bits(r.distances.cpu().numpy(), od, f"dists shapes={count},{fill}")
                ru = t.within_raw(q, r2, f | ka.KD_UNSORTED)
                ui, ud = ru.items.cpu().numpy(), ru.distances.cpu().numpy()
                qid = np.repeat(np.arange(len(q)), np.diff(oo.astype(np.int64)))
                key = np.lexsort((ui, ud, qid))
                assert_same_bits(ui[key], oi, f"unsorted items as multisets shapes={count},{fill}")
                assert_same_bits(ud[key], od, f"unsorted dists as multisets shapes={count},{fill}")
            t.tune_within(0, 0)
            chosen = t.autotune_within(q, r2, flags)
            r = t.within_raw(q, r2, flags | (ka.KD_EXCLUSIVE if excl else 0))
            assert_same_bits(r.items.cpu().numpy(), oi, f"items autotuned {chosen}")
            t.tune_within(0, 0)


@pytest.mark.parametrize("r2,dense", [(0.006, True), (0.0004, False)], ids=["dense", "sparse"])
def test_within_large_batch(ka, oracle_mod, cuda, r2, dense):
    """70,000 queries, dense (group-per-query fill) and sparse (thread fill) outputs:
    offsets and sorted results equal the oracle's."""
    rng = np.random.default_rng(23)
    pts = rng.random((30_000, 3))
    q = rng.random((70_000, 3))
    q[:64] = pts[:64]
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    ot = oracle_mod.Tree.build(pts, 32)
    for excl in (False, True):
        _announce("within large batch", r2, excl)
        oo, oi, od = ot.within(q, r2, excl, True)
        assert (int(oo[-1]) >= 32 * len(q)) == dense
        for flags in (0, ka.KD_MORTON):
            r = t.within_raw(q, r2, flags | (ka.KD_EXCLUSIVE if excl else 0))
            assert_same_bits(r.offsets.cpu().numpy().astype(np.uint64), oo, "offsets")
            assert_same_bits(r.items.cpu().numpy(), oi, "items")
            assert_same_bits(r.distances.cpu().numpy(), od, "dists")


SCREEN_CASES = [
    ("grid-dups-k3", lambda rng: rng.integers(0, 7, (40_000, 3)).astype(np.float64), 32),
    ("grid-dups-b8", lambda rng: rng.integers(0, 5, (6_000, 3)).astype(np.float64), 8),
    ("tiny-20", lambda rng: rng.random((20, 3)), 32),
    ("tiny-70", lambda rng: rng.random((70, 3)), 32),
    ("uniform-k2", lambda rng: rng.random((50_000, 2)), 32),
    ("clusters", lambda rng: np.concatenate([rng.normal(c, 0.01, (5000, 3)) for c in rng.random((8, 3))]), 16),
    # 200 exact copies of each of 10 points: a query on a copy has more survivors at
    # distance 0 in its first block than k, so the sorted first block stops at the k-th and
    # the traversal's visit of the block takes the rest
    ("hot-duplicates", lambda rng: np.concatenate([rng.random((20_000, 3)), np.repeat(rng.random((10, 3)), 200, axis=0)]), 32),
]


@pytest.mark.parametrize("name,make,B", SCREEN_CASES, ids=[c[0] for c in SCREEN_CASES])
def test_screened_edge_cases(ka, oracle_mod, cuda, name, make, B):
    """The FP32-screened kernel (pre-scan block as the first visit, survivor masks, entries
    past the masks screened one at a time) against the oracle: long leaves of equal points,
    trees smaller than the pre-scan block, k above n, 2-D trees, clusters; k from 1 to 64,
    with and without a radius."""
    rng = np.random.default_rng(31)
    pts = make(rng)
    q = _queries(pts, 3000, 5)
    t = ka.ImmutableKdTree.new_from_slice(pts, B, shadow_f32=True)
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, B)
    r2 = float(np.median(O.nearest_n(ot, q[:200], 8)[1][:, -1]))
    for k in (1, 5, 16, 32, 64):
        for radius, excl in ((math.inf, False), (r2, False), (r2, True)):
            _announce("screened edge", name, k, radius, excl)
            ref = O.nearest_n(ot, q, k, radius, excl)
            for flags in (ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON):
                f = flags | (ka.KD_EXCLUSIVE if excl else 0)
                r = t.nearest_n_raw(q, k, radius, f)
                got = (r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy())
                for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
                    assert_same_bits(a, b, f"{name} k={k} r={radius} excl={excl} flags={f} {nm}")


def test_screened_redo_pass(ka, oracle_mod, cuda):
    """The FP32 collector keeps the candidates found after the pre-scan block in a list of
    k + 8 entries. On a 7 x 7 x 7 grid holding ~117 exact copies of each point, a query on a
    grid point ties with copies spread over many leaves: its list overflows and the exact
    FP64 redo pass answers it. 300 exact copies of a point mostly share the query's pre-scan
    block, whose ties are resolved in the kernel's final pass. Both against the oracle; the
    redo count is reported by kd_query_redos."""
    rng = np.random.default_rng(7)
    grid = rng.integers(0, 7, (40_000, 3)).astype(np.float64)
    hot = np.concatenate([rng.random((30_000, 3)), np.repeat(rng.random((3, 3)), 300, axis=0)])
    for name, pts in (("grid", grid), ("hot", hot)):
        q = rng.random((4000, 3)) * float(pts.max())
        q[:1000] = pts[rng.integers(0, len(pts), 1000)]
        t = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
        # one thread per query: the FP32 collector and its redo pass (a batch this small
        # would take the group form automatically)
        t.tune(1, 0, -1)
        O = OracleEngine(oracle_mod)
        ot = O.build(pts, 32)
        for k in (1, 4, 8, 16):
            ref = O.nearest_n(ot, q, k)
            for flags in (ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON):
                _announce("screened redo", name, k, flags)
                r = t.nearest_n_raw(q, k, math.inf, flags)
                got = (r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy())
                for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
                    assert_same_bits(a, b, f"redo {name} k={k} flags={flags} {nm}")
                redos = t.query_redos()
                if name == "grid":  # every query ties with a whole stack of copies
                    assert 0 < redos <= len(q), f"grid k={k}: {redos} queries redone"
                else:
                    assert redos <= 1100, (name, k, redos)
    # no ties: nothing is redone
    u = rng.random((20_000, 3))
    tu = ka.ImmutableKdTree.new_from_slice(u, 32, shadow_f32=True)
    tu.tune(1, 0, -1)
    tu.nearest_n_raw(rng.random((4000, 3)), 16, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON)
    assert tu.query_redos() == 0


def test_screened_redo_concurrent_streams(ka, oracle_mod, cuda):
    """Screened calls on one tree from two streams at once, on a tree whose queries go to the
    redo pass: the redo pass reads only its own call's counts, so neither call can lose the
    other's marks when the tree's shared counters are cleared. Every call equals the oracle."""
    import torch
    rng = np.random.default_rng(11)
    pts = rng.integers(0, 7, (40_000, 3)).astype(np.float64)
    t = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    t.tune(1, 0, -1)  # one thread per query: the FP32 collector and its redo pass
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, 32)
    qs = [rng.random((30_000, 3)) * 6.0 for _ in range(2)]
    refs = [O.nearest_n(ot, q, 16) for q in qs]
    qd = [torch.from_numpy(q).cuda() for q in qs]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for rep in range(4):
        _announce("screened redo streams", rep)
        outs = []
        for s, q in zip(streams, qd):
            with torch.cuda.stream(s):
                outs.append(t.nearest_n_raw(q, 16, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON, stream=s))
        torch.cuda.synchronize()
        for r, ref in zip(outs, refs):
            got = (r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy())
            for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
                assert_same_bits(a, b, f"streams rep={rep} {nm}")
