"""GPU parity for k and n above the register collectors (k > 64: a max-heap in each
query's output row), through the C ABI, bit for bit against the oracle: nearest_n with and
without a radius (inclusive and exclusive), FP32-screened, Morton-ordered, f32 / f64 / u16
trees, k above the tree size, periodic nearest_n, another metric, best_n_within, and a
served request. Kiddo takes any k (builder.rs:2019-2028); it keeps up to 192 results in a
sorted vector and more in a BinaryHeap (nearest_n_within.rs:24-27, :161-176)."""
import math

import numpy as np
import pytest

from engines import GpuEngine, OracleEngine, assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


def _cmp(got, ref, what):
    for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
        assert_same_bits(a, b, f"{what} {nm}")


def _res(r):
    return r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy()


@pytest.fixture(scope="module")
def dup_case(ka, oracle_mod, cuda):
    rng = np.random.default_rng(41)
    base = rng.random((40_000, 3))
    pts = np.concatenate([base, base[rng.integers(0, 40_000, 4000)]])
    q = rng.random((3000, 3))
    q[:300] = pts[rng.integers(0, len(pts), 300)]
    return pts, q, oracle_mod.Tree.build(pts, 32)


def test_large_k_f64_plain_screened_morton(ka, dup_case, cuda):
    pts, q, ot = dup_case
    t = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    for k in (65, 128, 192, 256):
        ref = ot.nearest_n(q, k, threads=8)
        for flags in (0, ka.KD_MORTON, ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON):
            _announce("large k f64", k, flags)
            _cmp(_res(t.nearest_n_raw(q, k, flags=flags)), ref, f"k={k} flags={flags}")
        if k == 128:
            # the FP32 screen re-measures fewer entries than it scans
            t.nearest_n_raw(q, k, flags=ka.KD_SCREEN_F32)
            st = t.query_stats()
            assert 0 < st["rescored"] < st["points"], st
    for excl in (False, True):
        for k in (100, 300):
            ref = ot.nearest_n(q, k, 1.5e-3, excl, threads=8)
            for flags in (0, ka.KD_SCREEN_F32 | ka.KD_MORTON):
                f = flags | (ka.KD_EXCLUSIVE if excl else 0)
                _announce("large k within", k, excl, flags)
                _cmp(_res(t.nearest_n_raw(q, k, 1.5e-3, f)), ref, f"within k={k} excl={excl} flags={f}")


def test_large_k_visits_match_oracle(ka, dup_case, cuda):
    """The heap collector follows Kiddo's DFS: the same leaves and entries as the oracle."""
    pts, q, ot = dup_case
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    _, _, _, ost = ot.nearest_n(q, 100, stats=True)
    t.nearest_n_raw(q, 100)
    st = t.query_stats()
    assert (st["leaves"], st["points"]) == (int(ost[0]), int(ost[1])), (st, ost)


@pytest.mark.parametrize("dtype,K", [(np.float32, 3), (np.uint16, 4), (np.float64, 2)])
def test_large_k_other_dtypes(ka, oracle_mod, cuda, dtype, K):
    rng = np.random.default_rng(7)
    if dtype == np.uint16:
        pts = rng.integers(0, 65535, (30_000, K)).astype(dtype)
        q = rng.integers(0, 65535, (1500, K)).astype(dtype)
        r = 4.0e7
    else:
        pts = rng.random((30_000, K)).astype(dtype)
        q = rng.random((1500, K)).astype(dtype)
        r = 2e-3
    q[:100] = pts[:100]
    ot = oracle_mod.Tree.build(pts, 32)
    G = GpuEngine(ka, flags=ka.KD_MORTON)
    gt = G.build(pts, 32)
    for k, rad in ((65, math.inf), (200, math.inf), (150, r)):
        _announce("large k", np.dtype(dtype).name, K, k, rad)
        _cmp(G.nearest_n(gt, q, k, rad), ot.nearest_n(q, k, rad, threads=8), f"k={k} r={rad}")
    for n in (65, 100, 250):
        _announce("large n best_n_within", np.dtype(dtype).name, n)
        _cmp(G.best_n_within(gt, q, r, n), ot.best_n_within(q, r, n, threads=8), f"best_n_within n={n}")


def test_k_above_tree_size(ka, oracle_mod, cuda):
    rng = np.random.default_rng(3)
    pts = rng.random((90, 3))
    q = rng.random((500, 3))
    ot = oracle_mod.Tree.build(pts, 32)
    G = GpuEngine(ka)
    gt = G.build(pts, 32)
    for k in (65, 90, 91, 400):
        got = G.nearest_n(gt, q, k)
        _cmp(got, ot.nearest_n(q, k), f"k={k} above n")
        assert np.all(got[2] == min(k, 90))
    _cmp(G.best_n_within(gt, q, 0.3, 100), ot.best_n_within(q, 0.3, 100), "best_n_within n > n")
    empty = G.build(np.zeros((0, 3)), 32)
    it, d, c = G.nearest_n(empty, q[:4], 100)
    assert c.tolist() == [0] * 4 and np.all(it == 0xFFFFFFFF) and np.all(np.isinf(d))


def test_large_k_periodic_and_metric(ka, oracle_mod, cuda):
    rng = np.random.default_rng(5)
    pts = rng.random((25_000, 3))
    q = rng.random((1200, 3))
    box = np.ones(3)
    ot = oracle_mod.Tree.build(pts, 32)
    G = GpuEngine(ka, flags=ka.KD_MORTON)
    gt = G.build(pts, 32)
    for k in (80, 200):
        _announce("periodic large k", k)
        _cmp(G.periodic_nearest_n(gt, q, box, k), ot.periodic_nearest_n(q, box, k, threads=8),
             f"periodic k={k}")
    _announce("manhattan large k")
    r = gt.nearest_n_raw(q, 128, flags=ka.KD_MORTON | ka.KD_METRIC_MANHATTAN)
    om = oracle_mod.Tree.build(pts, 32).set_metric(ka.KD_METRIC_MANHATTAN)
    _cmp(_res(r), om.nearest_n(q, 128, threads=8), "manhattan k=128")


def test_large_k_served(ka, oracle_mod, cuda):
    rng = np.random.default_rng(9)
    pts = rng.random((20_000, 3))
    t = ka.ImmutableKdTree.new_from_slice(pts, 32)
    ot = oracle_mod.Tree.build(pts, 32)
    srv = ka.Server([t], arena_bytes=64 << 20, slots=2)
    qs = [rng.random((m, 3)) for m in (1, 300, 2000)]
    tickets = [srv.submit(0, ka.FAMILY_NEAREST_N, qq, 100) for qq in qs]
    tickets += [srv.submit(0, ka.FAMILY_BEST_N_WITHIN, qq, 90, 0.01) for qq in qs]
    res = [srv.wait(tk) for tk in tickets]
    srv.close()
    for qq, r in zip(qs, res[:3]):
        _cmp(r[:3], ot.nearest_n(qq, 100), f"served nearest_n(100) m={len(qq)}")
    for qq, r in zip(qs, res[3:]):
        _cmp(r[:3], ot.best_n_within(qq, 0.01, 90), f"served best_n_within(90) m={len(qq)}")
