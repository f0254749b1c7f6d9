"""The FP32 collector on f32 trees (KD_SCREEN_F32 on a tree with an FP32 shadow): FP32
distances only during the traversal, items resolved at the end. Bit for bit against the
oracle (items, order, distance bits, counts) on the parity cases, with radii, exclusive
radii, exact ties that overflow the candidate list (the redo pass), and every launch form."""
import math

import numpy as np
import pytest

from engines import GpuEngine, OracleEngine, assert_same_bits
from test_gpu_parity import _points, _queries, _radius

pytestmark = pytest.mark.gpu

NAMES = ("items", "dists", "counts")

F32_CASES = [
    ("uniform", 100_000, 3, 32),
    ("dups", 60_000, 3, 32),
    ("grid", 30_000, 2, 8),
    ("grid", 40_000, 3, 32),
    ("mixture", 80_000, 3, 32),
    ("uniform", 50_000, 4, 16),
    ("all_same", 800, 3, 32),
    ("uniform", 20, 3, 32),
    ("uniform", 1, 3, 32),
]


def _announce(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("kind,n,K,B", F32_CASES, ids=[f"{c[0]}-{c[1]}-K{c[2]}-B{c[3]}" for c in F32_CASES])
def test_f32_collector_matches_oracle(ka, oracle_mod, cuda, kind, n, K, B):
    _announce("f32 collector", kind, n, K, B)
    pts = _points(kind, n, K, np.float32, seed=n + K + 11)
    q = _queries(pts, 4000, seed=n + 5)
    r = np.float32(_radius(pts))
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, B)
    for flags in (ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON):
        G = GpuEngine(ka, flags=flags, shadow=True)
        gt = G.build(pts, B)
        for k in (1, 4, 5, 8, 16):
            for radius, excl in ((math.inf, False), (float(r), False), (float(r), True)):
                ref = O.nearest_n(ot, q, k, radius, excl)
                for variant in (1, 0):  # one thread per query (the collector); automatic
                    gt.tune(variant, 0, -1)
                    res = G.nearest_n(gt, q, k, radius, excl)
                    for a, b, nm in zip(res, ref, NAMES):
                        assert_same_bits(a, b, f"k={k} r={radius} excl={excl} variant={variant} flags={flags} {nm}")
        gt.tune(0, 0, -1)


def test_f32_collector_redo_and_stats(ka, oracle_mod, cuda):
    """A 7 x 7 x 7 grid holding ~117 copies of each point: queries on grid points tie with
    copies over many leaves, their lists overflow and the exact redo pass answers them.
    Uniform points: nothing is redone, and results equal the plain kernel's."""
    _announce("f32 collector redo")
    rng = np.random.default_rng(17)
    grid = rng.integers(0, 7, (40_000, 3)).astype(np.float32)
    q = (rng.random((4000, 3)) * 6.0).astype(np.float32)
    q[:1000] = grid[rng.integers(0, len(grid), 1000)]
    t = ka.ImmutableKdTree.new_from_slice(grid, 32, shadow_f32=True)
    t.tune(1, 0, -1)
    O = OracleEngine(oracle_mod)
    ot = O.build(grid, 32)
    for k in (1, 4, 16):
        ref = O.nearest_n(ot, q, k)
        r = t.nearest_n_raw(q, k, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON)
        got = (r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy())
        for a, b, nm in zip(got, ref, NAMES):
            assert_same_bits(a, b, f"grid k={k} {nm}")
        assert 0 < t.query_redos() <= len(q)
    u = rng.random((100_000, 3)).astype(np.float32)
    qu = rng.random((20_000, 3)).astype(np.float32)
    tu = ka.ImmutableKdTree.new_from_slice(u, 32, shadow_f32=True)
    tu.tune(1, 0, -1)
    plain = tu.nearest_n_raw(qu, 16, math.inf, ka.KD_MORTON)
    coll = tu.nearest_n_raw(qu, 16, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON)
    assert tu.query_redos() == 0
    for a, b, nm in zip((coll.items, coll.distances, coll.counts), (plain.items, plain.distances, plain.counts), NAMES):
        assert_same_bits(a.cpu().numpy(), b.cpu().numpy(), f"uniform collector vs plain {nm}")


def test_f32_shadow_device_build(ka, oracle_mod, cuda):
    """A tree built on the GPU with the shadow answers like the host-built one."""
    import torch
    _announce("f32 collector device build")
    rng = np.random.default_rng(23)
    pts = rng.random((70_000, 3)).astype(np.float32)
    q = rng.random((5000, 3)).astype(np.float32)
    gt = ka.ImmutableKdTree.build_on_device(torch.from_numpy(pts).cuda(), 32, shadow_f32=True)
    gt.tune(1, 0, -1)
    ref = OracleEngine(oracle_mod).build(pts, 32).nearest_n(q, 8)
    r = gt.nearest_n_raw(q, 8, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON)
    for a, b, nm in zip((r.items, r.distances, r.counts), ref, NAMES):
        assert_same_bits(a.cpu().numpy(), b, f"device-built f32 collector {nm}")
