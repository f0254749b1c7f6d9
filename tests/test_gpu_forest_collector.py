"""The FP32 collector in the batched forest (f32 trees created with an FP32 shadow,
KD_SCREEN_F32): every tree's results bit-identical to the oracle run on that tree alone,
under every thread-form launch shape, Morton order, radii, hipGraph replay, and with a tree
of exact ties whose queries overflow the candidate list (the forest's redo pass)."""
import math

import numpy as np
import pytest

from engines import assert_same_bits
from test_gpu_forest import _check_knn, _forest_inputs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def collector_case(ka, oracle_mod, cuda):
    torch = cuda
    print("case: forest collector build", flush=True)
    pts, qs = _forest_inputs(ntrees=40, seed=9)
    # a tree of exact ties: a 6 x 6 x 6 grid holding ~46 copies of each point
    rng = np.random.default_rng(3)
    pts[5] = (rng.integers(0, 6, (10_000, 3)) * 0.25).astype(np.float32)
    qs[5] = pts[5][rng.integers(0, len(pts[5]), 600)].copy()
    forest = ka.Forest.from_points(pts, bucket=32, shadow_f32=True)
    offs = np.zeros(len(qs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(q) for q in qs])
    qcat = torch.from_numpy(np.concatenate(qs)).cuda()
    refs = [oracle_mod.Tree.build(p, 32) for p in pts]
    return forest, qcat, offs, pts, qs, refs


def test_forest_collector_matches_per_tree_oracle(ka, collector_case):
    forest, qcat, offs, pts, qs, refs = collector_case
    f = ka.KD_SCREEN_F32
    for morton in (False, True):
        forest.plan(qcat, offs, chunk=256, morton=morton)
        for lds, waves in ((0, 0), (8, 4), (11, 5)):
            forest.tune(1, 0, lds, waves)
            for k in (1, 5, 8, 16):
                print(f"case: forest collector morton={morton} lds={lds} waves={waves} k={k}", flush=True)
                res = forest.nearest_n(qcat, k, flags=f)
                _check_knn(res, offs, qs, refs, k, tag=f"collector m={morton} lds={lds} w={waves} k={k}")
            for excl in (False, True):
                res = forest.nearest_n(qcat, 8, 0.01, flags=f | (ka.KD_EXCLUSIVE if excl else 0))
                _check_knn(res, offs, qs, refs, 8, 0.01, excl, tag=f"collector radius excl={excl}")
    forest.tune(1, 0, 8, 0)


def test_forest_collector_graph_replays(ka, cuda, collector_case):
    torch = cuda
    forest, qcat, offs, pts, qs, refs = collector_case
    forest.plan(qcat, offs, chunk=256, morton=True)
    forest.tune(1, 0, 8, 4)
    eager = forest.nearest_n(qcat, 8, flags=ka.KD_SCREEN_F32)
    torch.cuda.synchronize()
    for rep in range(3):
        print(f"case: forest collector graph replay {rep}", flush=True)
        g = forest.nearest_n(qcat, 8, flags=ka.KD_SCREEN_F32 | ka.KD_GRAPH)
        torch.cuda.synchronize()
        for a, b, nm in ((g.items, eager.items, "items"), (g.distances, eager.distances, "dists"),
                         (g.counts, eager.counts, "counts")):
            assert_same_bits(a.cpu().numpy(), b.cpu().numpy(), f"graph replay {rep} {nm}")
        forest.nearest_n(qcat, 4)  # other eager work between replays
        torch.cuda.synchronize()
    forest.tune(1, 0, 8, 0)


def test_forest_collector_needs_shadow(ka, cuda):
    torch = cuda
    pts = [np.random.default_rng(1).random((500, 3)).astype(np.float32)]
    forest = ka.Forest.from_points(pts, bucket=32)
    q = torch.from_numpy(pts[0][:10].copy()).cuda()
    forest.plan(q, np.array([0, 10], np.uint64), morton=False)
    with pytest.raises(Exception):
        forest.nearest_n(q, 4, flags=ka.KD_SCREEN_F32)
    with pytest.raises(Exception):
        ka.Forest.from_points([np.random.default_rng(2).random((100, 3))], bucket=32, shadow_f32=True)  # f64
