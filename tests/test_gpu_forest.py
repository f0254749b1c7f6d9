"""Batched form on the GPU: many trees of uneven sizes with their own query sets in one
persistent launch (thread and group forms, LDS stems, Morton order, hipGraph replay),
each tree's results bit-identical to the oracle run on that tree alone."""
import numpy as np
import pytest

from engines import assert_same_bits

pytestmark = pytest.mark.gpu


def _forest_inputs(ntrees=48, seed=5):
    from kiddo_amd.datasets import noisy_sphere
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(40), np.log(20000), ntrees)).astype(int)
    sizes[0], sizes[1], sizes[2] = 0, 1, 31  # empty, single entry, single leaf
    nq = np.exp(rng.uniform(np.log(16), np.log(3000), ntrees)).astype(int)
    pts = [noisy_sphere(int(n), seed=seed * 1000 + i) for i, n in enumerate(sizes)]
    qs = [noisy_sphere(int(m), seed=seed * 7000 + i) for i, m in enumerate(nq)]
    for i in range(3, ntrees, 5):  # some queries exactly on entries
        if len(pts[i]):
            qs[i][: len(qs[i]) // 3] = pts[i][rng.integers(0, len(pts[i]), len(qs[i]) // 3)]
    return pts, qs


@pytest.fixture(scope="module")
def forest_case(ka, oracle_mod, cuda):
    torch = cuda
    print("case: forest build", flush=True)
    pts, qs = _forest_inputs()
    forest = ka.Forest.from_points(pts, bucket=32)
    offs = np.zeros(len(qs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(q) for q in qs])
    qcat = torch.from_numpy(np.concatenate(qs)).cuda()
    refs = [oracle_mod.Tree.build(p, 32) for p in pts]
    return forest, qcat, offs, pts, qs, refs


def _check_knn(res, offs, qs, refs, k, radius=np.inf, excl=False, tag=""):
    items = res.items.cpu().numpy()
    dists = res.distances.cpu().numpy()
    counts = res.counts.cpu().numpy()
    for t, (q, ref) in enumerate(zip(qs, refs)):
        s, e = int(offs[t]), int(offs[t + 1])
        ri, rd, rc = ref.nearest_n(q, k, radius, excl)
        assert_same_bits(items[s:e], ri, f"tree {t} items {tag}")
        assert_same_bits(dists[s:e], rd, f"tree {t} dists {tag}")
        assert_same_bits(counts[s:e], rc, f"tree {t} counts {tag}")


def test_forest_knn_matches_per_tree_oracle(ka, forest_case):
    forest, qcat, offs, pts, qs, refs = forest_case
    for morton in (False, True):
        forest.plan(qcat, offs, chunk=256, morton=morton)
        for variant, gs, lds, waves in ((1, 0, 0, 0), (1, 0, 8, 0), (1, 0, 11, 4), (2, 8, 8, 0), (2, 16, 0, 0)):
            print(f"case: forest knn morton={morton} variant={variant}/{gs}/{lds}/{waves}", flush=True)
            forest.tune(variant, gs, lds, waves)
            res = forest.nearest_n(qcat, 8)
            _check_knn(res, offs, qs, refs, 8, tag=f"m={morton} v={variant}/{gs}/{lds}/{waves}")
    forest.tune(1, 0, 8, 0)
    res = forest.nearest_n(qcat, 5, 4e-4, ka.KD_EXCLUSIVE)
    _check_knn(res, offs, qs, refs, 5, 4e-4, True, "within-radius exclusive")


def test_forest_graph_replays(ka, cuda, forest_case):
    torch = cuda
    forest, qcat, offs, pts, qs, refs = forest_case
    forest.plan(qcat, offs, chunk=128, morton=True)
    forest.tune(1, 0, 8, 0)
    eager = forest.nearest_n(qcat, 8)
    e_items, e_d = eager.items.clone(), eager.distances.clone()
    out = forest._knn_out(8)
    for rep in range(4):
        print(f"case: forest graph replay {rep}", flush=True)
        out.items.zero_()
        out.distances.zero_()
        forest.nearest_n(qcat, 8, flags=ka.KD_GRAPH, out=out)
        torch.cuda.synchronize()
        assert torch.equal(out.items, e_items)
        assert torch.equal(out.distances.view(torch.int32), e_d.view(torch.int32))
        # eager work between replays
        forest.nearest_n(qcat, 4)
        torch.cuda.synchronize()


def test_forest_within_matches_per_tree_oracle(ka, forest_case):
    forest, qcat, offs, pts, qs, refs = forest_case
    forest.plan(qcat, offs, chunk=256, morton=True)
    forest.tune(1, 0, 8, 0)
    r2 = 4e-4
    res = forest.within(qcat, r2)
    go = res.offsets.cpu().numpy().astype(np.int64)
    gi = res.items.cpu().numpy()
    gd = res.distances.cpu().numpy()
    for t, (q, ref) in enumerate(zip(qs, refs)):
        s, e = int(offs[t]), int(offs[t + 1])
        oo, oi, od = ref.within(q, r2)
        base = go[s]
        assert_same_bits((go[s:e + 1] - base).astype(np.uint64), oo, f"tree {t} offsets")
        assert_same_bits(gi[go[s]:go[e]], oi, f"tree {t} items")
        assert_same_bits(gd[go[s]:go[e]], od, f"tree {t} dists")


def test_forest_autotune(ka, forest_case):
    forest, qcat, offs, pts, qs, refs = forest_case
    forest.plan(qcat, offs, chunk=256, morton=True)
    chosen = forest.autotune_nearest_n(qcat, 8)
    assert chosen["variant"] in (1, 2)
    res = forest.nearest_n(qcat, 8)
    _check_knn(res, offs, qs, refs, 8, tag=f"autotuned {chosen}")
