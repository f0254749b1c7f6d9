"""Bucket-size autotuners (kd_autotune_bucket, kd_forest_autotune_bucket): each returns one
of the candidates with a time for every candidate, and the tree or forest built at the
chosen bucket answers exactly as the oracle (the canonical answer does not depend on the
bucket size)."""
import math

import numpy as np
import pytest

from engines import OracleEngine, assert_same_bits

pytestmark = pytest.mark.gpu


def test_autotune_bucket_single_tree(ka, oracle_mod, cuda):
    import torch
    from kiddo_amd.datasets import uniform_with_duplicates
    pts = uniform_with_duplicates(200_000, 0.05, seed=5)
    q = np.random.default_rng(6).random((50_000, 3))
    qd = torch.from_numpy(q).cuda()
    cands = (8, 32, 128)
    for flags, shadow in ((ka.KD_MORTON, False), (ka.KD_MORTON | ka.KD_SCREEN_F32, True)):
        print("case: autotune bucket", flags, flush=True)
        best, ms = ka.ImmutableKdTree.autotune_bucket(torch.from_numpy(pts).cuda(), qd, 8, flags=flags,
                                                      buckets=cands, shadow_f32=shadow)
        assert best in cands and sorted(ms) == sorted(cands)
        assert all(0.0 < v < 1e4 for v in ms.values()), ms
        assert ms[best] == min(ms.values())
        t = ka.ImmutableKdTree.build_on_device(pts, best, shadow_f32=shadow)
        r = t.nearest_n_raw(qd, 8, math.inf, flags)
        O = OracleEngine(oracle_mod)
        ref = O.nearest_n(O.build(pts, 32), q, 8)
        for a, b, nm in zip((r.items.cpu().numpy(), r.distances.cpu().numpy(), r.counts.cpu().numpy()), ref,
                            ("items", "dists", "counts")):
            assert_same_bits(a, b, f"bucket {best} {nm}")


def test_autotune_bucket_forest(ka, oracle_mod, cuda):
    import torch
    from kiddo_amd.datasets import noisy_sphere
    rng = np.random.default_rng(3)
    sizes = rng.integers(1_000, 20_000, 24)
    nq = rng.integers(256, 3_000, 24)
    pts = [noisy_sphere(int(n), seed=100 + i) for i, n in enumerate(sizes)]
    qs = [noisy_sphere(int(n), seed=900 + i) for i, n in enumerate(nq)]
    offs = np.zeros(len(qs) + 1, np.uint64)
    offs[1:] = np.cumsum(nq)
    q = torch.from_numpy(np.concatenate(qs)).cuda()
    print("case: forest autotune bucket", flush=True)
    best, ms = ka.Forest.autotune_bucket(pts, q, offs, 8, buckets=(16, 64))
    assert best in (16, 64) and set(ms) == {16, 64} and ms[best] == min(ms.values())
    f = ka.Forest.from_points(pts, bucket=best)
    f.plan(q, offs, chunk=256, morton=True)
    r = f.nearest_n(q, 8)
    items, d = r.items.cpu().numpy(), r.distances.cpu().numpy()
    O = OracleEngine(oracle_mod)
    for i in range(0, len(pts), 6):
        ri, rd, rc = O.nearest_n(O.build(pts[i], 32), qs[i], 8)
        lo, hi = int(offs[i]), int(offs[i + 1])
        assert_same_bits(items[lo:hi], ri, f"tree {i} items")
        assert_same_bits(d[lo:hi], rd, f"tree {i} dists")
