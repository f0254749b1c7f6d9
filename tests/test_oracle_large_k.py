"""Pins the oracle's k-NN and best_n_within collectors above the sizes the other tests use
(Kiddo switches from its sorted vector to a BinaryHeap above 192 results,
src/kd_tree/query/nearest_n_within.rs:24-27 and :161-176): nearest_n, nearest_n().within
and best_n_within against a linear scan with the canonical rules, for k and n up to 1000,
k above the tree size, duplicate-heavy data and exclusive radii."""
import numpy as np
import pytest

from engines import assert_same_bits


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_large_k_matches_brute_force(oracle_mod, dtype):
    rng = np.random.default_rng(13)
    base = rng.random((6000, 3))
    pts = np.concatenate([base, base[rng.integers(0, 6000, 1500)]]).astype(dtype)
    q = rng.random((400, 3)).astype(dtype)
    q[:50] = pts[:50]
    t = oracle_mod.Tree.build(pts, 32)
    for k in (65, 128, 192, 193, 256, 1000):
        got = t.nearest_n(q, k, threads=4)
        ref = t.brute_nearest_n(q, k, threads=4)
        for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
            assert_same_bits(a, b, f"k={k} {nm}")
    for excl in (False, True):
        got = t.nearest_n(q, 300, 2e-3, excl, threads=4)
        ref = t.brute_nearest_n(q, 300, 2e-3, excl, threads=4)
        for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
            assert_same_bits(a, b, f"within k=300 excl={excl} {nm}")
        for n in (65, 100, 300):
            got = t.best_n_within(q, 4e-3, n, excl, threads=4)
            ref = t.brute_best_n_within(q, 4e-3, n, excl, threads=4)
            for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
                assert_same_bits(a, b, f"best_n_within n={n} excl={excl} {nm}")
    small = oracle_mod.Tree.build(pts[:70], 32)
    got = small.nearest_n(q, 100)
    assert np.all(got[2] == 70)
    assert np.all(got[0][:, 70:] == 0xFFFFFFFF)
