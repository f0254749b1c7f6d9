"""Donnelly stem layout (kiddo src/stem_strategy/donnelly/), on the CPU:

  * the oracle's restatement of the Donnelly index step against the reference's own test
    vectors (tests/golden/donnelly_kats.json, transcribed from core.rs by
    tests/golden/extract_donnelly_kats.py): path -> stem index, path -> both child indices,
    the unrolled head/tail form, and block steps;
  * the reference's regression properties (core.rs:704-839): a block step of BH levels
    lands where BH single steps do, and a node's two children are the single steps left
    and right;
  * the stem-depth padding of a block layout (kd_tree/mod.rs:733-789: 100 points,
    Donnelly<4>, bucket 32 -> depth 4, max_stem_level 3) and the padded tree's layout
    invariants and query results against brute force;
  * the product's kd_layout_from_donnelly: the Donnelly-ordered tree the oracle exports,
    re-laid out in Eytzinger order, equals the oracle's padded Eytzinger tree byte for
    byte (stems, extents, arena), and bad layouts are refused.
"""
import json
import os

import numpy as np
import pytest

from engines import assert_same_bits

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
with open(os.path.join(GOLDEN, "donnelly_kats.json")) as f:
    KATS = json.load(f)
BH = KATS["block_height"]


def test_reference_path_vectors(oracle_mod):
    O = oracle_mod
    g = KATS["groups"]
    n = 0
    for name in ("donnelly_core_get_child_idx_produces_correct_values",
                 "donnelly_core_get_child_idx_unrolled_produces_correct_values"):
        for c in g[name]:
            assert O.donnelly_path(BH, c["input"])[0] == c["expect"], (name, c)
            n += 1
    for c in g["donnelly_core_get_both_child_idxs_produces_correct_values"]:
        left = O.donnelly_path(BH, c["input"] + [False])[0]
        right = O.donnelly_path(BH, c["input"] + [True])[0]
        assert [left, right] == c["expect"], c
        n += 1
    for c in g["donnelly_core_traverse_block_produces_correct_values"]:
        assert O.donnelly_block_path(BH, c["input"], 3) == c["expect"], c
        n += 1
    assert n == 154


@pytest.mark.parametrize("bh", [3, 4, 5])
def test_block_step_equals_repeated_steps(oracle_mod, bh):
    """core.rs:704-766 (block height 4 there): from every block boundary reached by block
    steps, a block step to child c lands where bh single steps along c's bits do."""
    O = oracle_mod
    for a in range(1 << bh):
        for b in range(0, 1 << bh, max(1, (1 << bh) // 4)):
            base = [a, b]
            bits = [((x >> s) & 1) == 1 for x in base for s in range(bh - 1, -1, -1)]
            for c in range(1 << bh):
                block = O.donnelly_block_path(bh, base + [c], bh)
                rep = O.donnelly_path(bh, bits + [((c >> s) & 1) == 1 for s in range(bh - 1, -1, -1)])
                assert rep == (block, 0), (bh, base, c)


def test_layout_is_a_bijection(oracle_mod):
    """Distinct paths of up to 12 levels reach distinct slots, and every child slot is
    larger than its parent's (a tree layout)."""
    O = oracle_mod
    for bh in (3, 4):
        seen = {}
        for depth in range(0, 11):
            for code in range(1 << depth):
                path = [((code >> (depth - 1 - i)) & 1) == 1 for i in range(depth)]
                slot = O.donnelly_path(bh, path)[0]
                assert slot not in seen, (bh, path, seen.get(slot))
                seen[slot] = path
                if depth:
                    assert slot > O.donnelly_path(bh, path[:-1])[0]


def test_padding_known_answer(oracle_mod):
    """kd_tree/mod.rs:733-789: 100 points [x, 2x, 3x, 4x] (x = i / 100) in f32, bucket 32,
    Donnelly<4>: 4 leaves -> stem depth 2, padded to 4; max_stem_level 3; a query lands in
    one of the 4 leaves."""
    x = (np.arange(100) / 100).astype(np.float32)
    pts = np.stack([x, x * 2, x * 3, x * 4], 1).astype(np.float32)
    t = oracle_mod.Tree.build(pts, 32, block=4)
    assert t.max_stem_level == 3 and t.pad_levels == 2
    stems, ext, arena, msl = t.donnelly_layout(4)
    assert msl == 3 and len(ext) == 4
    assert sum(int(e[1]) for e in ext) == 100
    i, d = t.approx_nearest_one(np.array([[0.5, 1.0, 1.5, 2.0]], np.float32))
    leaves = t.leaves()
    owner = [li for li, (_, its) in enumerate(leaves) if int(i[0]) in its.tolist()]
    assert owner and owner[0] < 4


def _padded_cases():
    rng = np.random.default_rng(17)
    return [
        ("uniform-f64-k3-bh3", rng.random((20_000, 3)), 32, 3),
        ("uniform-f32-k4-bh4", rng.random((9_000, 4)).astype(np.float32), 16, 4),
        ("grid-f64-k2-bh5", rng.integers(0, 9, (6_000, 2)).astype(np.float64), 8, 5),
        ("dups-f64-k3-bh3", np.repeat(rng.random((2_000, 3)), 3, axis=0), 32, 3),
        ("uniform-f64-k3-bh6", rng.random((30_000, 3)), 8, 6),
        ("u16-k4-bh3", rng.integers(0, 65535, (8_000, 4)).astype(np.uint16), 32, 3),
        ("small-f64-k3-bh4", rng.random((70, 3)), 32, 4),
    ]


@pytest.mark.parametrize("name,pts,B,bh", _padded_cases(), ids=[c[0] for c in _padded_cases()])
def test_padded_tree_invariants_and_queries(oracle_mod, name, pts, B, bh):
    from test_oracle import _check_layout_invariants
    t = oracle_mod.Tree.build(pts, B, block=bh)
    D = t.max_stem_level + 1
    assert D % bh == 0 and 0 <= t.pad_levels < bh
    _check_layout_invariants(t, pts)
    stems = t.stems()
    # the padding levels hold no pivot on the path to the real root
    for level in range(t.pad_levels):
        s = stems[1 << level]
        assert (np.isinf(s) if s.dtype.kind == "f" else s == np.iinfo(s.dtype).max)
    rng = np.random.default_rng(3)
    if pts.dtype.kind == "u":
        q = rng.integers(0, 65535, (300, pts.shape[1])).astype(pts.dtype)
        r = 4.0e7
    else:
        q = rng.random((300, pts.shape[1])).astype(pts.dtype)
        r = 0.01
    q[:60] = pts[rng.integers(0, len(pts), 60)]
    for k in (1, 8, 20):
        for a, b in zip(t.nearest_n(q, k), t.brute_nearest_n(q, k)):
            assert_same_bits(a, b, f"{name} nearest_n({k})")
    for a, b in zip(t.within(q, r), t.brute_within(q, r)):
        assert_same_bits(a, b, f"{name} within")
    for a, b in zip(t.best_n_within(q, r, 5), t.brute_best_n_within(q, r, 5)):
        assert_same_bits(a, b, f"{name} best_n_within")


@pytest.mark.parametrize("name,pts,B,bh", _padded_cases(), ids=[c[0] for c in _padded_cases()])
def test_product_relayout_equals_oracle(ka, oracle_mod, name, pts, B, bh):
    t = oracle_mod.Tree.build(pts, B, block=bh)
    stems, ext, arena, msl = t.donnelly_layout(bh)
    lay = ka.HostLayout.from_donnelly(pts.dtype, pts.shape[1], B, len(pts), msl, stems, ext, arena, bh)
    assert lay.max_stem_level == t.max_stem_level
    assert_same_bits(lay.stems(), t.stems(), f"{name} stems")
    assert_same_bits(lay.extents(), t.extents(), f"{name} extents")
    a, b = lay.arena(), t.arena()
    assert_same_bits(a[: len(b)], b, f"{name} arena")
    assert not a[len(b):].any()


def test_product_relayout_refuses_bad_layouts(ka, oracle_mod):
    rng = np.random.default_rng(1)
    pts = rng.random((5000, 3))
    t = oracle_mod.Tree.build(pts, 32, block=3)
    stems, ext, arena, msl = t.donnelly_layout(3)
    with pytest.raises(ka.KiddoAmdError):  # depth 9 is not a multiple of 4
        ka.HostLayout.from_donnelly(np.float64, 3, 32, 5000, msl, stems, ext, arena, 4)
    with pytest.raises(ka.KiddoAmdError):  # not a power-of-two leaf count
        ka.HostLayout.from_donnelly(np.float64, 3, 32, 5000, msl, stems, ext[:-1], arena, 3)
    with pytest.raises(ka.KiddoAmdError):  # block heights below 3 have no Donnelly layout
        ka.HostLayout.from_donnelly(np.float64, 3, 32, 5000, msl, stems, ext, arena, 2)
    with pytest.raises(ka.KiddoAmdError):  # stems array too short for the real nodes
        ka.HostLayout.from_donnelly(np.float64, 3, 32, 5000, msl, stems[:16], ext, arena, 3)
    # a stemless tree passes through unchanged
    small = oracle_mod.Tree.build(pts[:20], 32, block=3)
    s2, e2, a2, m2 = small.donnelly_layout(3)
    lay = ka.HostLayout.from_donnelly(np.float64, 3, 32, 20, m2, s2, e2, a2, 3)
    assert lay.max_stem_level == -1 and lay.extents().tolist() == small.extents().tolist()
