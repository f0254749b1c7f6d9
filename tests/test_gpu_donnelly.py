"""GPU parity on trees built with Donnelly<BH> stems (kiddo src/stem_strategy/donnelly/):
the oracle builds the padded tree and exports it in Donnelly order (what a Rust shim reads
from a KdTree<_, _, Donnelly<BH>, ...>), kd_layout_from_donnelly re-lays it out, and every
family's results on the GPU equal the oracle's on the same tree, bit for bit, with the same
leaf visits."""
import math

import numpy as np
import pytest

from engines import GpuEngine, assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


def _cmp(got, ref, what):
    for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
        assert_same_bits(a, b, f"{what} {nm}")


CASES = [
    ("uniform-f64-k3-bh3", lambda r: r.random((60_000, 3)), 32, 3),
    ("dups-f64-k3-bh4", lambda r: np.repeat(r.random((10_000, 3)), 3, axis=0), 32, 4),
    ("uniform-f32-k4-bh4", lambda r: r.random((40_000, 4)).astype(np.float32), 16, 4),
    ("grid-f32-k2-bh5", lambda r: r.integers(0, 30, (30_000, 2)).astype(np.float32), 8, 5),
    ("u16-k4-bh3", lambda r: r.integers(0, 65535, (30_000, 4)).astype(np.uint16), 32, 3),
]


@pytest.mark.parametrize("name,make,B,bh", CASES, ids=[c[0] for c in CASES])
def test_donnelly_tree_families(ka, oracle_mod, cuda, name, make, B, bh):
    rng = np.random.default_rng(23)
    pts = make(rng)
    ot = oracle_mod.Tree.build(pts, B, block=bh)
    stems, ext, arena, msl = ot.donnelly_layout(bh)
    shadow = pts.dtype == np.float64 and pts.shape[1] <= 3
    gt = ka.ImmutableKdTree.from_donnelly(pts.dtype, pts.shape[1], B, len(pts), msl, stems, ext, arena, bh,
                                          shadow_f32=shadow)
    assert gt.info.max_stem_level == ot.max_stem_level
    assert_same_bits(gt.export_layout().stems(), ot.stems(), "device stems")
    if pts.dtype.kind == "u":
        q = rng.integers(0, 65535, (3000, pts.shape[1])).astype(pts.dtype)
        r = 4.0e7
    else:
        lo, hi = float(pts.min()), float(pts.max())
        q = (lo + rng.random((3000, pts.shape[1])) * (hi - lo)).astype(pts.dtype)
        r = 2.0 if hi > 1.5 else 0.004
    q[:500] = pts[rng.integers(0, len(pts), 500)]
    for flags in (0, ka.KD_MORTON):
        G = GpuEngine(ka, flags=flags)
        _announce("donnelly", name, flags)
        gi, gd = G.nearest_one(gt, q)
        oi, od = ot.nearest_one(q)
        assert_same_bits(gi, oi, "nearest_one items")
        assert_same_bits(gd, od, "nearest_one dists")
        for k in (1, 8, 16):
            ref = ot.nearest_n(q, k)
            for variant in (1, 2):
                gt.tune(variant, 0, -1)
                _cmp(G.nearest_n(gt, q, k), ref, f"nearest_n({k}) variant {variant}")
            gt.tune(0, 0, -1)
            if shadow:
                rs = gt.nearest_n_raw(q, k, flags=flags | ka.KD_SCREEN_F32)
                _cmp((rs.items.cpu().numpy(), rs.distances.cpu().numpy(), rs.counts.cpu().numpy()), ref,
                     f"screened nearest_n({k})")
        _cmp(G.nearest_n(gt, q, 100), ot.nearest_n(q, 100), "nearest_n(100)")
        _cmp(G.nearest_n(gt, q, 8, r, True), ot.nearest_n(q, 8, r, True), "nearest_n within exclusive")
        go, gi, gd = G.within(gt, q, r)
        oo, oi, od = ot.within(q, r)
        assert_same_bits(go, oo, "within offsets")
        assert_same_bits(gi, oi, "within items")
        assert_same_bits(gd, od, "within dists")
        _cmp(G.best_n_within(gt, q, r, 5), ot.best_n_within(q, r, 5), "best_n_within")
        if pts.dtype.kind == "f":
            gi, gd = G.approx_nearest_one(gt, q)
            oi, od = ot.approx_nearest_one(q)
            assert_same_bits(gi, oi, "approx items")
            assert_same_bits(gd, od, "approx dists")
    # the same leaves and entries as the oracle's DFS on this tree
    _, _, _, ost = ot.nearest_n(q, 16, stats=True)
    gt.tune(1, 0, -1)
    gt.nearest_n_raw(q, 16)
    st = gt.query_stats()
    assert (st["leaves"], st["points"]) == (int(ost[0]), int(ost[1])), (st, ost)
