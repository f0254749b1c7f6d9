"""Tree construction on the GPU (kd_tree_build_device) against the host builder and the
oracle: identical stems (bits), extents and per-leaf entry sets, identical query results,
and the same argument checks (NaN, sentinel values, leaf overflow)."""
import math

import numpy as np
import pytest

from engines import assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


def _leaf_rows(lay, K, dtype):
    """Per leaf: the (item, coordinates) rows sorted, decoded from a Kiddo arena."""
    a = np.dtype(dtype)
    arena = lay.arena()
    out = []
    for off, ln in lay.extents():
        off, ln = int(off), int(ln)
        pts = np.zeros((ln, K), a)
        its = np.zeros(ln, np.uint32)
        base, p = 0, off
        while base < ln:
            rem = ln - base
            tl = next(w for w in (32, 8, 4, 2, 1) if rem >= w)
            for d in range(K):
                pts[base:base + tl, d] = np.frombuffer(arena[p:p + tl * a.itemsize].tobytes(), a)
                p += tl * a.itemsize
            its[base:base + tl] = np.frombuffer(arena[p:p + 4 * tl].tobytes(), np.uint32)
            p += 4 * tl
            base += tl
        rows = sorted(zip(its.tolist(), map(tuple, pts.view(f"u{a.itemsize}").tolist())))
        out.append(rows)
    return out


def _cases():
    rng = np.random.default_rng(9)
    yield "uniform-f64-k3", rng.random((100_000, 3)), 32, None
    base = rng.random((5000, 3))
    yield "dups-f64-k3", base[rng.integers(0, 5000, 60_000)], 32, None
    yield "grid-f32-k2", rng.integers(0, 9, (40_000, 2)).astype(np.float32), 8, None
    yield "palette-u16-k4", rng.integers(0, 65535, (65_536, 4), dtype=np.uint16), 32, None
    yield "u32-k3-items", rng.integers(0, 1 << 20, (30_000, 3)).astype(np.uint32), 16, \
        rng.permutation(30_000).astype(np.uint32)
    yield "all-equal-f64", np.full((800, 3), 0.25), 32, None
    yield "axis-degenerate-f32-k4", np.concatenate(
        [np.zeros((20_000, 1)), rng.random((20_000, 3))], 1).astype(np.float32), 32, None
    for n in (1, 31, 33, 1000):
        yield f"small-{n}", rng.random((n, 3)), 32, None


@pytest.mark.parametrize("name,pts,B,items", [pytest.param(*c, id=c[0]) for c in _cases()])
def test_device_build_matches_host(ka, oracle_mod, cuda, name, pts, B, items):
    _announce("build", name)
    K = pts.shape[1]
    t = ka.ImmutableKdTree.build_on_device(pts, B, items)
    lg = t.export_layout()
    lh = ka.HostLayout(pts, B, items)
    assert lg.max_stem_level == lh.max_stem_level
    assert lg.stems().tobytes() == lh.stems().tobytes(), "stems differ"
    assert_same_bits(lg.extents(), lh.extents(), "extents")
    assert len(lg.arena()) == len(lh.arena())
    assert _leaf_rows(lg, K, pts.dtype) == _leaf_rows(lh, K, pts.dtype), "leaf membership differs"
    # the oracle's own builder agrees on stems and extents too
    ot = oracle_mod.Tree.build(pts, B, items)
    assert ot.stems().tobytes() == lg.stems().tobytes()
    # queries on the device-built tree equal the oracle's
    q = np.random.default_rng(1).random((2000, K)).astype(pts.dtype) if pts.dtype.kind == "f" else \
        np.random.default_rng(1).integers(0, 1000, (2000, K)).astype(pts.dtype)
    r = t.nearest_n_raw(q, 8)
    oi, od, oc = ot.nearest_n(q, 8)
    assert_same_bits(r.items.cpu().numpy(), oi, "items")
    assert_same_bits(r.distances.cpu().numpy(), od, "dists")
    assert_same_bits(r.counts.cpu().numpy(), oc, "counts")


def test_device_build_shadow_and_screen(ka, cuda):
    pts = np.random.default_rng(4).random((200_000, 3))
    q = np.random.default_rng(5).random((20_000, 3))
    _announce("build shadow")
    t = ka.ImmutableKdTree.build_on_device(pts, 32, shadow_f32=True)
    th = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=True)
    a = t.nearest_n_raw(q, 16, math.inf, ka.KD_SCREEN_F32 | ka.KD_MORTON)
    b = th.nearest_n_raw(q, 16, math.inf, 0)
    assert cuda.equal(a.items, b.items)
    assert cuda.equal(a.distances.view(cuda.int64), b.distances.view(cuda.int64))


def test_device_build_argument_checks(ka, cuda):
    pts = np.random.default_rng(6).random((1000, 3))
    pts[17, 1] = np.nan
    with pytest.raises(ka.KiddoAmdError):
        ka.ImmutableKdTree.build_on_device(pts, 32)
    u = np.zeros((100, 3), np.uint16)
    u[5, 0] = 65535  # the no-pivot sentinel
    with pytest.raises(ka.KiddoAmdError):
        ka.ImmutableKdTree.build_on_device(u, 32)
    with pytest.raises(ka.KiddoAmdError):  # 1100 equal points cannot be split: one leaf > 1024
        ka.ImmutableKdTree.build_on_device(np.full((1100, 3), 0.5), 32)
    t = ka.ImmutableKdTree.build_on_device(np.zeros((0, 3)), 32)
    r = t.nearest_one_raw(np.full((2, 3), 0.5))
    assert r.item.cpu().tolist() == [0, 0]
