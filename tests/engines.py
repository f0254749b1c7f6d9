"""Adapters that give the oracle and the GPU engine one call shape, so the same cases
run through both. Everything returned is host numpy."""
from __future__ import annotations

import json
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NP_DT = {"f32": np.float32, "f64": np.float64, "u16": np.uint16, "u32": np.uint32}


def load_kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


class OracleEngine:
    name = "oracle"

    def __init__(self, O):
        self.O = O

    def build(self, points, bucket=32, items=None):
        return self.O.Tree.build(points, bucket, items)

    def nearest_one(self, t, q):
        return t.nearest_one(q)

    def approx_nearest_one(self, t, q):
        return t.approx_nearest_one(q)

    def nearest_n(self, t, q, k, radius=math.inf, exclusive=False, sorted=True):
        return t.nearest_n(q, k, radius, exclusive, sorted)

    def within(self, t, q, radius, exclusive=False, sorted=True):
        return t.within(q, radius, exclusive, sorted)

    def best_n_within(self, t, q, radius, n, exclusive=False):
        return t.best_n_within(q, radius, n, exclusive)

    def periodic_nearest_one(self, t, q, box):
        return t.periodic_nearest_one(q, box)

    def periodic_nearest_n(self, t, q, box, k, radius=math.inf, exclusive=False, sorted=True):
        return t.periodic_nearest_n(q, box, k, radius, exclusive, sorted)

    def periodic_within(self, t, q, box, radius, exclusive=False, sorted=True):
        return t.periodic_within(q, box, radius, exclusive, sorted)


class GpuEngine:
    name = "gpu"

    def __init__(self, ka, flags=0, shadow=False):
        self.ka = ka
        self.flags = flags
        self.shadow = shadow

    def build(self, points, bucket=32, items=None):
        if items is None:
            return self.ka.ImmutableKdTree.new_from_slice(points, bucket, shadow_f32=self.shadow)
        return self.ka.ImmutableKdTree.new_from_entries(points, items, bucket, shadow_f32=self.shadow)

    @staticmethod
    def _host(x):
        return x.cpu().numpy()

    def nearest_one(self, t, q):
        r = t.nearest_one_raw(np.asarray(q, t.a_dtype), self.flags & ~self.ka.KD_SCREEN_F32)
        return self._host(r.item), self._host(r.distance)

    def approx_nearest_one(self, t, q):
        r = t.nearest_one_raw(np.asarray(q, t.a_dtype), self.flags & ~self.ka.KD_SCREEN_F32, approx=True)
        return self._host(r.item), self._host(r.distance)

    def nearest_n(self, t, q, k, radius=math.inf, exclusive=False, sorted=True):
        f = self.flags | (self.ka.KD_EXCLUSIVE if exclusive else 0)
        if not sorted:
            f |= self.ka.KD_UNSORTED
        r = t.nearest_n_raw(np.asarray(q, t.a_dtype), k, radius, f)
        return self._host(r.items), self._host(r.distances), self._host(r.counts)

    def within(self, t, q, radius, exclusive=False, sorted=True):
        f = (self.flags & ~self.ka.KD_SCREEN_F32) | (self.ka.KD_EXCLUSIVE if exclusive else 0)
        if not sorted:
            f |= self.ka.KD_UNSORTED
        r = t.within_raw(np.asarray(q, t.a_dtype), radius, f)
        return (self._host(r.offsets).astype(np.uint64), self._host(r.items),
                self._host(r.distances))

    def best_n_within(self, t, q, radius, n, exclusive=False):
        f = (self.flags & ~self.ka.KD_SCREEN_F32) | (self.ka.KD_EXCLUSIVE if exclusive else 0)
        r = t.best_n_within_raw(np.asarray(q, t.a_dtype), radius, n, f)
        return self._host(r.items), self._host(r.distances), self._host(r.counts)

    def _pflags(self, exclusive=False, sorted=True):
        f = (self.flags & ~self.ka.KD_SCREEN_F32) | (self.ka.KD_EXCLUSIVE if exclusive else 0)
        return f if sorted else f | self.ka.KD_UNSORTED

    def periodic_nearest_one(self, t, q, box):
        r = t.periodic_nearest_one_raw(np.asarray(q, t.a_dtype), box, self._pflags())
        return self._host(r.item), self._host(r.distance)

    def periodic_nearest_n(self, t, q, box, k, radius=math.inf, exclusive=False, sorted=True):
        r = t.periodic_nearest_n_raw(np.asarray(q, t.a_dtype), box, k, radius,
                                     self._pflags(exclusive, sorted))
        return self._host(r.items), self._host(r.distances), self._host(r.counts)

    def periodic_within(self, t, q, box, radius, exclusive=False, sorted=True):
        r = t.periodic_within_raw(np.asarray(q, t.a_dtype), box, radius,
                                  self._pflags(exclusive, sorted))
        return (self._host(r.offsets).astype(np.uint64), self._host(r.items),
                self._host(r.distances))


def run_kat_case(engine, case):
    """Runs one reference known-answer case; asserts on mismatch."""
    dt = NP_DT[case["dtype"]]
    pts = np.array(case["points"], dtype=dt)
    items = np.array(case["items"], np.uint32) if "items" in case else None
    t = engine.build(pts, case.get("bucket", 32), items)
    for qc in case["queries"]:
        q = np.array([qc["q"]], dtype=dt)
        fam = qc["family"]
        if "box" in qc:
            _run_periodic_kat(engine, t, case, qc, q, np.array(qc["box"], dtype=dt))
            continue
        if fam == "nearest_one":
            it, d = engine.nearest_one(t, q)
            assert int(it[0]) == qc["item"], (case["name"], qc, it)
            if "dist_exact" in qc:
                assert float(d[0]) == qc["dist_exact"], (case["name"], float(d[0]))
            if "dist_abs_tol" in qc:
                assert abs(float(d[0]) - qc["dist"]) < qc["dist_abs_tol"]
            if "dist_rel_tol" in qc:
                assert abs(float(d[0]) - qc["dist"]) <= qc["dist_rel_tol"] * abs(qc["dist"])
        elif fam in ("nearest_n", "nearest_n_first", "nearest_n_within", "nearest_n_within_unsorted"):
            radius = qc.get("radius", math.inf)
            it, d, c = engine.nearest_n(t, q, qc["k"], radius, qc.get("exclusive", False),
                                        fam != "nearest_n_within_unsorted")
            got = [int(x) for x in it[0][: int(c[0])]]
            if fam == "nearest_n_first":
                assert got[0] == qc["item"], (case["name"], qc, got)
            elif fam == "nearest_n_within_unsorted":
                assert sorted(got) == qc["items_sorted"], (case["name"], qc, got)
            else:
                assert got == qc["items"], (case["name"], qc, got)
            if "dists_exact" in qc:
                assert [float(x) for x in d[0][: int(c[0])]] == qc["dists_exact"]
        elif fam == "within":
            off, it, d = engine.within(t, q, qc["radius"], qc["exclusive"], True)
            got = [int(x) for x in it[int(off[0]): int(off[1])]]
            assert got == qc["items"], (case["name"], qc, got)
        elif fam == "best_n_within":
            it, d, c = engine.best_n_within(t, q, qc["radius"], qc["n"], qc["exclusive"])
            got = [int(x) for x in it[0][: int(c[0])]]
            assert got == qc["items"], (case["name"], qc, got)
        else:
            raise AssertionError(f"unknown family {fam}")


def _run_periodic_kat(engine, t, case, qc, q, box):
    """Known answers under periodic boundary conditions (query/builder.rs tests)."""
    fam = qc["family"]
    radius = qc.get("radius", math.inf)
    if qc.get("radius_from_nearest_one"):
        _, d = engine.periodic_nearest_one(t, q, box)
        radius = float(d[0])
    if fam == "nearest_one":
        it, d = engine.periodic_nearest_one(t, q, box)
        assert int(it[0]) == qc["item"], (case["name"], qc, it)
        assert abs(float(d[0]) - qc["dist"]) < qc["dist_abs_tol"], (case["name"], float(d[0]))
    elif fam == "nearest_n_within_unsorted":
        it, d, c = engine.periodic_nearest_n(t, q, box, qc["k"], radius, qc.get("exclusive", False),
                                             False)
        got = sorted(int(x) for x in it[0][: int(c[0])])
        assert got == qc["items_sorted"], (case["name"], qc, got)
    elif fam == "within_unsorted":
        off, it, d = engine.periodic_within(t, q, box, radius, qc.get("exclusive", False), False)
        got = sorted(int(x) for x in it[int(off[0]): int(off[1])])
        assert got == qc["items_sorted"], (case["name"], qc, got)
    else:
        raise AssertionError(f"unknown periodic family {fam}")


def min_image_sqdist(points: np.ndarray, q: np.ndarray, box: np.ndarray, out_dtype) -> np.ndarray:
    """Squared distance under periodic boundaries, evaluated independently with numpy in
    the order Kiddo's periodic leaf scan uses: per axis delta = |q - p|, the smaller of
    delta and L - delta, then acc = acc + m*m from 0, each op rounded."""
    p = points.astype(out_dtype)
    qq = q.astype(out_dtype)
    L = box.astype(out_dtype)
    acc = np.zeros(np.broadcast_shapes(p.shape[:-1], qq.shape[:-1]), out_dtype)
    for d in range(p.shape[-1]):
        delta = np.abs((qq[..., d] - p[..., d]).astype(out_dtype))
        wrapped = (L[d] - delta).astype(out_dtype)
        m = np.where(wrapped < delta, wrapped, delta).astype(out_dtype)
        acc = (acc + (m * m).astype(out_dtype)).astype(out_dtype)
    return acc


def assert_same_bits(a, b, what=""):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if a.dtype.kind == "f":
        av = a.view(np.uint64 if a.dtype.itemsize == 8 else np.uint32)
        bv = b.view(np.uint64 if b.dtype.itemsize == 8 else np.uint32)
        bad = np.nonzero(av != bv)
    else:
        bad = np.nonzero(a != b)
    assert bad[0].size == 0, f"{what}: {bad[0].size} mismatches, first at {[x[:5] for x in bad]}: {a[bad][:5]} vs {b[bad][:5]}"


def scalar_order_sqdist(points: np.ndarray, q: np.ndarray, out_dtype) -> np.ndarray:
    """Kiddo's SquaredEuclidean in its scalar order, evaluated independently with numpy:
    acc = 0; acc = acc + (p_d - q_d) * (p_d - q_d) for d in 0..K, each op rounded."""
    p = points.astype(out_dtype)
    qq = q.astype(out_dtype)
    acc = np.zeros(p.shape[:-1], out_dtype)
    for d in range(p.shape[-1]):
        diff = (p[..., d] - qq[..., d]).astype(out_dtype)
        acc = (acc + (diff * diff).astype(out_dtype)).astype(out_dtype)
    return acc
