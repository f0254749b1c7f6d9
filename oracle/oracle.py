"""ctypes front end of the CPU oracle (liboracle.so, built from kiddo_oracle.cpp).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg use it as the checker. The kiddo_amd product never imports it.

Every function mirrors one family of Kiddo's query builder (src/kd_tree/query/builder.rs
:2008-2077) run as a batch over host numpy arrays. ``mode`` selects the collector:
``CANONICAL`` orders ties by lower item, ``FAITHFUL`` keeps Kiddo's own first-visit rules.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

CANONICAL = 0
FAITHFUL = 1

DTYPES = {np.dtype(np.float32): 0, np.dtype(np.float64): 1,
          np.dtype(np.uint16): 2, np.dtype(np.uint32): 3}
OUT_DTYPE = {0: np.float32, 1: np.float64, 2: np.float64, 3: np.float64}
A_DTYPE = {0: np.float32, 1: np.float64, 2: np.uint16, 3: np.uint32}

_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, u64, i32, f64 = C.c_void_p, C.c_uint64, C.c_int, C.c_double
        L.or_build.argtypes = [i32, i32, i32, u64, vp, vp, C.POINTER(vp)]
        L.or_build_blocked.argtypes = [i32, i32, i32, u64, vp, vp, i32, C.POINTER(vp)]
        L.or_pad_levels.argtypes = [vp]
        L.or_donnelly_path.argtypes = [i32, vp, u64, vp]
        L.or_donnelly_block_path.argtypes = [i32, vp, u64, i32, vp]
        L.or_donnelly_stems.argtypes = [vp, i32, vp, C.POINTER(u64)]
        L.or_from_layout.argtypes = [i32, i32, i32, u64, i32, vp, u64, vp, u64, vp, u64,
                                     C.POINTER(vp)]
        L.or_free.argtypes = [vp]
        L.or_update_pivot.argtypes = [i32, i32, vp, u64, vp, i32, u64]
        L.or_update_pivot.restype = C.c_int64
        L.or_info.argtypes = [vp, vp]
        L.or_stems.argtypes = [vp]
        L.or_stems.restype = vp
        L.or_extents.argtypes = [vp]
        L.or_extents.restype = vp
        L.or_arena.argtypes = [vp]
        L.or_arena.restype = vp
        L.or_nearest_one.argtypes = [vp, vp, u64, i32, vp, vp, i32, vp]
        L.or_approx_nearest_one.argtypes = [vp, vp, u64, i32, vp, vp, i32]
        L.or_nearest_n.argtypes = [vp, vp, u64, u64, f64, i32, i32, i32, vp, vp, vp, i32, vp]
        L.or_within_count.argtypes = [vp, vp, u64, f64, i32, vp, i32, vp]
        L.or_within_fill.argtypes = [vp, vp, u64, f64, i32, i32, i32, vp, vp, vp, i32]
        L.or_best_n_within.argtypes = [vp, vp, u64, f64, u64, i32, vp, vp, vp, i32, vp]
        L.or_periodic_nearest_one.argtypes = [vp, vp, u64, vp, i32, vp, vp, i32, vp]
        L.or_periodic_nearest_n.argtypes = [vp, vp, u64, vp, u64, f64, i32, i32, i32, vp, vp, vp,
                                            i32, vp]
        L.or_periodic_within_count.argtypes = [vp, vp, u64, vp, f64, i32, vp, i32, vp]
        L.or_periodic_within_fill.argtypes = [vp, vp, u64, vp, f64, i32, i32, i32, vp, vp, vp, i32]
        L.or_set_metric.argtypes = [vp, C.c_uint32]
        L.or_metric_dist.argtypes = [C.c_uint32, vp, vp, i32, vp]
        L.or_metric_rect.argtypes = [C.c_uint32, f64, vp, i32, i32, f64, vp]
        L.or_brute_nearest_n.argtypes = [vp, vp, u64, u64, f64, i32, vp, vp, vp, i32]
        L.or_brute_within.argtypes = [vp, vp, u64, f64, i32, vp, vp, vp, vp, i32]
        L.or_brute_best_n_within.argtypes = [vp, vp, u64, f64, u64, i32, vp, vp, vp, i32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with status {rc}")


def metric_dist(flags: int, a, b) -> float:
    """Distance of two f64 vectors under the metric the flags name (bits 8-19)."""
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = np.zeros(1)
    _check(lib().or_metric_dist(flags, _p(a), _p(b), len(a), _p(out)), "metric_dist")
    return float(out[0])


def metric_rect(flags: int, rd: float, off, dim: int, new_off: float) -> float:
    """Rect distance after replacing off[dim] by new_off (rect_dist_after_update)."""
    off = np.ascontiguousarray(off, np.float64)
    out = np.zeros(1)
    _check(lib().or_metric_rect(flags, float(rd), _p(off), len(off), dim, float(new_off), _p(out)),
           "metric_rect")
    return float(out[0])


def donnelly_path(block_height: int, path) -> tuple:
    """(stem index, minor level) reached from the root of a Donnelly<block_height> stem
    array by the left (False) / right (True) choices of `path` (donnelly/core.rs
    traverse)."""
    bits = np.ascontiguousarray([1 if b else 0 for b in path] or [0], np.uint8)
    out = np.zeros(2, np.uint32)
    _check(lib().or_donnelly_path(block_height, _p(bits), len(path), _p(out)), "donnelly_path")
    return int(out[0]), int(out[1])


def donnelly_block_path(block_height: int, children, block_size: int) -> int:
    """Stem index reached by block steps of `block_size` levels (traverse_block)."""
    ch = np.ascontiguousarray(list(children) or [0], np.uint8)
    out = np.zeros(1, np.uint32)
    _check(lib().or_donnelly_block_path(block_height, _p(ch), len(children), block_size, _p(out)),
           "donnelly_block_path")
    return int(out[0])


def update_pivot(points: np.ndarray, sort_index: np.ndarray, dim: int, init: int) -> int:
    """construction.rs:1089-1175 on one slice; sort_index is reordered in place."""
    pts = np.ascontiguousarray(points)
    assert sort_index.dtype == np.uint64 and sort_index.flags.c_contiguous
    return int(lib().or_update_pivot(DTYPES[pts.dtype], pts.shape[1], _p(pts), pts.shape[0],
                                     _p(sort_index), dim, init))


class Tree:
    """A Kiddo-layout ImmutableKdTree held by the oracle."""

    def __init__(self, handle, dtype_code, K, B):
        self._h = handle
        self.dtype_code = dtype_code
        self.K = K
        self.B = B
        info = np.zeros(6, np.int64)
        lib().or_info(self._h, _p(info))
        (self.size, self.max_stem_level, self.stems_len, self.leaf_count,
         self.arena_len, self.max_leaf_len) = (int(x) for x in info)

    # -- construction -------------------------------------------------------------
    @classmethod
    def build(cls, points: np.ndarray, bucket: int = 32, items: np.ndarray | None = None,
              block: int | None = None):
        """ImmutableKdTree::new_from_slice (items None) / new_from_entries (items given).
        block: the stem depth padded to a multiple of `block` levels, as Kiddo builds a tree
        with a block-based stem strategy (Donnelly<block>); held in Eytzinger order."""
        pts = np.ascontiguousarray(points)
        code = DTYPES[pts.dtype]
        n, K = pts.shape
        it = None if items is None else np.ascontiguousarray(items, dtype=np.uint32)
        h = C.c_void_p()
        if block is None:
            _check(lib().or_build(code, K, bucket, n, _p(pts), _p(it), C.byref(h)), "build")
        else:
            _check(lib().or_build_blocked(code, K, bucket, n, _p(pts), _p(it), block, C.byref(h)),
                   "build_blocked")
        return cls(h, code, K, bucket)

    @property
    def pad_levels(self) -> int:
        return int(lib().or_pad_levels(self._h))

    def donnelly_layout(self, block_height: int):
        """The tree as Kiddo holds a KdTree<_, _, Donnelly<block_height>, ...>: stems in
        Donnelly order, the real leaves' extents (the empty ones under the padding levels
        dropped), the arena and max_stem_level. Needs a tree built with block=block_height."""
        n = C.c_uint64()
        _check(lib().or_donnelly_stems(self._h, block_height, None, C.byref(n)), "donnelly_stems")
        stems = np.zeros(n.value, A_DTYPE[self.dtype_code])
        if n.value:
            _check(lib().or_donnelly_stems(self._h, block_height, _p(stems), C.byref(n)), "donnelly_stems")
        ext = self.extents()
        if self.max_stem_level >= 0:
            ext = ext[: 1 << (self.max_stem_level + 1 - self.pad_levels)]
        return stems, np.ascontiguousarray(ext), self.arena(), self.max_stem_level

    @classmethod
    def from_layout(cls, dtype, K, B, n, max_stem_level, stems, extents, arena):
        code = DTYPES[np.dtype(dtype)]
        stems = np.ascontiguousarray(stems)
        extents = np.ascontiguousarray(extents, dtype=np.uint64)
        arena = np.ascontiguousarray(arena, dtype=np.uint8)
        h = C.c_void_p()
        _check(lib().or_from_layout(code, K, B, n, max_stem_level, _p(stems), stems.size,
                                    _p(extents), extents.size // 2, _p(arena), arena.size,
                                    C.byref(h)), "from_layout")
        return cls(h, code, K, B)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().or_free(self._h)
            self._h = None

    # -- layout -------------------------------------------------------------------
    def stems(self) -> np.ndarray:
        a = A_DTYPE[self.dtype_code]
        buf = (C.c_uint8 * (self.stems_len * np.dtype(a).itemsize)).from_address(
            lib().or_stems(self._h)) if self.stems_len else b""
        return np.frombuffer(bytes(buf), dtype=a).copy()

    def extents(self) -> np.ndarray:
        if self.leaf_count == 0:
            return np.zeros((0, 2), np.uint64)
        buf = (C.c_uint64 * (2 * self.leaf_count)).from_address(lib().or_extents(self._h))
        return np.frombuffer(buf, dtype=np.uint64).reshape(-1, 2).copy()

    def arena(self) -> np.ndarray:
        if self.arena_len == 0:
            return np.zeros(0, np.uint8)
        buf = (C.c_uint8 * self.arena_len).from_address(lib().or_arena(self._h))
        return np.frombuffer(buf, dtype=np.uint8).copy()

    def leaf_lens(self) -> np.ndarray:
        return self.extents()[:, 1].astype(np.int64)

    def leaves(self):
        """Decode the arena into per-leaf (points[len, K], items[len])."""
        a = np.dtype(A_DTYPE[self.dtype_code])
        arena = self.arena()
        out = []
        for off, ln in self.extents():
            off, ln = int(off), int(ln)
            pts = np.zeros((ln, self.K), a)
            its = np.zeros(ln, np.uint32)
            base = 0
            p = off
            while base < ln:
                rem = ln - base
                tl = next(w for w in (32, 8, 4, 2, 1) if rem >= w)
                for d in range(self.K):
                    pts[base:base + tl, d] = np.frombuffer(arena[p:p + tl * a.itemsize], a)
                    p += tl * a.itemsize
                its[base:base + tl] = np.frombuffer(arena[p:p + 4 * tl], np.uint32)
                p += 4 * tl
                base += tl
            out.append((pts, its))
        return out

    # -- queries ------------------------------------------------------------------
    def _q(self, q):
        q = np.ascontiguousarray(q, dtype=A_DTYPE[self.dtype_code])
        assert q.ndim == 2 and q.shape[1] == self.K
        return q

    def _out(self, *shape):
        return np.zeros(shape, OUT_DTYPE[self.dtype_code])

    def nearest_one(self, q, mode=CANONICAL, threads=1, stats=False):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros(m, np.uint32)
        d = self._out(m)
        st = np.zeros(3, np.uint64)
        _check(lib().or_nearest_one(self._h, _p(q), m, mode, _p(items), _p(d), threads,
                                    _p(st)), "nearest_one")
        return (items, d, st) if stats else (items, d)

    def approx_nearest_one(self, q, mode=CANONICAL, threads=1):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros(m, np.uint32)
        d = self._out(m)
        _check(lib().or_approx_nearest_one(self._h, _p(q), m, mode, _p(items), _p(d), threads),
               "approx_nearest_one")
        return items, d

    def nearest_n(self, q, k, radius=np.inf, exclusive=False, sorted=True, mode=CANONICAL,
                  threads=1, stats=False):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros((m, k), np.uint32)
        d = self._out(m, k)
        counts = np.zeros(m, np.uint32)
        st = np.zeros(3, np.uint64)
        _check(lib().or_nearest_n(self._h, _p(q), m, k, float(radius), int(exclusive),
                                  int(sorted), mode, _p(items), _p(d), _p(counts), threads,
                                  _p(st)), "nearest_n")
        return (items, d, counts, st) if stats else (items, d, counts)

    def within(self, q, radius, exclusive=False, sorted=True, mode=CANONICAL, threads=1,
               stats=False):
        """Returns CSR (offsets[m+1], items, dists)."""
        q = self._q(q)
        m = q.shape[0]
        counts = np.zeros(m, np.uint64)
        st = np.zeros(3, np.uint64)
        _check(lib().or_within_count(self._h, _p(q), m, float(radius), int(exclusive),
                                     _p(counts), threads, _p(st)), "within_count")
        offsets = np.zeros(m + 1, np.uint64)
        np.cumsum(counts, out=offsets[1:])
        total = int(offsets[-1])
        items = np.zeros(total, np.uint32)
        d = self._out(total)
        _check(lib().or_within_fill(self._h, _p(q), m, float(radius), int(exclusive),
                                    int(sorted), mode, _p(offsets), _p(items), _p(d), threads),
               "within_fill")
        return (offsets, items, d, st) if stats else (offsets, items, d)

    def within_count(self, q, radius, exclusive=False, threads=1):
        """Per-query result counts of within(radius) (the count pass alone)."""
        q = self._q(q)
        m = q.shape[0]
        counts = np.zeros(m, np.uint64)
        st = np.zeros(3, np.uint64)
        _check(lib().or_within_count(self._h, _p(q), m, float(radius), int(exclusive),
                                     _p(counts), threads, _p(st)), "within_count")
        return counts

    def set_metric(self, flags: int):
        """Metric of the following queries (flags bits 8-19, as the GPU engine's)."""
        _check(lib().or_set_metric(self._h, int(flags)), "set_metric")
        return self

    # -- periodic boundary conditions (float trees; box: K positive lengths) -------------
    def _box(self, box):
        b = np.ascontiguousarray(box, dtype=A_DTYPE[self.dtype_code])
        assert b.shape == (self.K,)
        return b

    def periodic_nearest_one(self, q, box, mode=CANONICAL, threads=1, stats=False):
        q, b = self._q(q), self._box(box)
        m = q.shape[0]
        items = np.zeros(m, np.uint32)
        d = self._out(m)
        st = np.zeros(3, np.uint64)
        _check(lib().or_periodic_nearest_one(self._h, _p(q), m, _p(b), mode, _p(items), _p(d),
                                             threads, _p(st)), "periodic_nearest_one")
        return (items, d, st) if stats else (items, d)

    def periodic_nearest_n(self, q, box, k, radius=np.inf, exclusive=False, sorted=True,
                           mode=CANONICAL, threads=1, stats=False):
        q, b = self._q(q), self._box(box)
        m = q.shape[0]
        items = np.zeros((m, k), np.uint32)
        d = self._out(m, k)
        counts = np.zeros(m, np.uint32)
        st = np.zeros(3, np.uint64)
        _check(lib().or_periodic_nearest_n(self._h, _p(q), m, _p(b), k, float(radius),
                                           int(exclusive), int(sorted), mode, _p(items), _p(d),
                                           _p(counts), threads, _p(st)), "periodic_nearest_n")
        return (items, d, counts, st) if stats else (items, d, counts)

    def periodic_within(self, q, box, radius, exclusive=False, sorted=True, mode=CANONICAL,
                        threads=1, stats=False):
        """Returns CSR (offsets[m+1], items, dists)."""
        q, b = self._q(q), self._box(box)
        m = q.shape[0]
        counts = np.zeros(m, np.uint64)
        st = np.zeros(3, np.uint64)
        _check(lib().or_periodic_within_count(self._h, _p(q), m, _p(b), float(radius),
                                              int(exclusive), _p(counts), threads, _p(st)),
               "periodic_within_count")
        offsets = np.zeros(m + 1, np.uint64)
        np.cumsum(counts, out=offsets[1:])
        total = int(offsets[-1])
        items = np.zeros(total, np.uint32)
        d = self._out(total)
        _check(lib().or_periodic_within_fill(self._h, _p(q), m, _p(b), float(radius),
                                             int(exclusive), int(sorted), mode, _p(offsets),
                                             _p(items), _p(d), threads), "periodic_within_fill")
        return (offsets, items, d, st) if stats else (offsets, items, d)

    def best_n_within(self, q, radius, n, exclusive=False, threads=1, stats=False):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros((m, n), np.uint32)
        d = self._out(m, n)
        counts = np.zeros(m, np.uint32)
        st = np.zeros(3, np.uint64)
        _check(lib().or_best_n_within(self._h, _p(q), m, float(radius), n, int(exclusive),
                                      _p(items), _p(d), _p(counts), threads, _p(st)),
               "best_n_within")
        return (items, d, counts, st) if stats else (items, d, counts)

    # -- brute force (canonical) ----------------------------------------------------
    def brute_nearest_n(self, q, k, radius=np.inf, exclusive=False, threads=1):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros((m, k), np.uint32)
        d = self._out(m, k)
        counts = np.zeros(m, np.uint32)
        _check(lib().or_brute_nearest_n(self._h, _p(q), m, k, float(radius), int(exclusive),
                                        _p(items), _p(d), _p(counts), threads), "brute_nn")
        return items, d, counts

    def brute_within(self, q, radius, exclusive=False, threads=1):
        q = self._q(q)
        m = q.shape[0]
        counts = np.zeros(m, np.uint64)
        _check(lib().or_brute_within(self._h, _p(q), m, float(radius), int(exclusive), None,
                                     _p(counts), None, None, threads), "brute_within")
        offsets = np.zeros(m + 1, np.uint64)
        np.cumsum(counts, out=offsets[1:])
        total = int(offsets[-1])
        items = np.zeros(total, np.uint32)
        d = self._out(total)
        _check(lib().or_brute_within(self._h, _p(q), m, float(radius), int(exclusive),
                                     _p(offsets), None, _p(items), _p(d), threads),
               "brute_within")
        return offsets, items, d

    def brute_best_n_within(self, q, radius, n, exclusive=False, threads=1):
        q = self._q(q)
        m = q.shape[0]
        items = np.zeros((m, n), np.uint32)
        d = self._out(m, n)
        counts = np.zeros(m, np.uint32)
        _check(lib().or_brute_best_n_within(self._h, _p(q), m, float(radius), n,
                                            int(exclusive), _p(items), _p(d), _p(counts),
                                            threads), "brute_best_n")
        return items, d, counts
