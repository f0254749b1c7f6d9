"""CPU-only checks of the product library: every symbol declared in include/kiddo_amd.h
is exported, the host builder produces exactly the oracle's Kiddo layout (stems and leaf
extents byte for byte, the same entries in every leaf), and bad inputs are refused with
an error code rather than an abort. No GPU call is made here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from engines import assert_same_bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    with open(os.path.join(ROOT, "include", "kiddo_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(ka):
    lib = ka.lib()
    declared = _declared_symbols()
    assert len(declared) >= 14
    for name in declared:
        assert hasattr(lib, name), f"{name} declared but not exported"
        assert name in ka.SIGNATURES, f"{name} has no Python binding"


def test_error_reporting_without_gpu(ka):
    lib = ka.lib()
    h = C.c_void_p()
    pts = np.zeros((4, 3))
    rc = lib.kd_build_layout(1, 3, 0, 4, pts.ctypes.data_as(C.c_void_p), None, 1, C.byref(h))
    assert rc == -1
    assert b"bucket" in lib.kd_last_error()
    rc = lib.kd_build_layout(1, 7, 32, 4, pts.ctypes.data_as(C.c_void_p), None, 1, C.byref(h))
    assert rc == -2  # dims 7 unsupported
    nan = np.array([[0.0, np.nan, 0.0]])
    rc = lib.kd_build_layout(1, 3, 32, 1, nan.ctypes.data_as(C.c_void_p), None, 1, C.byref(h))
    assert rc == -1 and b"NaN" in lib.kd_last_error()
    sentinel = np.array([[1, 65535, 2]], np.uint16)
    rc = lib.kd_build_layout(2, 3, 32, 1, sentinel.ctypes.data_as(C.c_void_p), None, 1, C.byref(h))
    assert rc == -1 and b"sentinel" in lib.kd_last_error()
    same = np.zeros((1025, 3))
    rc = lib.kd_build_layout(1, 3, 32, 1025, same.ctypes.data_as(C.c_void_p), None, 1, C.byref(h))
    assert rc == -4  # more than 1024 equal points end in one leaf (construction.rs:526)
    # null-tree query calls fail cleanly
    assert lib.kd_nearest_n(None, None, 1, 4, float("inf"), 0, None, None, None, None) == -1
    # the bucket autotuners check their arguments before touching the device
    best = C.c_int32()
    b = np.array([0, 32], np.int32)
    dummy = C.c_void_p(8)
    rc = lib.kd_autotune_bucket(1, 3, 4, dummy, None, 0, dummy, 4, 4, float("inf"), 0,
                                b.ctypes.data_as(C.c_void_p), 2, C.byref(best), None, None)
    assert rc == -1 and b"bucket 0" in lib.kd_last_error()
    rc = lib.kd_autotune_bucket(1, 3, 4, dummy, None, 0, dummy, 4, 4, float("inf"), 0,
                                b.ctypes.data_as(C.c_void_p), 0, C.byref(best), None, None)
    assert rc == -1
    offs = np.array([0, 4], np.uint64)
    sizes = np.array([4], np.uint64)
    ptrs = (C.c_void_p * 1)(pts.ctypes.data)
    rc = lib.kd_forest_autotune_bucket(1, 3, 1, ptrs, sizes.ctypes.data_as(C.c_void_p), dummy,
                                       offs.ctypes.data_as(C.c_void_p), 65, float("inf"), 0,
                                       b.ctypes.data_as(C.c_void_p), 2, C.byref(best), None, None)
    assert rc == -1 and b"k must be" in lib.kd_last_error()


def _layout_equal(ka, O, pts, B, items=None, threads=4):
    lay = ka.HostLayout(pts, B, items, threads)
    ref = O.Tree.build(pts, B, items)
    assert lay.max_stem_level == ref.max_stem_level
    assert_same_bits(lay.stems(), ref.stems(), "stems")
    assert_same_bits(lay.extents(), ref.extents(), "leaf extents")
    assert lay.desc.leaf_bytes_len == ref.arena_len
    # same entries per leaf (the order inside a leaf is the builder's own)
    got = O.Tree.from_layout(pts.dtype, pts.shape[1], B, len(pts), lay.max_stem_level,
                             lay.stems(), lay.extents(), lay.arena())
    for (gp, gi), (rp, ri) in zip(got.leaves(), ref.leaves()):
        og, orf = np.argsort(gi), np.argsort(ri)
        assert_same_bits(gi[og], ri[orf], "leaf items")
        assert_same_bits(gp[og], rp[orf], "leaf points")
    return lay


@pytest.mark.parametrize("n,K,dtype,B,kind", [
    (5000, 3, np.float64, 32, "uniform"),
    (100000, 3, np.float64, 32, "uniform"),
    (70000, 3, np.float32, 32, "mixture"),
    (20000, 2, np.float64, 8, "grid"),
    (20000, 4, np.uint16, 32, "uniform"),
    (9, 2, np.float32, 2, "kat"),
    (33, 3, np.float64, 32, "uniform"),
    (31, 3, np.float64, 32, "uniform"),
    (3000, 3, np.float64, 32, "dups"),
    (1000, 3, np.float64, 32, "all_same"),
])
def test_host_builder_matches_oracle_layout(ka, oracle_mod, n, K, dtype, B, kind):
    rng = np.random.default_rng(n)
    if kind == "uniform":
        pts = rng.random((n, K))
        if np.dtype(dtype).kind == "u":
            pts = pts * 60000
    elif kind == "mixture":
        from kiddo_amd.datasets import gaussian_mixture
        pts = gaussian_mixture(n, 16, 0.01, K, seed=3)
    elif kind == "grid":
        pts = rng.integers(0, 9, (n, K))
    elif kind == "dups":
        base = rng.random((n // 5, K))
        pts = base[rng.integers(0, len(base), n)]
    elif kind == "all_same":
        pts = np.full((n, K), 0.5)
    else:
        pts = np.array([[3.0, 0.0], [1.0, 0.6], [1.0, 1.4], [3.0, 3.3], [3.0, 3.8],
                        [0.0, 1.8], [3.0, 1.5], [3.0, 2.7], [1.0, 3.3]])
    pts = np.ascontiguousarray(pts.astype(dtype))
    _layout_equal(ka, oracle_mod, pts, B)


def test_host_builder_explicit_items(ka, oracle_mod):
    rng = np.random.default_rng(5)
    pts = rng.random((4000, 3))
    items = rng.permutation(10_000_000)[:4000].astype(np.uint32)
    _layout_equal(ka, oracle_mod, pts, 32, items)


def test_oracle_queries_on_product_layout(ka, oracle_mod):
    """The oracle answers identically on the product builder's layout and its own."""
    O = oracle_mod
    rng = np.random.default_rng(9)
    pts = rng.random((20000, 3))
    lay = ka.HostLayout(pts, 32)
    a = O.Tree.from_layout(np.float64, 3, 32, len(pts), lay.max_stem_level, lay.stems(),
                           lay.extents(), lay.arena())
    b = O.Tree.build(pts, 32)
    q = rng.random((500, 3))
    for x, y in zip(a.nearest_n(q, 16), b.nearest_n(q, 16)):
        assert_same_bits(x, y)
