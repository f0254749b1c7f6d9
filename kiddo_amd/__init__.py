"""kiddo_amd — MI355X batch query engine for Kiddo's ImmutableKdTree.

Python mirror of Kiddo's query builder (kiddo src/kd_tree/query/builder.rs:2318-2331
``KdTree::query`` and the families at :2008-2077) over the C ABI in include/kiddo_amd.h.
Where Kiddo answers one query per ``execute()``, ``tree.query(queries)`` here takes a
whole batch (an m x K array) and runs it on the GPU:

    tree = ImmutableKdTree.new_from_slice(points)                 # host build + upload
    r = tree.query(q).nearest_n(16).execute()                     # r.items, r.distances
    r = tree.query(q).nearest_n(8).within(0.01).exclusive_boundaries().execute()
    r = tree.query(q).nearest_one().execute()                     # r.item, r.distance
    r = tree.query(q).nearest_one().approx().execute()
    r = tree.query(q).within(0.01).unsorted().execute()           # CSR: r.offsets
    r = tree.query(q).best_n_within(0.01, 5).execute()

Radii are squared distances (SquaredEuclidean output units, like Kiddo's ``within``).
torch provides device memory and streams only; every query runs in the HIP kernels of
libkiddo_amd.so, and importing this module fails when that library is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkiddo_amd.so")

KD_OK = 0
KD_F32, KD_F64, KD_U16, KD_U32 = 0, 1, 2, 3
KD_EXCLUSIVE = 1 << 0
KD_UNSORTED = 1 << 1
KD_MORTON = 1 << 2
KD_SCREEN_F32 = 1 << 3
KD_GRAPH = 1 << 4
KD_TREE_SHADOW_F32 = 1 << 0
# batch limit of kd_nearest_n_noalloc under KD_MORTON (include/kiddo_amd.h)
KD_NOALLOC_MAX_QUERIES = 1 << 20
# distance metrics (flag bits 8-19, include/kiddo_amd.h KD_METRIC_*)
KD_METRIC_SQEUCLIDEAN = 0
KD_METRIC_MANHATTAN = 1 << 8
KD_METRIC_CHEBYSHEV = 2 << 8
KD_METRIC_MASK = 0x000FFF00


def KD_METRIC_MINKOWSKI(p: int) -> int:
    return (3 << 8) | (int(p) << 12)


def metric_flags(metric) -> int:
    """'sqeuclidean' | 'manhattan' | 'chebyshev' | ('minkowski', P) | flag value -> flag bits."""
    if metric is None:
        return 0
    if isinstance(metric, int):
        return metric
    if isinstance(metric, tuple) and metric[0] == "minkowski":
        return KD_METRIC_MINKOWSKI(metric[1])
    return {"sqeuclidean": 0, "squared_euclidean": 0, "manhattan": KD_METRIC_MANHATTAN,
            "chebyshev": KD_METRIC_CHEBYSHEV}[metric]

_NP_CODE = {np.dtype(np.float32): KD_F32, np.dtype(np.float64): KD_F64,
            np.dtype(np.uint16): KD_U16, np.dtype(np.uint32): KD_U32}
_A_NP = {KD_F32: np.float32, KD_F64: np.float64, KD_U16: np.uint16, KD_U32: np.uint32}
_O_NP = {KD_F32: np.float32, KD_F64: np.float64, KD_U16: np.float64, KD_U32: np.float64}


class KiddoAmdError(RuntimeError):
    pass


class TreeDesc(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("dims", C.c_int32), ("bucket", C.c_int32),
                ("max_stem_level", C.c_int32), ("size", C.c_uint64),
                ("stems_len", C.c_uint64), ("stems", C.c_void_p),
                ("leaf_count", C.c_uint64), ("leaf_extents", C.c_void_p),
                ("leaf_bytes_len", C.c_uint64), ("leaf_bytes", C.c_void_p)]


class TreeInfo(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("dims", C.c_int32), ("max_stem_level", C.c_int32),
                ("has_shadow_f32", C.c_int32), ("size", C.c_uint64),
                ("leaf_count", C.c_uint64), ("max_leaf_len", C.c_uint64),
                ("device_bytes", C.c_uint64), ("bbox_min", C.c_double * 4),
                ("bbox_max", C.c_double * 4)]


# The exported entry points and their C signatures (include/kiddo_amd.h).
_vp, _u64, _u32, _i32, _f64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32, C.c_double
SIGNATURES = {
    "kd_last_error": ([], C.c_char_p),
    "kd_version": ([], C.c_int),
    "kd_build_layout": ([_i32, _i32, _i32, _u64, _vp, _vp, _i32, C.POINTER(_vp)], C.c_int),
    "kd_layout_desc": ([_vp, C.POINTER(TreeDesc)], C.c_int),
    "kd_layout_free": ([_vp], None),
    "kd_layout_from_donnelly": ([C.POINTER(TreeDesc), _i32, C.POINTER(_vp)], C.c_int),
    "kd_tree_create": ([C.POINTER(TreeDesc), _u32, C.POINTER(_vp)], C.c_int),
    "kd_tree_build_device": ([_i32, _i32, _i32, _u64, _vp, _vp, _u32, C.POINTER(_vp), _vp], C.c_int),
    "kd_tree_export_layout": ([_vp, C.POINTER(_vp)], C.c_int),
    "kd_tree_destroy": ([_vp], C.c_int),
    "kd_tree_get_info": ([_vp, C.POINTER(TreeInfo)], C.c_int),
    "kd_tree_tune": ([_vp, _i32, _i32, _i32], C.c_int),
    "kd_tree_tune2": ([_vp, _i32, _i32, _i32, _i32], C.c_int),
    "kd_tree_tune_within": ([_vp, _i32, _i32], C.c_int),
    "kd_tree_tune_best_n_within": ([_vp, _i32], C.c_int),
    "kd_autotune_within": ([_vp, _vp, _u64, _f64, _u32, _vp, _vp], C.c_int),
    "kd_autotune_nearest_n": ([_vp, _vp, _u64, _u32, _f64, _u32, _vp, _vp], C.c_int),
    "kd_autotune_bucket": ([_i32, _i32, _u64, _vp, _vp, _u32, _vp, _u64, _u32, _f64, _u32, _vp, _u32, _vp, _vp,
                            _vp], C.c_int),
    "kd_forest_autotune_bucket": ([_i32, _i32, _u32, _vp, _vp, _vp, _vp, _u32, _f64, _u32, _vp, _u32, _vp, _vp,
                                   _vp], C.c_int),
    "kd_nearest_one": ([_vp, _vp, _u64, _u32, _vp, _vp, _vp], C.c_int),
    "kd_approx_nearest_one": ([_vp, _vp, _u64, _u32, _vp, _vp, _vp], C.c_int),
    "kd_nearest_n": ([_vp, _vp, _u64, _u32, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_nearest_n_noalloc": ([_vp, _vp, _u64, _u32, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_within_count": ([_vp, _vp, _u64, _f64, _u32, _vp, _vp], C.c_int),
    "kd_morton_order": ([_vp, _vp, _u64, _vp, _vp], C.c_int),
    "kd_within_fill": ([_vp, _vp, _u64, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_best_n_within": ([_vp, _vp, _u64, _f64, _u32, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_periodic_nearest_one": ([_vp, _vp, _u64, _vp, _u32, _vp, _vp, _vp], C.c_int),
    "kd_periodic_nearest_n": ([_vp, _vp, _u64, _vp, _u32, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_periodic_within_count": ([_vp, _vp, _u64, _vp, _f64, _u32, _vp, _vp], C.c_int),
    "kd_periodic_within_fill": ([_vp, _vp, _u64, _vp, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_query_stats": ([_vp, _vp], C.c_int),
    "kd_query_redos": ([_vp, _vp], C.c_int),
    "kd_forest_create": ([_vp, _u32, _u32, C.POINTER(_vp)], C.c_int),
    "kd_forest_destroy": ([_vp], C.c_int),
    "kd_forest_plan": ([_vp, _vp, _vp, _u32, _u32, _vp], C.c_int),
    "kd_forest_nearest_n": ([_vp, _vp, _u32, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_forest_within": ([_vp, _vp, _f64, _u32, _vp, _vp, _vp, _vp], C.c_int),
    "kd_forest_tune": ([_vp, _i32, _i32, _i32, _i32], C.c_int),
    "kd_forest_autotune_nearest_n": ([_vp, _vp, _u32, _f64, _u32, _vp, _vp, _vp, _vp, _vp], C.c_int),
    "kd_forest_set_timing": ([_vp, _i32], C.c_int),
    "kd_forest_last_kernel_ms": ([_vp, C.POINTER(C.c_float)], C.c_int),
    "kd_forest_query_stats": ([_vp, _vp], C.c_int),
    "kd_server_create": ([_vp, _u32, _u64, _u32, _u64, C.POINTER(_vp)], C.c_int),
    "kd_server_submit": ([_vp, _vp, C.POINTER(_u64)], C.c_int),
    "kd_server_wait": ([_vp, _u64, C.POINTER(C.c_double)], C.c_int),
    "kd_server_stats": ([_vp, _vp], C.c_int),
    "kd_server_set_part": ([_vp, _u64], C.c_int),
    "kd_server_destroy": ([_vp], C.c_int),
    "kd_tree_set_timing": ([_vp, _i32], C.c_int),
    "kd_last_kernel_ms": ([_vp, C.POINTER(C.c_float)], C.c_int),
}

_lib = None


def lib():
    """Loads libkiddo_amd.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KiddoAmdError(
                f"{LIB_PATH} is missing: build it with `python -m kiddo_amd._build`")
        # torch carries its own HIP runtime (sonames libamdhip64.so.7 and
        # libhsa-runtime64.so.1). Loading torch first makes the engine bind to that same
        # runtime by soname; otherwise a second runtime would be loaded and find no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != KD_OK:
        msg = lib().kd_last_error().decode(errors="replace")
        raise KiddoAmdError(f"{what} failed ({rc}): {msg}")


def _torch():
    import torch
    return torch


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _stream_ptr(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return C.c_void_p(stream.cuda_stream)


# ------------------------------------------------------------------ host layout ------

class HostLayout:
    """Kiddo-layout tree built on the host (ImmutableKdTree::new_from_slice /
    new_from_entries, construction.rs:278-303, :403-411)."""

    def __init__(self, points: np.ndarray = None, bucket: int = 32, items=None, threads: int = 0,
                 _handle=None):
        if _handle is not None:  # a layout exported from a device tree
            self._h = _handle
            self.desc = TreeDesc()
            _check(lib().kd_layout_desc(self._h, C.byref(self.desc)), "kd_layout_desc")
            self.dtype_code = self.desc.dtype
            self.n, self.K, self.bucket = self.desc.size, self.desc.dims, self.desc.bucket
            return
        pts = np.ascontiguousarray(points)
        if pts.ndim != 2:
            raise ValueError("points must be an (n, K) array")
        self.dtype_code = _NP_CODE[pts.dtype]
        self.n, self.K = pts.shape
        self.bucket = bucket
        it = None if items is None else np.ascontiguousarray(items, dtype=np.uint32)
        h = C.c_void_p()
        _check(lib().kd_build_layout(self.dtype_code, self.K, bucket, self.n, _np_ptr(pts),
                                     _np_ptr(it), threads, C.byref(h)), "kd_build_layout")
        self._h = h
        self.desc = TreeDesc()
        _check(lib().kd_layout_desc(self._h, C.byref(self.desc)), "kd_layout_desc")

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            try:
                _lib.kd_layout_free(self._h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @classmethod
    def from_donnelly(cls, dtype, K, bucket, size, max_stem_level, stems, extents, arena,
                      block_height: int):
        """Kiddo buffers of a Donnelly<block_height> tree -> the Eytzinger host layout
        (kd_layout_from_donnelly)."""
        d, keep = ImmutableKdTree._desc(dtype, K, bucket, size, max_stem_level, stems, extents, arena)
        h = C.c_void_p()
        _check(lib().kd_layout_from_donnelly(C.byref(d), block_height, C.byref(h)),
               "kd_layout_from_donnelly")
        del keep
        return cls(_handle=h)

    @property
    def max_stem_level(self):
        return self.desc.max_stem_level

    def stems(self) -> np.ndarray:
        a = np.dtype(_A_NP[self.dtype_code])
        n = self.desc.stems_len
        if n == 0:
            return np.zeros(0, a)
        buf = (C.c_uint8 * (n * a.itemsize)).from_address(self.desc.stems)
        return np.frombuffer(buf, a).copy()

    def extents(self) -> np.ndarray:
        n = self.desc.leaf_count
        if n == 0:
            return np.zeros((0, 2), np.uint64)
        buf = (C.c_uint64 * (2 * n)).from_address(self.desc.leaf_extents)
        return np.frombuffer(buf, np.uint64).reshape(-1, 2).copy()

    def arena(self) -> np.ndarray:
        n = self.desc.leaf_bytes_len
        if n == 0:
            return np.zeros(0, np.uint8)
        buf = (C.c_uint8 * n).from_address(self.desc.leaf_bytes)
        return np.frombuffer(buf, np.uint8).copy()


# ------------------------------------------------------------------ results ----------

@dataclass
class NearestOne:
    item: object       # (m,) uint32
    distance: object   # (m,) O


@dataclass
class NearestN:
    items: object      # (m, k) uint32, padded with 0xFFFFFFFF
    distances: object  # (m, k) O, padded with +inf
    counts: object     # (m,) uint32


@dataclass
class Within:
    offsets: object    # (m + 1,) int64 CSR offsets
    items: object
    distances: object


# ------------------------------------------------------------------ device tree ------

class ImmutableKdTree:
    """A Kiddo ImmutableKdTree<A, K> resident in HBM (kiddo src/lib.rs:270-271)."""

    def __init__(self, desc: TreeDesc, shadow_f32: bool = False, keepalive=None, _handle=None):
        if _handle is not None:
            h = _handle
        else:
            h = C.c_void_p()
            flags = KD_TREE_SHADOW_F32 if shadow_f32 else 0
            _check(lib().kd_tree_create(C.byref(desc), flags, C.byref(h)), "kd_tree_create")
        self._h = h
        self._keep = keepalive
        info = TreeInfo()
        _check(lib().kd_tree_get_info(self._h, C.byref(info)), "kd_tree_get_info")
        self.info = info
        self.dtype_code = info.dtype
        self.K = info.dims
        self.size = info.size

    @classmethod
    def new_from_slice(cls, points: np.ndarray, bucket: int = 32, shadow_f32: bool = False,
                       threads: int = 0):
        """ImmutableKdTree::new_from_slice: items are the source indices."""
        lay = HostLayout(points, bucket, None, threads)
        return cls(lay.desc, shadow_f32, keepalive=None)

    @classmethod
    def new_from_entries(cls, points: np.ndarray, items, bucket: int = 32,
                         shadow_f32: bool = False, threads: int = 0):
        lay = HostLayout(points, bucket, items, threads)
        return cls(lay.desc, shadow_f32, keepalive=None)

    @classmethod
    def build_on_device(cls, points, bucket: int = 32, items=None, shadow_f32: bool = False,
                        stream=None):
        """new_from_slice / new_from_entries built on the GPU (kd_tree_build_device) from a
        cuda tensor (or numpy array, copied to the device) of shape (n, K)."""
        torch = _torch()
        if isinstance(points, np.ndarray):
            points = torch.from_numpy(np.ascontiguousarray(points)).to("cuda")
        points = points.contiguous()
        code = {torch.float32: KD_F32, torch.float64: KD_F64, torch.uint16: KD_U16,
                torch.uint32: KD_U32}[points.dtype]
        n, K = points.shape
        it = None
        if items is not None:
            if isinstance(items, np.ndarray):
                items = torch.from_numpy(np.ascontiguousarray(items, dtype=np.uint32)).to("cuda")
            it = items.contiguous()
        h = C.c_void_p()
        _check(lib().kd_tree_build_device(code, K, bucket, n, points.data_ptr() if n else None,
                                          it.data_ptr() if it is not None and n else None,
                                          KD_TREE_SHADOW_F32 if shadow_f32 else 0, C.byref(h),
                                          _stream_ptr(stream)), "kd_tree_build_device")
        return cls(None, _handle=h)

    @classmethod
    def autotune_bucket(cls, points, queries, k: int, radius: float = math.inf, flags: int = 0,
                        buckets=(16, 32, 64, 128), items=None, shadow_f32: bool = False, stream=None):
        """Bucket size by measurement (kd_autotune_bucket): the tree is built on the GPU at
        each candidate bucket, its k-NN launch autotuned and nearest_n timed on the batch's
        tuning sample. Returns (best_bucket, {bucket: kernel ms})."""
        torch = _torch()
        if isinstance(points, np.ndarray):
            points = torch.from_numpy(np.ascontiguousarray(points)).to("cuda")
        points = points.contiguous()
        code = {torch.float32: KD_F32, torch.float64: KD_F64, torch.uint16: KD_U16,
                torch.uint32: KD_U32}[points.dtype]
        n, K = points.shape
        it = None
        if items is not None:
            if isinstance(items, np.ndarray):
                items = torch.from_numpy(np.ascontiguousarray(items, dtype=np.uint32)).to("cuda")
            it = items.contiguous()
        q = queries if not isinstance(queries, np.ndarray) else torch.from_numpy(np.ascontiguousarray(queries)).cuda()
        q = q.contiguous()
        b = np.ascontiguousarray(buckets, dtype=np.int32)
        ms = np.zeros(b.size, np.float32)
        best = C.c_int32()
        _check(lib().kd_autotune_bucket(code, K, n, points.data_ptr(), it.data_ptr() if it is not None else None,
                                        KD_TREE_SHADOW_F32 if shadow_f32 else 0, q.data_ptr(), q.shape[0], k,
                                        float(radius), flags, _np_ptr(b), b.size, C.byref(best), _np_ptr(ms),
                                        _stream_ptr(stream)), "kd_autotune_bucket")
        return int(best.value), {int(x): float(y) for x, y in zip(b, ms)}

    def export_layout(self) -> "HostLayout":
        """Kiddo's host layout (stems, extents, arena) of this device tree."""
        h = C.c_void_p()
        _check(lib().kd_tree_export_layout(self._h, C.byref(h)), "kd_tree_export_layout")
        return HostLayout(_handle=h)

    @staticmethod
    def _desc(dtype, K, bucket, size, max_stem_level, stems, extents, arena):
        stems = np.ascontiguousarray(stems)
        extents = np.ascontiguousarray(extents, dtype=np.uint64)
        arena = np.ascontiguousarray(arena, dtype=np.uint8)
        d = TreeDesc()
        d.dtype = _NP_CODE[np.dtype(dtype)]
        d.dims = K
        d.bucket = bucket
        d.max_stem_level = max_stem_level
        d.size = size
        d.stems_len = stems.size
        d.stems = stems.ctypes.data if stems.size else None
        d.leaf_count = extents.size // 2
        d.leaf_extents = extents.ctypes.data if extents.size else None
        d.leaf_bytes_len = arena.size
        d.leaf_bytes = arena.ctypes.data if arena.size else None
        return d, (stems, extents, arena)

    @classmethod
    def from_layout(cls, dtype, K, bucket, size, max_stem_level, stems, extents, arena,
                    shadow_f32: bool = False):
        """Uploads a tree given in Kiddo's own buffers (stems, leaf extents, leaf arena)."""
        d, keep = cls._desc(dtype, K, bucket, size, max_stem_level, stems, extents, arena)
        return cls(d, shadow_f32, keepalive=keep)

    @classmethod
    def from_donnelly(cls, dtype, K, bucket, size, max_stem_level, stems, extents, arena,
                      block_height: int, shadow_f32: bool = False):
        """Uploads a KdTree with Donnelly<block_height> stems (stems in Donnelly order, its
        real leaves' extents, the arena), re-laid out by kd_layout_from_donnelly."""
        lay = HostLayout.from_donnelly(dtype, K, bucket, size, max_stem_level, stems, extents, arena,
                                       block_height)
        return cls(lay.desc, shadow_f32, keepalive=lay)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kd_tree_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def a_dtype(self):
        return _A_NP[self.dtype_code]

    @property
    def out_dtype(self):
        return _O_NP[self.dtype_code]

    def query(self, queries, stream=None) -> "QueryBuilder":
        return QueryBuilder(self, queries, stream)

    def query_stats(self):
        """{leaves, points, stems, rescored} of the last query call on this tree."""
        out = np.zeros(4, np.uint64)
        _check(lib().kd_query_stats(self._h, _np_ptr(out)), "kd_query_stats")
        return dict(zip(("leaves", "points", "stems", "rescored"), (int(x) for x in out)))

    def query_redos(self) -> int:
        """Queries of the last call that the FP32-screened k-NN kernel handed to its exact
        FP64 redo pass (candidate list overflowed, e.g. by many exact ties)."""
        out = np.zeros(1, np.uint64)
        _check(lib().kd_query_redos(self._h, _np_ptr(out)), "kd_query_redos")
        return int(out[0])

    def tune(self, knn_variant: int = 0, group_size: int = 0, lds_levels: int = -1, waves: int = 0):
        """Kernel form for k-NN (0 auto, 1 thread/query, 2 group/query), LDS stem levels and
        the occupancy floor of the thread-per-query kernel."""
        _check(lib().kd_tree_tune2(self._h, knn_variant, group_size, lds_levels, waves),
               "kd_tree_tune2")

    def tune_within(self, count_group: int = 0, fill_group: int = 0):
        """Launch shapes of the within passes: 0 automatic, 1 thread per query, 4..32 lanes
        per query."""
        _check(lib().kd_tree_tune_within(self._h, count_group, fill_group), "kd_tree_tune_within")

    def tune_best_n_within(self, group: int = 0):
        """Launch shape of best_n_within: 0 automatic, 1 thread per query, 4..32 lanes per
        query (at least n; SquaredEuclidean)."""
        _check(lib().kd_tree_tune_best_n_within(self._h, group), "kd_tree_tune_best_n_within")

    def autotune_within(self, q, radius, flags=0, stream=None):
        """Times the within passes' launch shapes on a sample of q and keeps the fastest."""
        q = self._queries(q)
        chosen = np.zeros(2, np.int32)
        _check(lib().kd_autotune_within(self._h, q.data_ptr(), q.shape[0], float(radius), flags,
                                        _np_ptr(chosen), _stream_ptr(stream)), "kd_autotune_within")
        return dict(count_group=int(chosen[0]), fill_group=int(chosen[1]))

    def autotune_nearest_n(self, q, k, radius=math.inf, flags=0, stream=None):
        """Times the candidate launch shapes on a sample of q and keeps the fastest."""
        q = self._queries(q)
        chosen = np.zeros(4, np.int32)
        _check(lib().kd_autotune_nearest_n(self._h, q.data_ptr(), q.shape[0], k, float(radius),
                                           flags, _np_ptr(chosen), _stream_ptr(stream)),
               "kd_autotune_nearest_n")
        return dict(zip(("variant", "group_size", "lds_levels", "waves"), (int(x) for x in chosen)))

    def set_timing(self, enable: bool = True):
        _check(lib().kd_tree_set_timing(self._h, int(enable)), "kd_tree_set_timing")

    def last_kernel_ms(self) -> float:
        ms = C.c_float()
        _check(lib().kd_last_kernel_ms(self._h, C.byref(ms)), "kd_last_kernel_ms")
        return ms.value

    def screen_stats(self):
        st = self.query_stats()
        return st["points"], st["rescored"]

    # -- raw batch calls on device tensors (used by the builder, benches and tests) ------
    def _queries(self, q):
        torch = _torch()
        tdt = {np.float32: torch.float32, np.float64: torch.float64,
               np.uint16: torch.uint16, np.uint32: torch.uint32}[self.a_dtype]
        if isinstance(q, np.ndarray):
            q = torch.from_numpy(np.ascontiguousarray(q, dtype=self.a_dtype)).to("cuda")
        if q.dtype != tdt or q.dim() != 2 or q.shape[1] != self.K or not q.is_cuda:
            raise ValueError(f"queries must be a cuda ({tdt}) tensor of shape (m, {self.K})")
        return q.contiguous()

    def _out(self, *shape, dtype=None):
        torch = _torch()
        dt = dtype or {np.float32: torch.float32, np.float64: torch.float64}[self.out_dtype]
        return torch.empty(shape, dtype=dt, device="cuda")

    def _check_out(self, out, m, k):
        """Caller-provided result buffers must be contiguous cuda tensors of exactly the
        shapes and dtypes the kernel writes (it writes m x k items and distances, m counts)."""
        torch = _torch()
        odt = {np.float32: torch.float32, np.float64: torch.float64}[self.out_dtype]
        for name, t, shape, dt in (("items", out.items, (m, k), torch.uint32),
                                   ("distances", out.distances, (m, k), odt),
                                   ("counts", out.counts, (m,), torch.uint32)):
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or tuple(t.shape) != shape
                    or t.dtype != dt or not t.is_contiguous()):
                raise ValueError(f"out.{name} must be a contiguous cuda {dt} tensor of shape {shape}")

    def morton_order(self, q, stream=None):
        """The permutation KD_MORTON visits q in (a cuda uint32 tensor of q's indices)."""
        torch = _torch()
        q = self._queries(q)
        perm = torch.empty(q.shape[0], dtype=torch.uint32, device="cuda")
        _check(lib().kd_morton_order(self._h, q.data_ptr(), q.shape[0], perm.data_ptr(), _stream_ptr(stream)),
               "kd_morton_order")
        return perm

    def nearest_one_raw(self, q, flags=0, stream=None, approx=False):
        torch = _torch()
        q = self._queries(q)
        m = q.shape[0]
        items = torch.empty(m, dtype=torch.uint32, device="cuda")
        if approx:
            d = self._out(m, dtype={np.float32: torch.float32, np.float64: torch.float64}[self.a_dtype])
            fn, name = lib().kd_approx_nearest_one, "kd_approx_nearest_one"
        else:
            d = self._out(m)
            fn, name = lib().kd_nearest_one, "kd_nearest_one"
        _check(fn(self._h, q.data_ptr(), m, flags, items.data_ptr(), d.data_ptr(),
                  _stream_ptr(stream)), name)
        return NearestOne(items, d)

    def nearest_n_raw(self, q, k, radius=math.inf, flags=0, stream=None, out=None):
        torch = _torch()
        q = self._queries(q)
        m = q.shape[0]
        if out is None:
            out = NearestN(torch.empty((m, k), dtype=torch.uint32, device="cuda"),
                           self._out(m, k), torch.empty(m, dtype=torch.uint32, device="cuda"))
        else:
            self._check_out(out, m, k)
        _check(lib().kd_nearest_n(self._h, q.data_ptr(), m, k, float(radius), flags,
                                  out.items.data_ptr(), out.distances.data_ptr(),
                                  out.counts.data_ptr(), _stream_ptr(stream)), "kd_nearest_n")
        return out

    def within_raw(self, q, radius, flags=0, stream=None):
        torch = _torch()
        q = self._queries(q)
        m = q.shape[0]
        offsets = torch.empty(m + 1, dtype=torch.int64, device="cuda")
        sp = _stream_ptr(stream)
        _check(lib().kd_within_count(self._h, q.data_ptr(), m, float(radius), flags,
                                     offsets.data_ptr(), sp), "kd_within_count")
        if stream is not None:
            stream.synchronize()  # the total below is read on torch's current stream
        total = int(offsets[m].item()) if m else 0
        items = torch.empty(max(total, 1), dtype=torch.uint32, device="cuda")
        d = self._out(max(total, 1))
        _check(lib().kd_within_fill(self._h, q.data_ptr(), m, float(radius), flags,
                                    offsets.data_ptr(), items.data_ptr(), d.data_ptr(), sp),
               "kd_within_fill")
        return Within(offsets, items[:total], d[:total])

    def best_n_within_raw(self, q, radius, n, flags=0, stream=None):
        torch = _torch()
        q = self._queries(q)
        m = q.shape[0]
        out = NearestN(torch.empty((m, n), dtype=torch.uint32, device="cuda"),
                       self._out(m, n), torch.empty(m, dtype=torch.uint32, device="cuda"))
        _check(lib().kd_best_n_within(self._h, q.data_ptr(), m, float(radius), n, flags,
                                      out.items.data_ptr(), out.distances.data_ptr(),
                                      out.counts.data_ptr(), _stream_ptr(stream)),
               "kd_best_n_within")
        return out


    # -- periodic boundary conditions (query/periodic.rs) --------------------------------
    def _box(self, box):
        b = np.ascontiguousarray(box, dtype=self.a_dtype)
        if b.shape != (self.K,):
            raise ValueError(f"box must hold {self.K} lengths")
        return b

    def periodic_nearest_one_raw(self, q, box, flags=0, stream=None):
        torch = _torch()
        q, b = self._queries(q), self._box(box)
        m = q.shape[0]
        items = torch.empty(m, dtype=torch.uint32, device="cuda")
        d = self._out(m)
        _check(lib().kd_periodic_nearest_one(self._h, q.data_ptr(), m, _np_ptr(b), flags,
                                             items.data_ptr(), d.data_ptr(), _stream_ptr(stream)),
               "kd_periodic_nearest_one")
        return NearestOne(items, d)

    def periodic_nearest_n_raw(self, q, box, k, radius=math.inf, flags=0, stream=None):
        torch = _torch()
        q, b = self._queries(q), self._box(box)
        m = q.shape[0]
        out = NearestN(torch.empty((m, k), dtype=torch.uint32, device="cuda"),
                       self._out(m, k), torch.empty(m, dtype=torch.uint32, device="cuda"))
        _check(lib().kd_periodic_nearest_n(self._h, q.data_ptr(), m, _np_ptr(b), k, float(radius),
                                           flags, out.items.data_ptr(), out.distances.data_ptr(),
                                           out.counts.data_ptr(), _stream_ptr(stream)),
               "kd_periodic_nearest_n")
        return out

    def periodic_within_raw(self, q, box, radius, flags=0, stream=None):
        torch = _torch()
        q, b = self._queries(q), self._box(box)
        m = q.shape[0]
        offsets = torch.empty(m + 1, dtype=torch.int64, device="cuda")
        sp = _stream_ptr(stream)
        _check(lib().kd_periodic_within_count(self._h, q.data_ptr(), m, _np_ptr(b), float(radius),
                                              flags, offsets.data_ptr(), sp),
               "kd_periodic_within_count")
        if stream is not None:
            stream.synchronize()  # the total below is read on torch's current stream
        total = int(offsets[m].item()) if m else 0
        items = torch.empty(max(total, 1), dtype=torch.uint32, device="cuda")
        d = self._out(max(total, 1))
        _check(lib().kd_periodic_within_fill(self._h, q.data_ptr(), m, _np_ptr(b), float(radius),
                                             flags, offsets.data_ptr(), items.data_ptr(),
                                             d.data_ptr(), sp), "kd_periodic_within_fill")
        return Within(offsets, items[:total], d[:total])


class Forest:
    """Many independent trees (same dtype and K) in pooled HBM buffers, queried as one batch
    by a persistent kernel fed from a device work queue (kd_forest_* in kiddo_amd.h)."""

    def __init__(self, layouts, shadow_f32: bool = False):
        """layouts: HostLayout objects (or anything with a `.desc` TreeDesc). shadow_f32 (f32
        trees): keep an FP32 shadow of the entries for the FP32 collector (KD_SCREEN_F32)."""
        self._layouts = list(layouts)
        arr = (TreeDesc * len(self._layouts))(*[l.desc for l in self._layouts])
        h = C.c_void_p()
        _check(lib().kd_forest_create(arr, len(self._layouts), KD_TREE_SHADOW_F32 if shadow_f32 else 0,
                                      C.byref(h)), "kd_forest_create")
        self._h = h
        self.ntrees = len(self._layouts)
        self.dtype_code = self._layouts[0].dtype_code
        self.K = self._layouts[0].K
        self.m = 0
        self._layouts = None  # the device pools hold everything now

    @classmethod
    def from_points(cls, point_sets, bucket=32, threads=0, shadow_f32: bool = False):
        return cls([HostLayout(p, bucket, None, threads) for p in point_sets], shadow_f32)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kd_forest_destroy(self._h)
            self._h = None

    @property
    def out_dtype(self):
        return _O_NP[self.dtype_code]

    @staticmethod
    def autotune_bucket(point_sets, queries, query_offsets, k: int, radius: float = math.inf, flags: int = 0,
                        buckets=(16, 32, 64, 128), stream=None):
        """Bucket size of a batched forest by measurement (kd_forest_autotune_bucket): the
        given trees (host (n_i, K) arrays, e.g. a sample of the forest's) are built at each
        candidate bucket, planned with their concatenated device queries (query_offsets: one
        more than trees), the launch autotuned and nearest_n timed. Returns
        (best_bucket, {bucket: kernel ms})."""
        pts = [np.ascontiguousarray(p) for p in point_sets]
        code = _NP_CODE[pts[0].dtype]
        K = pts[0].shape[1]
        ptrs = (C.c_void_p * len(pts))(*[p.ctypes.data for p in pts])
        sizes = np.ascontiguousarray([p.shape[0] for p in pts], dtype=np.uint64)
        offs = np.ascontiguousarray(query_offsets, dtype=np.uint64)
        b = np.ascontiguousarray(buckets, dtype=np.int32)
        ms = np.zeros(b.size, np.float32)
        best = C.c_int32()
        _check(lib().kd_forest_autotune_bucket(code, K, len(pts), ptrs, _np_ptr(sizes), queries.data_ptr(),
                                               _np_ptr(offs), k, float(radius), flags, _np_ptr(b), b.size,
                                               C.byref(best), _np_ptr(ms), _stream_ptr(stream)),
               "kd_forest_autotune_bucket")
        return int(best.value), {int(x): float(y) for x, y in zip(b, ms)}

    def plan(self, queries, query_offsets, chunk=256, morton=True, stream=None):
        """queries: concatenated (m, K) device tensor; query_offsets: ntrees+1 host ints."""
        offs = np.ascontiguousarray(query_offsets, dtype=np.uint64)
        assert offs.size == self.ntrees + 1
        self.m = int(offs[-1])
        _check(lib().kd_forest_plan(self._h, queries.data_ptr(), _np_ptr(offs), chunk,
                                    KD_MORTON if morton else 0, _stream_ptr(stream)),
               "kd_forest_plan")
        self.query_offsets = offs

    def tune(self, variant=1, group_size=0, lds_levels=0, waves=0):
        _check(lib().kd_forest_tune(self._h, variant, group_size, lds_levels, waves), "kd_forest_tune")

    def _knn_out(self, k):
        torch = _torch()
        dt = {np.float32: torch.float32, np.float64: torch.float64}[self.out_dtype]
        return NearestN(torch.empty((self.m, k), dtype=torch.uint32, device="cuda"),
                        torch.empty((self.m, k), dtype=dt, device="cuda"),
                        torch.empty(self.m, dtype=torch.uint32, device="cuda"))

    def nearest_n(self, queries, k, radius=math.inf, flags=0, out=None, stream=None):
        out = out or self._knn_out(k)
        _check(lib().kd_forest_nearest_n(self._h, queries.data_ptr(), k, float(radius), flags,
                                         out.items.data_ptr(), out.distances.data_ptr(),
                                         out.counts.data_ptr(), _stream_ptr(stream)),
               "kd_forest_nearest_n")
        return out

    def autotune_nearest_n(self, queries, k, radius=math.inf, flags=0, stream=None):
        out = self._knn_out(k)
        chosen = np.zeros(4, np.int32)
        _check(lib().kd_forest_autotune_nearest_n(self._h, queries.data_ptr(), k, float(radius),
                                                  flags, out.items.data_ptr(),
                                                  out.distances.data_ptr(), out.counts.data_ptr(),
                                                  _np_ptr(chosen), _stream_ptr(stream)),
               "kd_forest_autotune_nearest_n")
        return dict(zip(("variant", "group_size", "lds_levels", "waves"), (int(x) for x in chosen)))

    def within(self, queries, radius, flags=0, stream=None):
        torch = _torch()
        offsets = torch.empty(self.m + 1, dtype=torch.int64, device="cuda")
        sp = _stream_ptr(stream)
        _check(lib().kd_forest_within(self._h, queries.data_ptr(), float(radius), flags,
                                      offsets.data_ptr(), None, None, sp), "kd_forest_within(count)")
        if stream is not None:
            stream.synchronize()  # the total below is read on torch's current stream
        total = int(offsets[self.m].item()) if self.m else 0
        dt = {np.float32: torch.float32, np.float64: torch.float64}[self.out_dtype]
        items = torch.empty(max(total, 1), dtype=torch.uint32, device="cuda")
        d = torch.empty(max(total, 1), dtype=dt, device="cuda")
        _check(lib().kd_forest_within(self._h, queries.data_ptr(), float(radius), flags,
                                      offsets.data_ptr(), items.data_ptr(), d.data_ptr(), sp),
               "kd_forest_within(fill)")
        return Within(offsets, items[:total], d[:total])

    def set_timing(self, enable=True):
        _check(lib().kd_forest_set_timing(self._h, int(enable)), "kd_forest_set_timing")

    def last_kernel_ms(self):
        ms = C.c_float()
        _check(lib().kd_forest_last_kernel_ms(self._h, C.byref(ms)), "kd_forest_last_kernel_ms")
        return ms.value

    def query_stats(self):
        out = np.zeros(4, np.uint64)
        _check(lib().kd_forest_query_stats(self._h, _np_ptr(out)), "kd_forest_query_stats")
        return dict(zip(("leaves", "points", "stems", "rescored"), (int(x) for x in out)))


class Request(C.Structure):
    _fields_ = [("tree", C.c_int32), ("family", C.c_int32), ("k", C.c_uint32),
                ("flags", C.c_uint32), ("radius", C.c_double), ("queries", C.c_void_p),
                ("m", C.c_uint64), ("items", C.c_void_p), ("dists", C.c_void_p),
                ("counts", C.c_void_p)]


FAMILY_NEAREST_ONE, FAMILY_NEAREST_N, FAMILY_BEST_N_WITHIN, FAMILY_APPROX = 1, 2, 3, 4


class Server:
    """Continuous-batching request loop over resident trees (kd_server_* in kiddo_amd.h).
    Requests take host numpy queries; results come back as host numpy arrays."""

    def __init__(self, trees, arena_bytes=256 << 20, slots=4, max_coalesce=1 << 18, max_part=None):
        self.trees = list(trees)
        arr = (C.c_void_p * len(self.trees))(*[t.handle.value for t in self.trees])
        h = C.c_void_p()
        _check(lib().kd_server_create(arr, len(self.trees), arena_bytes, slots, max_coalesce,
                                      C.byref(h)), "kd_server_create")
        self._h = h
        self._live = {}
        if max_part is not None:
            _check(lib().kd_server_set_part(self._h, max_part), "kd_server_set_part")

    def submit(self, tree, family, queries, k=1, radius=math.inf, flags=0):
        t = self.trees[tree]
        q = np.ascontiguousarray(queries, dtype=t.a_dtype)
        m = q.shape[0]
        w = 1 if family in (FAMILY_NEAREST_ONE, FAMILY_APPROX) else k
        out_dt = t.a_dtype if family == FAMILY_APPROX else t.out_dtype
        items = np.empty((m, w), np.uint32)
        dists = np.empty((m, w), out_dt)
        counts = np.empty(m, np.uint32)
        r = Request(tree, family, k, flags, float(radius), q.ctypes.data, m, items.ctypes.data,
                    dists.ctypes.data, counts.ctypes.data)
        ticket = C.c_uint64()
        _check(lib().kd_server_submit(self._h, C.byref(r), C.byref(ticket)), "kd_server_submit")
        self._live[ticket.value] = (q, items, dists, counts, family)
        return ticket.value

    def wait(self, ticket):
        """Returns (items, dists, counts, latency_ms); nearest_one rows are 1 wide."""
        lat = C.c_double()
        rc = lib().kd_server_wait(self._h, ticket, C.byref(lat))
        q, items, dists, counts, family = self._live.pop(ticket)
        _check(rc, "kd_server_wait")
        if family in (FAMILY_NEAREST_ONE, FAMILY_APPROX):
            return items[:, 0], dists[:, 0], None, lat.value
        return items, dists, counts, lat.value

    def stats(self):
        out = np.zeros(6, np.uint64)
        _check(lib().kd_server_stats(self._h, _np_ptr(out)), "kd_server_stats")
        return dict(zip(("done", "launches", "queries", "coalesced", "memory_waits",
                         "peak_arena_bytes"), (int(x) for x in out)))

    def close(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.kd_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class QueryBuilder:
    """Fluent batch builder mirroring kiddo src/kd_tree/query/builder.rs."""

    def __init__(self, tree: ImmutableKdTree, queries, stream=None):
        self.tree = tree
        self.queries = queries
        self.stream = stream
        self.family = None
        self.k = None
        self.radius = math.inf
        self.flags = 0
        self.is_approx = False
        self.box = None

    def _family(self, fam):
        if self.family is not None:
            raise ValueError("query family already chosen")
        self.family = fam
        return self

    # space (builder.rs:1280-1291, 1989-2003): chosen before the family
    def periodic_boundary_condition(self, box_size):
        if self.family is not None:
            raise ValueError("periodic_boundary_condition() must come before the query family")
        self.box = box_size
        return self

    # families (builder.rs:2008-2057)
    def nearest_one(self, metric=None):
        self.flags |= metric_flags(metric)
        return self._family("nearest_one")

    def nearest_n(self, k: int, metric=None):
        if k < 1:
            raise ValueError("max_qty must be non-zero")
        self.k = k
        self.flags |= metric_flags(metric)
        return self._family("nearest_n")

    def within(self, radius: float, metric=None):
        if self.family == "nearest_n":  # nearest_n(k).within(radius)
            self.radius = radius
            return self
        self.radius = radius
        self.flags |= metric_flags(metric)
        return self._family("within")

    def best_n_within(self, radius: float, n: int, metric=None):
        self.flags |= metric_flags(metric)
        if self.box is not None:
            raise ValueError("best_n_within is Cartesian-only (builder.rs:1287-1290)")
        if n < 1:
            raise ValueError("max_qty must be non-zero")
        self.radius = radius
        self.k = n
        return self._family("best_n_within")

    # options (builder.rs:2059-2081)
    def approx(self):
        if self.box is not None:
            raise ValueError("approx() is not available for periodic queries")
        if self.family != "nearest_one":
            raise ValueError("approx() is only available after nearest_one()")
        self.is_approx = True
        return self

    def unsorted(self):
        if self.family not in ("within", "nearest_n"):
            raise ValueError("unsorted() applies to within / nearest_n")
        self.flags |= KD_UNSORTED
        return self

    def exclusive_boundaries(self):
        if self.family not in ("within", "nearest_n", "best_n_within"):
            raise ValueError("exclusive_boundaries() applies to radius queries")
        self.flags |= KD_EXCLUSIVE
        return self

    # engine options (no Kiddo counterpart: they change speed, not results)
    def morton(self):
        self.flags |= KD_MORTON
        return self

    def screen_f32(self):
        self.flags |= KD_SCREEN_F32
        return self

    def execute(self):
        t = self.tree
        if self.box is not None:
            if self.family == "nearest_one":
                return t.periodic_nearest_one_raw(self.queries, self.box, self.flags, self.stream)
            if self.family == "nearest_n":
                return t.periodic_nearest_n_raw(self.queries, self.box, self.k, self.radius,
                                                self.flags, self.stream)
            if self.family == "within":
                return t.periodic_within_raw(self.queries, self.box, self.radius, self.flags,
                                             self.stream)
            raise ValueError("choose a query family first")
        if self.family == "nearest_one":
            return t.nearest_one_raw(self.queries, self.flags, self.stream, approx=self.is_approx)
        if self.family == "nearest_n":
            return t.nearest_n_raw(self.queries, self.k, self.radius, self.flags, self.stream)
        if self.family == "within":
            return t.within_raw(self.queries, self.radius, self.flags, self.stream)
        if self.family == "best_n_within":
            return t.best_n_within_raw(self.queries, self.radius, self.k, self.flags, self.stream)
        raise ValueError("choose a query family first")


__all__ = ["ImmutableKdTree", "HostLayout", "Forest", "Server", "KD_GRAPH", "FAMILY_NEAREST_ONE",
           "FAMILY_NEAREST_N", "FAMILY_BEST_N_WITHIN", "FAMILY_APPROX", "QueryBuilder", "NearestOne", "NearestN",
           "Within", "KiddoAmdError", "lib", "SIGNATURES", "KD_EXCLUSIVE", "KD_UNSORTED",
           "KD_MORTON", "KD_SCREEN_F32", "KD_TREE_SHADOW_F32", "KD_METRIC_SQEUCLIDEAN",
           "KD_METRIC_MANHATTAN", "KD_METRIC_CHEBYSHEV", "KD_METRIC_MINKOWSKI", "metric_flags"]
