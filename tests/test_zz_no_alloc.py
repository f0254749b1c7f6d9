"""kd_nearest_n_noalloc (the allocation-free form of kd_nearest_n, whose Morton-ordering
scratch lives in the library's static device memory) against kd_nearest_n, through the C
ABI, bit for bit: two calls in a row, the second on a smaller batch, for each flag set;
the batch limit under KD_MORTON; any batch size without it."""
import math

import numpy as np
import pytest

from engines import assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


def _call(ka, fn, tree, q, k, radius, flags):
    torch = ka._torch()
    m = q.shape[0]
    items = torch.full((m, k), 7, dtype=torch.uint32, device="cuda")
    dists = torch.full((m, k), 3.0, dtype=torch.float64 if tree.out_dtype == np.float64 else torch.float32,
                       device="cuda")
    counts = torch.full((m,), 9, dtype=torch.uint32, device="cuda")
    rc = fn(tree.handle, q.data_ptr(), m, k, float(radius), flags, items.data_ptr(), dists.data_ptr(),
            counts.data_ptr(), ka._stream_ptr(None))
    torch.cuda.synchronize()
    return rc, (items.cpu().numpy(), dists.cpu().numpy(), counts.cpu().numpy())


def _same(ka, tree, q, k, radius, flags, tag):
    L = ka.lib()
    rc0, ref = _call(ka, L.kd_nearest_n, tree, q, k, radius, flags)
    rc1, got = _call(ka, L.kd_nearest_n_noalloc, tree, q, k, radius, flags)
    assert rc0 == 0 and rc1 == 0, (tag, rc0, rc1, L.kd_last_error())
    for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
        assert_same_bits(a, b, f"{tag} {nm}")


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_noalloc_equals_nearest_n(ka, cuda, dtype):
    torch = cuda
    from kiddo_amd.datasets import uniform_with_duplicates
    pts = uniform_with_duplicates(200_000, 0.05, seed=4).astype(dtype)
    shadow = dtype == np.float64
    tree = ka.ImmutableKdTree.new_from_slice(pts, 32, shadow_f32=shadow)
    rng = np.random.default_rng(6)
    big = torch.from_numpy(rng.random((300_000, 3)).astype(dtype)).cuda()
    small = torch.from_numpy(rng.random((70_001, 3)).astype(dtype)).cuda()
    flag_sets = [0, ka.KD_MORTON, ka.KD_MORTON | ka.KD_EXCLUSIVE]
    if shadow:
        flag_sets += [ka.KD_SCREEN_F32, ka.KD_SCREEN_F32 | ka.KD_MORTON]
    for flags in flag_sets:
        for k, radius in ((16, math.inf), (5, 2e-4)):
            for name, q in (("first", big), ("second, smaller", small)):
                _announce("noalloc", np.dtype(dtype).name, flags, k, name)
                _same(ka, tree, q, k, radius, flags, f"{name} flags={flags} k={k}")


def test_noalloc_limits(ka, cuda):
    torch = cuda
    L = ka.lib()
    rng = np.random.default_rng(2)
    pts = rng.random((50_000, 3))
    tree = ka.ImmutableKdTree.new_from_slice(pts, 32)
    lim = int(ka.KD_NOALLOC_MAX_QUERIES)
    q = torch.from_numpy(rng.random((lim + 1, 3))).cuda()
    _announce("noalloc at the limit")
    _same(ka, tree, q[:lim], 4, math.inf, ka.KD_MORTON, "m = limit, Morton")
    _announce("noalloc above the limit")
    rc, _ = _call(ka, L.kd_nearest_n_noalloc, tree, q, 4, math.inf, ka.KD_MORTON)
    assert rc == -7 and b"KD_MORTON" in L.kd_last_error()
    _same(ka, tree, q, 4, math.inf, 0, "m above the limit, input order")
    _announce("noalloc small after large")
    _same(ka, tree, q[:3], 4, math.inf, ka.KD_MORTON, "m = 3, Morton")
    _same(ka, tree, q[:1], 4, math.inf, ka.KD_MORTON, "m = 1, Morton")
