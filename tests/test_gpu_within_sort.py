"""Sorted within: segments of every length from 0 to ~60 results (short ones sorted in
registers, longer ones in global memory), with exact-duplicate ties, against the oracle bit
for bit, under the thread and group fill shapes, Cartesian and periodic."""
import numpy as np
import pytest

from engines import GpuEngine, OracleEngine, assert_same_bits

pytestmark = pytest.mark.gpu


def _announce(*what):
    print("case:", *what, flush=True)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_within_sorted_segment_lengths(ka, oracle_mod, cuda, dtype):
    _announce("within sorted segment lengths", np.dtype(dtype).name)
    rng = np.random.default_rng(31)
    base = rng.random((40_000, 3))
    pts = np.concatenate([base, base[rng.integers(0, len(base), 4000)]]).astype(dtype)  # ties
    q = rng.random((6000, 3)).astype(dtype)
    q[:500] = pts[rng.integers(0, len(pts), 500)]
    O = OracleEngine(oracle_mod)
    ot = O.build(pts, 32)
    G = GpuEngine(ka, flags=ka.KD_MORTON)
    gt = G.build(pts, 32)
    seen = np.zeros(0, np.int64)
    # about 0.3, 2.7, 10 and 78 results per query on average (at most 5, 11, 27 and 122)
    for r in (1e-4, 6e-4, 1.5e-3, 6e-3):
        oo, oi, od = O.within(ot, q, r, False, sorted=True)
        seen = np.concatenate([seen, np.diff(oo.astype(np.int64))])
        for fill in (1, 8, 0):
            gt.tune_within(0, fill)
            go, gi, gd = G.within(gt, q, r, False, sorted=True)
            assert_same_bits(go, oo, f"offsets r={r} fill={fill}")
            assert_same_bits(gi, oi, f"items r={r} fill={fill}")
            assert_same_bits(gd, od, f"dists r={r} fill={fill}")
        gt.tune_within(0, 0)
    for lo, hi in ((0, 1), (1, 4), (4, 8), (8, 16), (16, 48), (48, 1 << 30)):
        assert np.any((seen > lo) & (seen <= hi)), (lo, hi)
    box = np.ones(3, dtype)
    oo, oi, od = ot.periodic_within(q, box, 1.5e-3, False, True)
    go, gi, gd = G.periodic_within(gt, q, box, 1.5e-3, False, True)
    for a, b, nm in ((go, oo, "offsets"), (gi, oi, "items"), (gd, od, "dists")):
        assert_same_bits(a, b, f"periodic within {nm}")
