"""Launch-shape sweep for nearest_n on the 16.8M-point f64 workload of bench.py.

Times the k-NN kernel (HIP events, best of `--reps`) for k in {1, 4, 8, 16} across the
plain FP64 kernel in input order, the plain kernel in Morton order, and the FP32-screened
kernel in Morton order at several occupancy floors, plus the group-per-query form.

    python -m kiddo_amd.sweep [--points 16000000] [--queries 16000000]
"""
from __future__ import annotations

import argparse
import json
import math

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=16_000_000)
    ap.add_argument("--queries", type=int, default=16_000_000)
    ap.add_argument("--reps", type=int, default=2)
    args = ap.parse_args()

    import torch
    import kiddo_amd as ka
    from kiddo_amd.datasets import uniform_with_duplicates

    pts = uniform_with_duplicates(args.points, 0.05, seed=1)
    qh = np.random.default_rng(2).random((args.queries, 3))
    lay = ka.HostLayout(pts, 32)
    t = ka.ImmutableKdTree(lay.desc, shadow_f32=True)
    t.set_timing(True)
    q = torch.from_numpy(qh).cuda()

    def run(k, flags, variant=1, waves=0, lds=-1, gs=0):
        t.tune(variant, gs, lds, waves)
        out, ms = None, []
        for _ in range(args.reps):
            out = t.nearest_n_raw(q, k, math.inf, flags, out=out)
            ms.append(t.last_kernel_ms())
        return round(min(ms), 2)

    S, M = ka.KD_SCREEN_F32, ka.KD_MORTON
    res = {}
    for k in (1, 4, 8, 16):
        res[f"plain_k{k}"] = run(k, 0)
        res[f"plain_morton_k{k}"] = run(k, M)
        for w in (3, 4, 5):
            res[f"screen_morton_k{k}_w{w}"] = run(k, S | M, waves=w)
        print(k, {a: b for a, b in res.items() if a.endswith(f"k{k}") or f"k{k}_" in a}, flush=True)
    res["screen_input_order_k16_w4"] = run(16, S, waves=4)
    res["group16_screen_morton_k16"] = run(16, S | M, variant=2, gs=16)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
