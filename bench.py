#!/usr/bin/env python3
"""Benchmark of the batch k-NN hot path on one MI355X.

Workload (the precision form of the headline metric): ImmutableKdTree<f64, 3> of
16,000,000 points uniform in [0,1)^3 plus 5% (800,000) exact copies of uniformly chosen
points (ties), bucket 32; 16,000,000 uniform queries; nearest_n(k=16). Inputs are seeded
and synthetic and already resident in HBM when timing starts.

A step is one batch of all queries through one kernel variant:
  * plain       — the plain FP64 kernel, queries in Morton order, launch shape autotuned
                  on the batch (thread or group per query, LDS stem levels, occupancy);
  * plain_input — the plain FP64 kernel in input order, one thread per query, no LDS stems;
  * screened    — the batched FP32-screened kernel: queries in Morton order, FP32 shadow
                  screening against a certified bound, FP64 rescore (identical results),
                  launch shape autotuned the same way.
`value` is the screened variant's queries/s; `screened_over_plain` compares it with the
plain kernel under the same query order and tuning (like for like), and
`screened_over_plain_input_order` with the plain kernel in input order. All three
variants must agree bit for bit, and a sample of 200,000 queries spread over the batch is
compared bit for bit (items, order, distances, counts) with the oracle in its canonical
mode: the bench exits with status 3 on any mismatch. Also reported: p50/p99 latency of
fixed 262,144-query sub-batches, the fraction of scanned entries rescored in FP64 and the
largest FP32 screen error seen, the HBM roofline of the screened kernel from its
algorithmic bytes, and the CPU baseline (the oracle's restatement of Kiddo's scalar query,
run on this host's cores over a sample of the same queries).

    python bench.py [--steps 5] [--warmup 2]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "queries/s, 16M-pt 3D kNN k=16: plain FP64 vs FP32-screened batched; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=16_000_000)
    ap.add_argument("--dup-frac", type=float, default=0.05)
    ap.add_argument("--queries", type=int, default=16_000_000)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--bucket", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--latency-batch", type=int, default=262_144)
    ap.add_argument("--cpu-sample", type=int, default=2_000_000)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["all", "plain", "plain_input", "screened"], default="all")
    ap.add_argument("--parity-sample", type=int, default=200_000,
                    help="knn16: queries compared with the canonical oracle (0: none)")
    ap.add_argument("--no-latency", action="store_true")
    ap.add_argument("--no-within", action="store_true", help="knn16: skip the within(r=0.002) pass")
    ap.add_argument("--no-device-build", action="store_true", help="knn16: skip timing the GPU tree build")
    ap.add_argument("--workload", choices=["knn16", "forest", "serving", "mixture", "small", "periodic",
                                           "metrics"],
                    default="knn16",
                    help="knn16: the headline 16M-point k-NN; forest: 4096-tree batched form; "
                         "serving: continuous batching over resident trees")
    ap.add_argument("--requests", type=int, default=3000)
    ap.add_argument("--result-budget", type=int, default=8 << 30,
                    help="mixture: bytes of within results held at once (queries run in chunks)")
    ap.add_argument("--rate", type=float, default=4000.0, help="serving: Poisson arrivals per second")
    ap.add_argument("--morton-min", type=int, default=4096,
                    help="serving: requests of at least this many queries are Morton-ordered")
    ap.add_argument("--max-part", type=int, default=16384,
                    help="serving: requests are cut into parts of at most this many queries (0: never)")
    ap.add_argument("--rates", default="2000,8000,16000,32000,64000",
                    help="serving: arrival rates (per second) of the rate sweep before the main run")
    ap.add_argument("--trees", type=int, default=4096)
    ap.add_argument("--forest-bucket", type=int, default=0,
                    help="forest: bucket size of every tree (0: autotuned on every 16th tree)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this engine runs a batch on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import kiddo_amd as ka
    from kiddo_amd.datasets import uniform_with_duplicates

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    if args.workload == "forest":
        return bench_forest(args, torch, ka)
    if args.workload == "serving":
        return bench_serving(args, torch, ka)
    if args.workload == "mixture":
        return bench_mixture(args, torch, ka)
    if args.workload == "small":
        return bench_small(args, torch, ka)
    if args.workload == "periodic":
        return bench_periodic(args, torch, ka)
    if args.workload == "metrics":
        return bench_metrics(args, torch, ka)

    if args.copy_probe:
        # A plain copy of known size: what FETCH_SIZE / WRITE_SIZE read for it scales the
        # same counters of the query kernels into bytes.
        src = torch.ones(args.copy_probe // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "copies": 3}), flush=True)
        return

    t0 = time.time()
    pts = uniform_with_duplicates(args.points, args.dup_frac, seed=args.seed)
    qh = np.random.default_rng(args.seed + 1).random((args.queries, 3))
    log(f"data {pts.shape[0]} points, {qh.shape[0]} queries in {time.time() - t0:.1f}s")
    t0 = time.time()
    lay = ka.HostLayout(pts, args.bucket)
    host_build_s = time.time() - t0
    log(f"host build {host_build_s:.1f}s, stems depth {lay.max_stem_level + 1}")
    t0 = time.time()
    tree = ka.ImmutableKdTree(lay.desc, shadow_f32=True)
    tree.set_timing(True)
    log(f"upload {time.time() - t0:.1f}s, {tree.info.device_bytes / 1e6:.0f} MB in HBM")

    m, k = args.queries, args.k
    q = torch.from_numpy(qh).to(dev)
    variants = ["plain_input", "plain", "screened"] if args.only == "all" else [args.only]
    outs = {name: ka.NearestN(torch.empty((m, k), dtype=torch.uint32, device=dev),
                              torch.empty((m, k), dtype=torch.float64, device=dev),
                              torch.empty(m, dtype=torch.uint32, device=dev))
            for name in variants}
    flags = {"plain_input": 0, "plain": ka.KD_MORTON, "screened": ka.KD_SCREEN_F32 | ka.KD_MORTON}

    # plain_input: one thread per query, input order, no LDS stems. plain and screened: the
    # launch shape the autotuner picks for each on a sample of the batch.
    tuning = {"plain_input": dict(variant=1, group_size=0, lds_levels=0, waves=0)}
    for name in ("plain", "screened"):
        if name in variants:
            tuning[name] = tree.autotune_nearest_n(q, k, math.inf, flags[name])
            log(f"autotuned {name} launch: {tuning[name]}")

    def run(name, qq=q, out=None):
        tree.nearest_n_raw(qq, k, math.inf, flags[name], out=out or outs[name])

    results = {}
    for name in variants:
        tree.tune(**{("knn_variant" if a == "variant" else a): v for a, v in tuning[name].items()})
        for _ in range(args.warmup):
            run(name)
        torch.cuda.synchronize()
        kernel_ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run(name)
            kernel_ms.append(tree.last_kernel_ms())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        stats = tree.query_stats()
        results[name] = dict(ms_per_step=1e3 * dt / args.steps, qps=m * args.steps / dt,
                             kernel_ms=float(np.mean(kernel_ms)), stats=stats)
        log(f"{name}: {results[name]['ms_per_step']:.1f} ms/step, "
            f"{results[name]['qps'] / 1e6:.1f} Mq/s, kernel {results[name]['kernel_ms']:.1f} ms, {stats}")

    # every variant returns the same results, bit for bit
    parity = None
    if len(variants) > 1:
        a = outs[variants[0]]
        parity = all(bool(torch.equal(a.items, outs[v].items) and torch.equal(a.counts, outs[v].counts) and
                          torch.equal(a.distances.view(torch.int64), outs[v].distances.view(torch.int64)))
                     for v in variants[1:])
        log(f"{' == '.join(variants)} bit for bit: {parity}")
    head = "screened" if "screened" in results else variants[-1]

    # a sample spread over the batch against the oracle's canonical answers
    oracle_parity = None
    if args.parity_sample:
        from oracle import oracle as O
        O.build()
        ot_c = O.Tree.from_layout(np.float64, 3, args.bucket, pts.shape[0], lay.max_stem_level,
                                  lay.stems(), lay.extents(), lay.arena())
        ns = min(args.parity_sample, m)
        rows = np.sort(np.random.default_rng(args.seed + 5).choice(m, ns, replace=False))
        ref = ot_c.nearest_n(np.ascontiguousarray(qh[rows]), k, mode=O.CANONICAL,
                             threads=min(16, os.cpu_count() or 1))
        del ot_c
        ridx = torch.from_numpy(rows).to(dev)
        oracle_parity = {}
        for name in variants:
            got = (outs[name].items.view(torch.int32)[ridx].cpu().numpy().view(np.uint32),
                   outs[name].distances[ridx].cpu().numpy(), outs[name].counts.view(torch.int32)[ridx].cpu().numpy().view(np.uint32))
            oracle_parity[name] = bool(np.array_equal(got[0], ref[0]) and np.array_equal(got[2], ref[2]) and
                                       np.array_equal(got[1].view(np.uint64), ref[1].view(np.uint64)))
        log(f"oracle (canonical) on {ns} sampled queries: {oracle_parity}")

    # latency of fixed-size sub-batches (screened path)
    tree.tune(**{("knn_variant" if a == "variant" else a): v for a, v in tuning[head].items()})
    lb = min(args.latency_batch, m)
    nb = 0 if args.no_latency else m // lb
    sub = ka.NearestN(torch.empty((lb, k), dtype=torch.uint32, device=dev),
                      torch.empty((lb, k), dtype=torch.float64, device=dev),
                      torch.empty(lb, dtype=torch.uint32, device=dev))
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(nb)]
    for b in range(nb):
        ev[b][0].record()
        run(head, q[b * lb:(b + 1) * lb], sub)
        ev[b][1].record()
    torch.cuda.synchronize()
    lat = sorted(e0.elapsed_time(e1) for e0, e1 in ev)
    p50 = float(np.percentile(lat, 50)) if lat else None
    p99 = float(np.percentile(lat, 99)) if lat else None

    # FP32 screen error on a sample of returned pairs, against the FP64 distances
    screen_err = None
    if "screened" in results:
        rng = np.random.default_rng(7)
        rows = rng.integers(0, m, 200_000)
        ridx = torch.from_numpy(rows).to(dev)
        it = outs["screened"].items.view(torch.int32)[ridx].cpu().numpy().view(np.uint32).astype(np.int64)
        d64 = outs["screened"].distances[ridx].cpu().numpy()
        p32 = pts[it].astype(np.float32)
        q32 = qh[rows][:, None, :].astype(np.float32)
        acc = np.zeros(d64.shape, np.float32)
        for d in range(3):
            df = (p32[..., d] - q32[..., d]).astype(np.float32)
            acc = (acc + df * df).astype(np.float32)
        err = np.abs(acc.astype(np.float64) - d64)
        screen_err = dict(max_abs=float(err.max()),
                          max_rel=float((err / np.maximum(d64, 1e-300)).max()))

    # algorithmic bytes per launch, from the reference DFS's visit counts (the plain kernel
    # visits exactly the leaves and stems Kiddo's traversal does; the screened kernel starts
    # from a certified bound and visits fewer): every scanned entry (K*8 coordinate bytes +
    # 4 item bytes), every stem read, the queries and the results
    st = results["plain_input"]["stats"] if "plain_input" in results else results[head]["stats"]
    own = results[head]["stats"]
    algo_bytes = st["points"] * (3 * 8 + 4) + st["stems"] * 8 + m * 3 * 8 + m * k * (4 + 8)
    kern_s = results[head]["kernel_ms"] / 1e3
    achieved = algo_bytes / kern_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "knn16_screened_c3_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                    kernel=("k_nearest_n_screen32" if head == "screened" else "k_nearest_n"),
                    kernel_ms=round(results[head]["kernel_ms"], 3),
                    algo_bytes_per_launch=int(algo_bytes))

    # construction on the GPU from the points already in HBM (kd_tree_build_device): the
    # second of two builds is timed; its stems must equal the host builder's bit for bit
    build = None
    if not args.no_device_build:
        dpts = torch.from_numpy(pts).to(dev)
        torch.cuda.synchronize()
        ka.ImmutableKdTree.build_on_device(dpts, args.bucket)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        gt = ka.ImmutableKdTree.build_on_device(dpts, args.bucket)
        torch.cuda.synchronize()
        gdt = time.perf_counter() - t0
        same = bool(gt.export_layout().stems().tobytes() == lay.stems().tobytes())
        build = dict(device_ms=round(gdt * 1e3, 1), host_ms=round(host_build_s * 1e3, 1),
                     host_threads=os.cpu_count(), stems_equal_host=same)
        log(f"device build {gdt * 1e3:.1f} ms (host {host_build_s * 1e3:.0f} ms), stems equal: {same}")
        del gt, dpts

    # compulsory footprint: tree bytes + queries + results, once each, over the kernel time
    footprint = (tree.info.device_bytes + m * 3 * 8 + m * k * (4 + 8) + m * 4)
    compulsory = dict(bytes=int(footprint), gbs=round(footprint / kern_s / 1e9, 1))

    # within(r = 0.002), sorted: count pass + scan + fill pass, same tree and queries
    within = None
    if not args.no_within:
        r2 = 0.002 * 0.002
        tree.within_raw(q[:4096], r2, ka.KD_MORTON)  # warm the within kernels
        torch.cuda.synchronize()
        wt = []
        for _ in range(3):
            t0 = time.perf_counter()
            w = tree.within_raw(q, r2, ka.KD_MORTON)
            torch.cuda.synchronize()
            wt.append(time.perf_counter() - t0)
        nres = int(w.offsets[-1].item())
        wdt = float(np.median(wt))
        within = dict(ms=round(wdt * 1e3, 2), results=nres, order="Morton", results_per_query=round(nres / m, 3),
                      queries_per_s=round(m / wdt, 1))
        log(f"within r=0.002 sorted: {wdt * 1e3:.1f} ms, {nres} results")
        del w

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ot = O.Tree.from_layout(np.float64, 3, args.bucket, pts.shape[0], lay.max_stem_level,
                                lay.stems(), lay.extents(), lay.arena())
        ns = min(args.cpu_sample, m)
        qs = np.ascontiguousarray(qh[:ns])
        runs = []
        for _ in range(3):  # median of three runs
            t0 = time.perf_counter()
            ci, cd, cc = ot.nearest_n(qs, k, mode=O.FAITHFUL, threads=threads)
            runs.append(time.perf_counter() - t0)
        dt = float(np.median(runs))
        cpu = dict(value=round(ns / dt, 1), unit="queries/s", cores=threads, kind="port",
                   sample=f"first {ns} of the {m} queries, nearest_n(k=16), Kiddo's scalar "
                          f"query restated in C++ (oracle/kiddo_oracle.cpp), {threads} threads, "
                          f"median of 3 runs",
                   seconds=round(dt, 2))
        # the CPU answers agree with the GPU on the sample's distances
        gd = outs[head].distances[:ns].cpu().numpy()
        cpu["distances_match_gpu"] = bool(np.array_equal(gd.view(np.uint64), cd.view(np.uint64)))
        log(f"cpu baseline {cpu['value'] / 1e6:.2f} Mq/s on {threads} threads ({dt:.1f}s)")

    out = {
        "metric": METRIC,
        "value": round(results[head]["qps"], 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(results[head]["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded: uniform + 5% exact duplicates; uniform queries)",
        "config": {"workload": "knn16-f64-3d-16M+5%dup-16Mq", "points": int(pts.shape[0]),
                   "queries": m, "k": k, "bucket": args.bucket,
                   "variant": "FP32-screened, Morton-ordered" if head == "screened" else "plain FP64",
                   "launch": tuning[head]},
        "plain_f64_qps": round(results["plain"]["qps"], 1) if "plain" in results else None,
        "plain_f64_launch": tuning.get("plain"),
        "plain_f64_input_order_qps": (round(results["plain_input"]["qps"], 1)
                                      if "plain_input" in results else None),
        "screened_over_plain": (round(results["screened"]["qps"] / results["plain"]["qps"], 3)
                                if "screened" in results and "plain" in results else None),
        "screened_over_plain_input_order": (round(results["screened"]["qps"] / results["plain_input"]["qps"], 3)
                                            if "screened" in results and "plain_input" in results else None),
        "kernel_ms": {name: round(results[name]["kernel_ms"], 3) for name in results},
        "variants_equal": parity,
        "oracle_parity_sample": ({"queries": min(args.parity_sample, m), "mode": "canonical",
                                  "equal": oracle_parity} if oracle_parity is not None else None),
        "p50_ms": p50, "p99_ms": p99, "latency_batch": lb,
        "fraction_rescored": (round(own["rescored"] / max(own["points"], 1), 4)
                              if head == "screened" else None),
        "screened_leaves_per_query": round(own["leaves"] / m, 2) if head == "screened" else None,
        "screened_entries_per_query": round(own["points"] / m, 2) if head == "screened" else None,
        "fp32_screen_error": screen_err,
        "leaves_per_query": round(st["leaves"] / m, 2),
        "entries_per_query": round(st["points"] / m, 2),
        "compulsory_footprint": compulsory,
        "tree_build": build,
        "within_r0.002_sorted": within,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if parity is False or (oracle_parity is not None and not all(oracle_parity.values())):
        log("results differ between variants or from the oracle")
        sys.exit(3)


def bench_forest(args, torch, ka):
    """Batched form: `--trees` independent f32 trees, sizes log-uniform in [1000, 262144]
    (seeded noisy unit-sphere clouds, radius jittered by N(0, 0.01)), each with a query set
    of log-uniform size in [256, 65536] from the same law; nearest_n(k=8) with the FP32
    collector (KD_SCREEN_F32 on the forest's FP32 shadow; the plain kernel timed beside it,
    results checked equal) and within(r=0.02, passed as r^2 = 4e-4), one persistent launch
    per batch, replayed from a hipGraph."""
    from concurrent.futures import ThreadPoolExecutor
    from kiddo_amd.datasets import noisy_sphere

    rng = np.random.default_rng(args.seed)
    T = args.trees
    sizes = np.exp(rng.uniform(np.log(1000), np.log(262144), T)).astype(np.int64)
    nq = np.exp(rng.uniform(np.log(256), np.log(65536), T)).astype(np.int64)
    t0 = time.time()
    with ThreadPoolExecutor(16) as ex:
        pts = list(ex.map(lambda i: noisy_sphere(int(sizes[i]), seed=10_000 + i), range(T)))
        qs = list(ex.map(lambda i: noisy_sphere(int(nq[i]), seed=50_000 + i), range(T)))
        log(f"forest data: {sizes.sum()} points, {nq.sum()} queries in {time.time() - t0:.1f}s")
        # bucket size by measurement on every 16th tree and its queries (kd_forest_autotune_bucket)
        bucket, bucket_ms = args.forest_bucket, None
        if bucket == 0:
            t0 = time.time()
            smp = list(range(0, T, 16))
            soffs = np.zeros(len(smp) + 1, np.uint64)
            soffs[1:] = np.cumsum([nq[i] for i in smp])
            sq = torch.from_numpy(np.concatenate([qs[i] for i in smp])).cuda()
            bucket, bucket_ms = ka.Forest.autotune_bucket([pts[i] for i in smp], sq, soffs, 8,
                                                          flags=ka.KD_SCREEN_F32, buckets=(16, 32, 64, 128))
            del sq
            log(f"bucket autotuned on {len(smp)} trees in {time.time() - t0:.1f}s: {bucket} {bucket_ms}")
        t0 = time.time()
        lays = list(ex.map(lambda p: ka.HostLayout(p, bucket, None, 1), pts))
    log(f"host build {time.time() - t0:.1f}s")
    t0 = time.time()
    forest = ka.Forest(lays, shadow_f32=True)
    del lays
    log(f"upload {time.time() - t0:.1f}s")
    offs = np.zeros(T + 1, np.uint64)
    offs[1:] = np.cumsum(nq)
    m = int(offs[-1])
    q = torch.from_numpy(np.concatenate(qs)).cuda()
    forest.plan(q, offs, chunk=256, morton=True)
    k = 8
    forest.set_timing(True)

    def timed_knn(flags):
        chosen = forest.autotune_nearest_n(q, k, flags=flags)
        out = forest._knn_out(k)
        for _ in range(args.warmup):
            forest.nearest_n(q, k, flags=flags | ka.KD_GRAPH, out=out)
        torch.cuda.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            forest.nearest_n(q, k, flags=flags | ka.KD_GRAPH, out=out)
            kms.append(forest.last_kernel_ms())
        torch.cuda.synchronize()
        return time.perf_counter() - t0, kms, forest.query_stats(), chosen, out

    dtp, kmsp, _, chosen_plain, outp = timed_knn(0)
    log(f"forest knn plain: {1e3 * dtp / args.steps:.1f} ms/step, {m * args.steps / dtp / 1e6:.1f} Mq/s, "
        f"launch {chosen_plain}")
    dt, kms, st, chosen, outc = timed_knn(ka.KD_SCREEN_F32)
    same = bool(torch.equal(outp.items, outc.items) and torch.equal(outp.distances, outc.distances)
                and torch.equal(outp.counts, outc.counts))
    del outp, outc
    log(f"forest knn FP32 collector: {1e3 * dt / args.steps:.1f} ms/step, {m * args.steps / dt / 1e6:.1f} Mq/s, "
        f"launch {chosen}, equal to plain {same}, {st}")
    if not same:
        log("collector results differ from the plain kernel's")
        sys.exit(3)
    if not same:
        log("collector results differ from the plain kernel's")
        sys.exit(3)
    # within(r = 0.02): count + scan + fill, timed together, best of 3
    wdt = float("inf")
    for _ in range(3):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        w = forest.within(q, 0.02 * 0.02, ka.KD_UNSORTED)
        torch.cuda.synchronize()
        wdt = min(wdt, time.perf_counter() - t1)
        nres = int(w.offsets[m].item())
        del w
    log(f"forest within: {wdt * 1e3:.1f} ms, {nres} results")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sample = list(range(0, T, max(1, T // 64)))
        cpu_q, cpu_dt = 0, 0.0
        for i in sample:
            ot = O.Tree.build(pts[i], bucket)
            t0 = time.perf_counter()
            ot.nearest_n(qs[i], k, mode=O.FAITHFUL, threads=threads)
            cpu_dt += time.perf_counter() - t0
            cpu_q += len(qs[i])
        cpu = dict(value=round(cpu_q / cpu_dt, 1), unit="queries/s", cores=threads, kind="port",
                   sample=f"{len(sample)} of the {T} trees (every {max(1, T // 64)}th), their "
                          f"{cpu_q} queries, nearest_n(k=8), query time only")
    kern_s = float(np.mean(kms)) / 1e3
    algo = st["points"] * (3 * 4 + 4) + st["stems"] * 4 + m * 3 * 4 + m * k * 8
    tpath = os.path.join(ROOT, "profiles", "forest_4096_f32_traffic.json")
    ftraffic = None
    if os.path.exists(tpath):
        with open(tpath) as fh:
            ftraffic = json.load(fh).get("k_forest_knn_s32", {}).get("bytes_per_launch")
    print(json.dumps({
        "metric": "queries/s, 4096-tree batched kNN k=8 (f32 3-D, persistent hipGraph launch)",
        "value": round(m * args.steps / dt, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * dt / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded noisy unit-sphere clouds)",
        "config": {"workload": "forest-4096-f32-knn8", "trees": T, "points": int(sizes.sum()),
                   "queries": m, "k": k, "launch": chosen, "bucket": bucket,
                   "bucket_autotune_ms": bucket_ms, "variant": "FP32 collector"},
        "plain_qps": round(m * args.steps / dtp, 1), "plain_kernel_ms": round(float(np.mean(kmsp)), 3),
        "plain_launch": chosen_plain, "collector_over_plain": round(dtp / dt, 3), "equal_to_plain": same,
        "within_r0.02": {"ms": round(wdt * 1e3, 2), "results": nres,
                         "results_per_s": round(nres / wdt, 1)},
        "roofline": {"bound": "hbm", "achieved": round(algo / kern_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(algo / kern_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": ftraffic, "kernel": "k_forest_knn_s32", "kernel_ms": round(kern_s * 1e3, 3)},
        "cpu_baseline": cpu,
    }), flush=True)


def bench_serving(args, torch, ka):
    """Serving form: a resident f64 3-D tree of 12,000,000 unit vectors (70% from 200
    Gaussian lat/lon clusters, sigma 0.5 deg, 30% uniform on the sphere; items are the rank of
    a seeded per-item f32 weight) and a u16 4-D palette tree of 65,536 colours. Requests
    arrive as a Poisson process; sizes are log-uniform in [1, 65536]; each is nearest_one,
    nearest_n(5) or best_n_within(r, 5) (r = 0.001 on the sphere, 2000 per channel on the
    palette), on either tree. Reports whole-run queries/s and per-request p50/p99 latency."""
    import threading
    from kiddo_amd.datasets import latlon_clusters, palette_u16, weight_rank_items

    t0 = time.time()
    pts = latlon_clusters(12_000_000, seed=args.seed)
    items = weight_rank_items(len(pts), seed=args.seed + 1)
    pal = palette_u16(65536, seed=args.seed + 2)
    sphere = ka.ImmutableKdTree.new_from_entries(pts, items, args.bucket)
    palette = ka.ImmutableKdTree.new_from_slice(pal, args.bucket)
    log(f"serving trees built and resident in {time.time() - t0:.1f}s")
    rng = np.random.default_rng(args.seed + 3)
    fams = [ka.FAMILY_NEAREST_ONE, ka.FAMILY_NEAREST_N, ka.FAMILY_BEST_N_WITHIN]
    reqs = []
    for i in range(args.requests):
        tree = int(rng.random() < 0.25)
        fam = fams[int(rng.integers(0, 3))]
        m = int(np.exp(rng.uniform(0, np.log(65536))))
        if tree == 0:
            q = latlon_clusters(m, seed=100_000 + i)
            radius = 0.001 ** 2
        else:
            q = palette_u16(m, seed=200_000 + i)
            radius = float(4 * 2000 ** 2)
        flags = ka.KD_MORTON if m >= args.morton_min else 0
        reqs.append((tree, fam, q, 5, radius if fam == ka.FAMILY_BEST_N_WITHIN else np.inf, flags))
    total_q = sum(r[2].shape[0] for r in reqs)
    log(f"{args.requests} requests, {total_q} queries")
    srv = ka.Server([sphere, palette], arena_bytes=1 << 30, slots=4, max_coalesce=1 << 16,
                    max_part=args.max_part)
    # warm-up: a sample of the same request mix (every family on both trees, small and large,
    # with and without Morton ordering), so that one-time set-up (first launches of each
    # kernel, the sort's memory pool) happens before the timed run
    warm = [reqs[i] for i in np.random.default_rng(args.seed + 9).choice(len(reqs), 60, replace=False)]
    for w in range(2):
        for r in [srv.submit(*x) for x in warm]:
            srv.wait(r)

    def run(rate, count):
        """Poisson arrivals at `rate` per second of the first `count` requests; returns the
        latencies (ms), the elapsed seconds and the arrival times."""
        sel = reqs[:count]
        arr = np.cumsum(np.random.default_rng(args.seed + 4 + int(rate)).exponential(1.0 / rate, count))
        tickets = [None] * count
        lat = [None] * count
        start = time.perf_counter()

        def producer():
            for i, (tree, fam, q, k, radius, flags) in enumerate(sel):
                while time.perf_counter() - start < arr[i]:
                    time.sleep(0.0001)
                tickets[i] = srv.submit(tree, fam, q, k, radius, flags)

        th = threading.Thread(target=producer)
        th.start()
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.0001)
            lat[i] = srv.wait(tickets[i])[3]
        end = time.perf_counter()
        th.join()
        return np.array(lat), end - start, arr

    # rate sweep: delivered queries/s and p50/p99 latency as the offered load grows
    sweep = []
    for rate in (float(x) for x in args.rates.split(",") if x):
        n = min(len(reqs), max(400, int(rate * 0.5)))
        sl, sdt, sarr = run(rate, n)
        nq = sum(r[2].shape[0] for r in reqs[:n])
        sweep.append({"rate_per_s": rate, "requests": n, "offered_qps": round(nq / sarr[-1], 1),
                      "delivered_qps": round(nq / sdt, 1), "p50_ms": round(float(np.percentile(sl, 50)), 3),
                      "p99_ms": round(float(np.percentile(sl, 99)), 3)})
        log(f"rate {rate:.0f}/s: {sweep[-1]}")
    st0 = srv.stats()
    lat, elapsed, arrivals = run(args.rate, len(reqs))
    st = srv.stats()
    st = {k: st[k] - st0[k] if k != "peak_arena_bytes" else st[k] for k in st}
    srv.close()
    end, start = elapsed, 0.0
    by_fam = {}
    names = {ka.FAMILY_NEAREST_ONE: "nearest_one", ka.FAMILY_NEAREST_N: "nearest_n5",
             ka.FAMILY_BEST_N_WITHIN: "best_n_within5"}
    for fam in fams:
        sel = np.array([r[1] == fam for r in reqs])
        by_fam[names[fam]] = {"requests": int(sel.sum()), "p50_ms": round(float(np.percentile(lat[sel], 50)), 3),
                              "p99_ms": round(float(np.percentile(lat[sel], 99)), 3)}
    sizes = np.array([r[2].shape[0] for r in reqs])
    by_size = {}
    for lo, hi in ((1, 256), (256, 4096), (4096, 16384), (16384, 65537)):
        sel = (sizes >= lo) & (sizes < hi)
        if sel.any():
            by_size[f"{lo}-{hi - 1}"] = {"requests": int(sel.sum()),
                                         "p50_ms": round(float(np.percentile(lat[sel], 50)), 3),
                                         "p99_ms": round(float(np.percentile(lat[sel], 99)), 3)}
    value = total_q / (end - start)
    log(f"serving: {value / 1e6:.1f} Mq/s, p50 {np.percentile(lat, 50):.2f} ms, p99 {np.percentile(lat, 99):.2f} ms, {st}")
    log(f"by family {by_fam}; by size {by_size}")
    slow = np.argsort(lat)[::-1][:12]
    tail = [{"arrival_ms": round(float(arrivals[i]) * 1e3, 2), "latency_ms": round(float(lat[i]), 2),
             "queries": int(sizes[i]), "family": names[reqs[i][1]]} for i in sorted(slow)]
    log(f"slowest requests: {tail}")
    roof = _serving_replay_roofline(reqs, [sphere, palette], args, torch, ka, fams, names)
    roof["gpu_busy_fraction"] = round(roof["kernel_ms"] / 1e3 / (end - start), 4)
    log(f"serving replay roofline: {roof}")
    print(json.dumps({
        "metric": "queries/s and p50/p99 request latency, continuous batching over resident trees",
        "value": round(value, 1), "unit": "queries/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": round((end - start) * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64 + u16(f64 output)", "data": "synthetic (seeded)",
        "config": {"workload": "serving-12M-f64-sphere+u16-palette", "requests": args.requests,
                   "rate_per_s": args.rate, "queries": int(total_q), "slots": 4,
                   "arena_bytes": 1 << 30, "max_part": args.max_part, "max_coalesce": 1 << 16},
        "rate_sweep": sweep,
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "per_family": by_fam, "per_size": by_size, "slowest": tail, "server": st,
        "offered_qps": round(total_q / arrivals[-1], 1),
        "roofline": roof,
    }), flush=True)


def _serving_replay_roofline(reqs, trees, args, torch, ka, fams, names):
    """Roofline of the serving mix's query kernels. The requests are replayed directly on the
    trees after the server has closed, cut into parts of `max_part` queries and coalesced per
    kind (tree, family, k, radius, flags) in arrival order into launches of at most 65,536
    queries, the server's launches when every part of a kind is pending at once. Each launch
    is timed with the tree's HIP events around its query kernel; its algorithmic bytes are the
    entries it scanned (coordinates + u32 item), the stems it read, its queries and its
    results (u32 item + distance, and a u32 count for the k-row families), from the tree's
    visit counters. `kernel_ms` is the sum over launches; `achieved` = bytes / kernel_ms."""
    groups = {}
    for tree, fam, q, k, radius, flags in reqs:
        for a in range(0, q.shape[0], args.max_part):
            groups.setdefault((tree, fam, k, radius, flags), []).append(q[a:a + args.max_part])
    for t in trees:
        t.set_timing(True)
    tot_ms, tot_bytes, launches, per = 0.0, 0, 0, {}
    for (tree, fam, k, radius, flags), parts in groups.items():
        t = trees[tree]
        launch_q, cur = [], []
        for p in parts:
            if cur and sum(x.shape[0] for x in cur) + p.shape[0] > 1 << 16:
                launch_q.append(np.concatenate(cur))
                cur = []
            cur.append(p)
        if cur:
            launch_q.append(np.concatenate(cur))
        for qh in launch_q:
            K, esz, mm = qh.shape[1], qh.dtype.itemsize, qh.shape[0]
            q = torch.from_numpy(np.ascontiguousarray(qh)).cuda()
            if fam == ka.FAMILY_NEAREST_ONE:
                out = t.nearest_one_raw(q, flags)
                res_bytes = mm * (4 + out.distance.element_size())
            elif fam == ka.FAMILY_NEAREST_N:
                out = t.nearest_n_raw(q, k, math.inf, flags)
                res_bytes = mm * k * (4 + out.distances.element_size()) + mm * 4
            else:
                out = t.best_n_within_raw(q, radius, k, flags)
                res_bytes = mm * k * (4 + out.distances.element_size()) + mm * 4
            ms = t.last_kernel_ms()
            st = t.query_stats()
            b = st["points"] * (K * esz + 4) + st["stems"] * esz + mm * K * esz + res_bytes
            key = ("sphere" if tree == 0 else "palette") + "/" + names[fam]
            g = per.setdefault(key, {"launches": 0, "queries": 0, "kernel_ms": 0.0, "algo_bytes": 0})
            g["launches"] += 1
            g["queries"] += mm
            g["kernel_ms"] += ms
            g["algo_bytes"] += int(b)
            tot_ms += ms
            tot_bytes += b
            launches += 1
    for g in per.values():
        g["achieved_gbs"] = round(g["algo_bytes"] / (g["kernel_ms"] / 1e3) / 1e9, 1)
        g["kernel_ms"] = round(g["kernel_ms"], 3)
    achieved = tot_bytes / (tot_ms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel_ms": round(tot_ms, 3),
            "algo_bytes": int(tot_bytes), "launches": launches,
            "kernel": "replay of the request mix, all query kernels", "per_kind": per}


def _time_knn(tree, q, k, flags, steps, warmup, torch, out=None):
    for _ in range(warmup):
        out = tree.nearest_n_raw(q, k, math.inf, flags, out=out)
    torch.cuda.synchronize()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out = tree.nearest_n_raw(q, k, math.inf, flags, out=out)
        kms.append(tree.last_kernel_ms())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, float(np.mean(kms)), out


def _row_roofline(st, m, k, K, esz, osz, kms, kernel, traffic=None):
    """HBM roofline of one k-row query launch from the tree's visit counters (the last call's
    query_stats): every entry scanned (K coordinates + u32 item), every stem read, the
    queries, and the rows written (k x (u32 item + distance)); achieved = those bytes / the
    kernel's HIP-event time."""
    algo = st["points"] * (K * esz + 4) + st["stems"] * esz + m * K * esz + m * k * (4 + osz)
    ach = algo / (kms / 1e3) / 1e9
    return {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kernel,
            "kernel_ms": round(kms, 3), "algo_bytes_per_launch": int(algo)}


def _traffic(name):
    """Measured HBM bytes per launch by kernel (rocprofv3 FETCH_SIZE / WRITE_SIZE summaries
    committed under profiles/), or {} when absent."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def bench_mixture(args, torch, ka):
    """ImmutableKdTree<f32, 3>: 10,000,000 points from a 64-component isotropic Gaussian
    mixture in the unit cube (centres uniform, sigma 0.01, clipped), bucket 32; 10,000,000
    queries from the same mixture; nearest_n(16) with the FP32 collector (KD_SCREEN_F32 on
    the tree's FP32 shadow), next to the plain f32 kernel in the same Morton order (both
    autotuned, results checked equal); and within(r = 0.005, passed as r^2 = 2.5e-5)
    unsorted, whose ~1.8e10 results are produced in query chunks that fit
    `--result-budget` bytes (count pass, then one fill pass per chunk)."""
    from kiddo_amd.datasets import gaussian_mixture
    n = m = 10_000_000
    centres = np.random.default_rng(args.seed ^ 0x5EED).random((64, 3))
    pts = gaussian_mixture(n, 64, 0.01, 3, seed=args.seed, centres=centres)
    qh = gaussian_mixture(m, 64, 0.01, 3, seed=args.seed + 1, centres=centres)
    tree = ka.ImmutableKdTree.new_from_slice(pts, args.bucket, shadow_f32=True)
    tree.set_timing(True)
    q = torch.from_numpy(qh).cuda()
    # the plain f32 kernel, autotuned, in Morton order
    chosen_plain = tree.autotune_nearest_n(q, 16, math.inf, ka.KD_MORTON)
    dtp, kmsp, outp = _time_knn(tree, q, 16, ka.KD_MORTON, args.steps, args.warmup, torch)
    stp = tree.query_stats()
    log(f"mixture knn16 plain: {dtp * 1e3:.1f} ms/step ({m / dtp / 1e6:.1f} Mq/s), kernel {kmsp:.1f} ms")
    # the FP32 collector, autotuned
    flags = ka.KD_MORTON | ka.KD_SCREEN_F32
    chosen = tree.autotune_nearest_n(q, 16, math.inf, flags)
    dt, kms, outc = _time_knn(tree, q, 16, flags, args.steps, args.warmup, torch)
    st = tree.query_stats()
    same = bool(torch.equal(outp.items, outc.items) and torch.equal(outp.distances, outc.distances)
                and torch.equal(outp.counts, outc.counts))
    redos = tree.query_redos()
    del outp, outc
    log(f"mixture knn16 collector: {dt * 1e3:.1f} ms/step ({m / dt / 1e6:.1f} Mq/s), kernel {kms:.1f} ms, "
        f"equal to plain {same}, redone {redos}, {st}")
    if not same:
        log("collector results differ from the plain kernel's")
        sys.exit(3)
    # within, chunked by result memory; launch shapes of both passes autotuned on a sample
    r2 = 0.005 ** 2
    wflags = ka.KD_UNSORTED | ka.KD_MORTON
    wshape = tree.autotune_within(q, r2, wflags)
    log(f"autotuned within shapes: {wshape}")
    cap = max(1, args.result_budget // 8)  # u32 item + f32 distance per result
    items = torch.empty(cap, dtype=torch.uint32, device="cuda")
    dists = torch.empty(cap, dtype=torch.float32, device="cuda")

    def within_pass(qq):
        """count pass over the batch, then fill passes over chunks that each fit the result
        budget; chunks are contiguous ranges of the batch's Morton order (kd_morton_order),
        so each fill launch sees spatially coherent queries, and each chunk's results are a
        CSR over its queries in that order (query perm[j] of the batch for row j - a)"""
        mm = qq.shape[0]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        offs = torch.empty(mm + 1, dtype=torch.int64, device="cuda")
        ka._check(ka.lib().kd_within_count(tree.handle, qq.data_ptr(), mm, r2, wflags, offs.data_ptr(),
                                           ka._stream_ptr(None)), "kd_within_count")
        count_kernel_ms = tree.last_kernel_ms()
        st = tree.query_stats()
        t_count = time.perf_counter() - t0
        perm = tree.morton_order(qq).view(torch.int32).long()
        cum = torch.zeros(mm + 1, dtype=torch.int64, device="cuda")
        torch.cumsum((offs[1:] - offs[:-1])[perm], 0, out=cum[1:])
        cum_h = cum.cpu().numpy()
        tot = int(cum_h[-1])
        a, chunks, fill_kernel_ms = 0, 0, 0.0
        while a < mm:
            b = int(np.searchsorted(cum_h, cum_h[a] + cap, side="right")) - 1
            b = max(min(b, mm), a + 1)
            sub = (cum[a:b + 1] - cum[a]).contiguous()
            qs = qq.index_select(0, perm[a:b])
            ka._check(ka.lib().kd_within_fill(tree.handle, qs.data_ptr(), b - a, r2, wflags,
                                              sub.data_ptr(), items.data_ptr(), dists.data_ptr(),
                                              ka._stream_ptr(None)), "kd_within_fill")
            fill_kernel_ms += tree.last_kernel_ms()
            chunks += 1
            a = b
        torch.cuda.synchronize()
        return dict(t_within=time.perf_counter() - t0, t_count=t_count, count_kernel_ms=count_kernel_ms,
                    fill_kernel_ms=fill_kernel_ms, total=tot, chunks=chunks, stats=st)

    within_pass(q[:200_000])  # warm-up: first launches of the pipeline's kernels
    W = within_pass(q)
    t_within, t_count, total, chunks = W["t_within"], W["t_count"], W["total"], W["chunks"]
    count_kernel_ms, fill_kernel_ms, wst = W["count_kernel_ms"], W["fill_kernel_ms"], W["stats"]
    # algorithmic bytes: every entry scanned (3 f32 coordinates), the item of every result
    # (read) and the result itself (u32 item + f32 distance, written), stems read, queries;
    # the count pass reads the same entries and stems and writes one u64 count per query
    fill_bytes = wst["points"] * 12 + total * (4 + 8) + wst["stems"] * 4 + m * 12
    count_bytes = wst["points"] * 12 + wst["stems"] * 4 + m * (12 + 8)
    log(f"mixture within: {total} results in {chunks} chunks, count {t_count * 1e3:.0f} ms, "
        f"total {t_within * 1e3:.0f} ms ({total / t_within / 1e9:.2f} G results/s)")
    algo = st["points"] * 16 + st["stems"] * 4 + m * 12 + m * 16 * 8
    # measured HBM traffic of these kernels (rocprofv3 FETCH_SIZE / WRITE_SIZE, committed)
    tr = {}
    tpath = os.path.join(ROOT, "profiles", "mixture_10M_f32_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
    print(json.dumps({
        "metric": "queries/s, 10M-pt f32 Gaussian-mixture kNN k=16; within(r=0.005) results/s",
        "value": round(m / dt, 1), "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded mixture)",
        "config": {"workload": "mixture-10M-f32-knn16+within0.005", "points": n, "queries": m,
                   "k": 16, "launch": chosen, "variant": "FP32 collector, Morton-ordered"},
        "plain_f32_qps": round(m / dtp, 1), "plain_f32_kernel_ms": round(kmsp, 3), "plain_f32_launch": chosen_plain,
        "collector_over_plain": round(dtp / dt, 3), "equal_to_plain": same, "redone": redos,
        "within_r0.005_unsorted": {"results": total, "chunks": chunks,
                                   "result_budget_bytes": args.result_budget,
                                   "launch": wshape,
                                   "count_ms": round(t_count * 1e3, 1),
                                   "count_kernel_ms": round(count_kernel_ms, 2),
                                   "fill_kernel_ms": round(fill_kernel_ms, 2),
                                   "total_ms": round(t_within * 1e3, 1),
                                   "results_per_s": round(total / t_within, 1),
                                   "entries_per_query": round(wst["points"] / m, 1),
                                   "roofline_fill": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                                                     "achieved": round(fill_bytes / (fill_kernel_ms / 1e3) / 1e9, 1),
                                                     "frac": round(fill_bytes / (fill_kernel_ms / 1e3) / 1e9
                                                                   / HBM_PEAK_GBS, 4),
                                                     "algo_bytes": int(fill_bytes),
                                                     "traffic": tr.get("k_within_fill_group_4lanes"
                                                                       if wshape.get("fill_group") == 4
                                                                       else "k_within_fill_group", {}).get("bytes_per_pass"),
                                                     "kernel": "k_within_fill_group"},
                                   "roofline_count": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                                                      "achieved": round(count_bytes / (count_kernel_ms / 1e3) / 1e9, 1),
                                                      "frac": round(count_bytes / (count_kernel_ms / 1e3) / 1e9
                                                                    / HBM_PEAK_GBS, 4),
                                                      "algo_bytes": int(count_bytes),
                                                      "traffic": tr.get("k_within_count_group", {}).get("bytes_per_launch"),
                                                      "kernel": "k_within_count_group"}},
        "roofline": {"bound": "hbm", "achieved": round(algo / (kms / 1e3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(algo / (kms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": tr.get("k_nearest_n_screen32<float,3,16>", {}).get("bytes_per_launch"),
                     "kernel": "k_nearest_n_screen32<float,3,16>", "kernel_ms": round(kms, 3)},
    }), flush=True)


def bench_small(args, torch, ka):
    """ImmutableKdTree<f64, 3>: 100,000 uniform points, bucket 32, 100,000 uniform queries;
    nearest_one and nearest_n(8) (the reference's CPU-runnable case)."""
    n = m = 100_000
    pts = np.random.default_rng(args.seed).random((n, 3))
    qh = np.random.default_rng(args.seed + 1).random((m, 3))
    tree = ka.ImmutableKdTree.new_from_slice(pts, args.bucket)
    tree.set_timing(True)
    q = torch.from_numpy(qh).cuda()
    res = {}
    for name, k in (("nearest_n8", 8),):
        dt, kms, _ = _time_knn(tree, q, k, ka.KD_MORTON, max(args.steps, 20), args.warmup, torch)
        res[name] = {"ms": round(dt * 1e3, 4), "kernel_ms": round(kms, 4), "qps": round(m / dt, 1)}
        # the 100k-point tree (2.8 MB) stays in the L2: no HBM traffic figure
        roof = _row_roofline(tree.query_stats(), m, k, 3, 8, 8, kms, "k_nearest_n<double,3,8>")
    for _ in range(args.warmup):
        tree.nearest_one_raw(q, ka.KD_MORTON)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(args.steps, 20)):
        tree.nearest_one_raw(q, ka.KD_MORTON)
    torch.cuda.synchronize()
    dt1 = (time.perf_counter() - t0) / max(args.steps, 20)
    res["nearest_one"] = {"ms": round(dt1 * 1e3, 4), "qps": round(m / dt1, 1)}
    log(f"small: {res}")
    print(json.dumps({
        "metric": "queries/s, 100k-pt f64 3D nearest_one and nearest_n(8)",
        "value": res["nearest_n8"]["qps"], "unit": "queries/s", "n_gpus": 1,
        "steps": max(args.steps, 20), "warmup": args.warmup, "ms_per_step": res["nearest_n8"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded uniform)", "config": {"workload": "small-100k-f64", "points": n,
                                                         "queries": m},
        "families": res, "roofline": roof,
    }), flush=True)


def bench_periodic(args, torch, ka):
    """Periodic boundary conditions (query/periodic.rs) on the unit box: 16M uniform f64
    points, 16M uniform queries, periodic nearest_n(16), nearest_one and within(r = 0.002),
    Morton-ordered, next to the Cartesian nearest_n(16) of the same tree and queries. CPU
    baseline: the oracle's periodic nearest_n on a sample of the queries."""
    n, m, k = args.points, args.queries, args.k
    pts = np.random.default_rng(args.seed).random((n, 3))
    qh = np.random.default_rng(args.seed + 1).random((m, 3))
    box = np.ones(3)
    tree = ka.ImmutableKdTree.new_from_slice(pts, args.bucket)
    tree.set_timing(True)
    q = torch.from_numpy(qh).cuda()
    M = ka.KD_MORTON

    def timed(fn, reps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms = []
        for _ in range(reps):
            fn()
            kms.append(tree.last_kernel_ms())
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps, float(np.mean(kms))

    res = {}
    dt, kms = timed(lambda: tree.periodic_nearest_n_raw(q, box, k, flags=M), args.steps)
    st = tree.query_stats()
    res[f"periodic_nearest_n{k}"] = {"ms": round(dt * 1e3, 2), "kernel_ms": round(kms, 2),
                                    "qps": round(m / dt, 1), "leaves_per_query": round(st["leaves"] / m, 2)}
    kname = f"k_periodic_nearest_n<double,3,{k}>"
    roof = _row_roofline(st, m, k, 3, 8, 8, kms, kname,
                         _traffic("periodic_16M_f64_traffic.json").get(kname, {}).get("bytes_per_launch"))
    dtc, kmsc = timed(lambda: tree.nearest_n_raw(q, k, flags=M), args.steps)
    res[f"cartesian_nearest_n{k}"] = {"ms": round(dtc * 1e3, 2), "kernel_ms": round(kmsc, 2),
                                     "qps": round(m / dtc, 1)}
    dt1, _ = timed(lambda: tree.periodic_nearest_one_raw(q, box, M), args.steps)
    res["periodic_nearest_one"] = {"ms": round(dt1 * 1e3, 2), "qps": round(m / dt1, 1)}
    t0 = time.perf_counter()
    w = tree.periodic_within_raw(q, box, 0.002 ** 2, M)
    torch.cuda.synchronize()
    dtw = time.perf_counter() - t0
    res["periodic_within_r0.002"] = {"ms": round(dtw * 1e3, 2), "results": int(w.offsets[-1].item())}
    del w
    log(f"periodic: {res}")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ot = O.Tree.build(pts, args.bucket)
        ns = min(args.cpu_sample, m)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ot.periodic_nearest_n(qh[:ns], box, k, mode=O.FAITHFUL, threads=threads)
            runs.append(time.perf_counter() - t0)
        cdt = float(np.median(runs))
        cpu = dict(value=round(ns / cdt, 1), unit="queries/s", cores=threads, kind="port",
                   sample=f"first {ns} of the {m} queries, periodic nearest_n(k={k}), Kiddo's "
                          f"periodic query restated in C++ (oracle/kiddo_oracle.cpp), {threads} "
                          f"threads, median of 3 runs", seconds=round(cdt, 2))
    print(json.dumps({
        "metric": f"queries/s, {n // 1_000_000}M-pt f64 3D periodic (unit box) nearest_n({k})",
        "value": res[f"periodic_nearest_n{k}"]["qps"], "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": res[f"periodic_nearest_n{k}"]["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded uniform in the unit box)",
        "config": {"workload": f"periodic-f64-3d-{n}-{m}q", "points": n, "queries": m, "k": k,
                   "box": [1.0, 1.0, 1.0]},
        "families": res, "roofline": roof, "cpu_baseline": cpu,
    }), flush=True)


def bench_metrics(args, torch, ka):
    """Other distance metrics on the 16M-point f64 tree of the periodic workload (no
    duplicates): nearest_n(16) and within under SquaredEuclidean, Manhattan, Chebyshev and
    Minkowski<3>, Morton-ordered thread-per-query kernels; within at the median k-th
    distance of the metric, so a median query holds about k entries. CPU baseline: the oracle's
    nearest_n(16) under Manhattan on a sample."""
    n, m, k = args.points, args.queries, args.k
    pts = np.random.default_rng(args.seed).random((n, 3))
    qh = np.random.default_rng(args.seed + 1).random((m, 3))
    tree = ka.ImmutableKdTree.new_from_slice(pts, args.bucket)
    tree.set_timing(True)
    q = torch.from_numpy(qh).cuda()
    flags = {"sqeuclidean": 0, "manhattan": ka.KD_METRIC_MANHATTAN, "chebyshev": ka.KD_METRIC_CHEBYSHEV,
             "minkowski3": ka.KD_METRIC_MINKOWSKI(3)}
    tr = _traffic("metrics_16M_f64_traffic.json")
    res = {}
    for name, f in flags.items():
        for _ in range(args.warmup):
            tree.nearest_n_raw(q, k, math.inf, f | ka.KD_MORTON)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms = []
        for _ in range(args.steps):
            out = tree.nearest_n_raw(q, k, math.inf, f | ka.KD_MORTON)
            kms.append(tree.last_kernel_ms())
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        st = tree.query_stats()
        kth = float(out.distances[:, -1].float().median().item())
        kname = f"k_nearest_n<double,3,{k},{name}>"
        res[name] = {"nearest_n_ms": round(dt * 1e3, 2), "kernel_ms": round(float(np.mean(kms)), 2),
                     "qps": round(m / dt, 1), "leaves_per_query": round(st["leaves"] / m, 2),
                     "median_kth_distance": kth,
                     "roofline": _row_roofline(st, m, k, 3, 8, 8, float(np.mean(kms)), kname,
                                               tr.get(kname, {}).get("bytes_per_launch"))}
        t0 = time.perf_counter()
        w = tree.within_raw(q, kth, f | ka.KD_MORTON)
        torch.cuda.synchronize()
        res[name]["within_at_median_kth_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
        res[name]["within_results"] = int(w.offsets[-1].item())
        del w, out
    log(f"metrics: {res}")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.build()
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        ot = O.Tree.build(pts, args.bucket).set_metric(ka.KD_METRIC_MANHATTAN)
        ns = min(args.cpu_sample, m)
        runs = []
        for _ in range(3):
            t0 = time.perf_counter()
            ot.nearest_n(qh[:ns], k, mode=O.FAITHFUL, threads=threads)
            runs.append(time.perf_counter() - t0)
        cdt = float(np.median(runs))
        cpu = dict(value=round(ns / cdt, 1), unit="queries/s", cores=threads, kind="port",
                   sample=f"first {ns} of the {m} queries, Manhattan nearest_n(k={k}), Kiddo's query "
                          f"restated in C++ (oracle/kiddo_oracle.cpp), {threads} threads, median of 3 runs",
                   seconds=round(cdt, 2))
    print(json.dumps({
        "metric": f"queries/s, {n // 1_000_000}M-pt f64 3D Manhattan nearest_n({k})",
        "value": res["manhattan"]["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["manhattan"]["nearest_n_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded uniform)",
        "config": {"workload": f"metrics-f64-3d-{n}-{m}q", "points": n, "queries": m, "k": k},
        "by_metric": res, "roofline": res["manhattan"]["roofline"], "cpu_baseline": cpu,
    }), flush=True)


if __name__ == "__main__":
    main()
