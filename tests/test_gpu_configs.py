"""GPU parity at the full sizes of the five benchmark configurations, through the C ABI,
against the oracle on the same seeded inputs (the generators and seeds bench.py uses).

The whole batch always runs on the GPU. The oracle answers a sample of the queries (its
single-threaded build of the full tree takes seconds, its queries run on the host's
cores), compared bit for bit: items, result order (distance, then lower item), distance
bits, counts. Properties that need no oracle are checked over the whole batch:
  * the FP32-screened kernel returns exactly the plain FP64 kernel's results;
  * within and nearest_n agree: for every query, the number of its k nearest within the
    radius is min(within count, k), and (sorted within, C3) a query with fewer than k
    results within the radius has exactly the first entries of its k nearest, in order;
  * the host builder's stems and leaf sizes equal the oracle's at full size.

Sizes: C1 100k f64 (tests/test_gpu_edge.py runs it whole); C2 10M f32 Gaussian mixture,
10M queries; C3 16M + 5% duplicates f64, 16M queries; C4 4096 f32 trees of 1k..262k
points with 256..65,536 queries each; C5 12M f64 sphere + u16 palette through the server.
Radii are squared distances (Euclidean r passed as r^2, as Kiddo's within takes them).
"""
import math
import threading
import time

import numpy as np
import pytest

from engines import assert_same_bits

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


def _announce(*what):
    print("case:", *what, f"(t={time.strftime('%H:%M:%S')})", flush=True)


def _rows(t, rows):
    """t[rows] on the device, copied to the host (uint32 through an int32 view: torch
    indexes few uint32 kernels)."""
    import torch
    idx = torch.from_numpy(np.asarray(rows, np.int64)).to(t.device)
    if t.dtype == torch.uint32:
        return t.view(torch.int32)[idx].cpu().numpy().view(np.uint32)
    return t[idx].cpu().numpy()


def _check_sample(res, rows, ref, what):
    got = (_rows(res.items, rows), _rows(res.distances, rows), _rows(res.counts, rows))
    for a, b, nm in zip(got, ref, ("items", "dists", "counts")):
        assert_same_bits(a, b, f"{what} {nm}")


def _csr_slices(offsets, rows):
    """Flat indices of the results of `rows` in a CSR batch, and their per-row lengths."""
    offsets = np.asarray(offsets, np.int64)
    s, e = offsets[rows], offsets[np.asarray(rows) + 1]
    lens = e - s
    start = np.repeat(s - np.concatenate([[0], np.cumsum(lens)[:-1]]), lens)
    return start + np.arange(int(lens.sum())), lens


def _within_knn_consistent(knn, offsets, r2, k, what):
    """For every query: #(k nearest with d <= r2) == min(within count, k)."""
    import torch
    counts_w = (offsets[1:] - offsets[:-1]).to(torch.int64)
    inside = (knn.distances <= r2).sum(dim=1).to(torch.int64)
    expect = torch.clamp(counts_w, max=k)
    bad = int((inside != expect).sum().item())
    assert bad == 0, f"{what}: {bad} queries where kNN and within disagree"


def _within_prefix_equals_knn(knn, w, k, what):
    """Sorted within: a query with c < k results has exactly knn[:c] (items and distance
    bits, same order), over the whole batch."""
    import torch
    offs = w.offsets
    c = offs[1:] - offs[:-1]
    rows = torch.nonzero((c > 0) & (c < k)).flatten()
    cr = c[rows]
    qid = torch.repeat_interleave(rows, cr)
    start = torch.repeat_interleave(torch.cumsum(cr, 0) - cr, cr)
    j = torch.arange(qid.numel(), device=qid.device) - start
    flat = offs[qid] + j
    wi = w.items.view(torch.int32)[flat]
    ki = knn.items.view(torch.int32)[qid, j]
    wd = w.distances.view(torch.int64)[flat]
    kd = knn.distances.view(torch.int64)[qid, j]
    bad = int(((wi != ki) | (wd != kd)).sum().item())
    assert bad == 0, f"{what}: {bad} within results differ from the k-NN prefix"
    return int(rows.numel())


def _sample(m, n, seed):
    return np.sort(np.random.default_rng(seed).choice(m, n, replace=False))


def _layout_equal(lay, ot, what):
    assert lay.max_stem_level == ot.max_stem_level, what
    assert lay.stems().tobytes() == ot.stems().tobytes(), f"{what}: stems differ"
    assert np.array_equal(lay.extents()[:, 1], ot.extents()[:, 1]), f"{what}: leaf sizes differ"


def test_c3_headline_16m_f64_full_size(ka, oracle_mod, cuda):
    """ImmutableKdTree<f64, 3>: 16,000,000 uniform points + 800,000 exact copies, bucket 32;
    16,000,000 uniform queries; nearest_n(16) screened (Morton, autotuned) and plain (input
    order); within(r = 0.002 -> 4e-6) sorted."""
    torch = cuda
    from kiddo_amd.datasets import uniform_with_duplicates
    _announce("C3 data")
    pts = uniform_with_duplicates(16_000_000, 0.05, seed=1)
    m, k, r2 = 16_000_000, 16, 0.002 ** 2
    qh = np.random.default_rng(2).random((m, 3))
    lay = ka.HostLayout(pts, 32)
    _announce("C3 oracle build")
    ot = oracle_mod.Tree.build(pts, 32)
    _layout_equal(lay, ot, "C3")
    tree = ka.ImmutableKdTree(lay.desc, shadow_f32=True)
    del lay
    q = torch.from_numpy(qh).cuda()
    _announce("C3 screened")
    sflags = ka.KD_SCREEN_F32 | ka.KD_MORTON
    tree.autotune_nearest_n(q, k, math.inf, sflags)
    rs = tree.nearest_n_raw(q, k, math.inf, sflags)
    _announce("C3 plain")
    tree.tune(1, 0, 0, 0)
    rp = tree.nearest_n_raw(q, k, math.inf, 0)
    torch.cuda.synchronize()
    assert torch.equal(rs.items, rp.items), "screened items != plain items"
    assert torch.equal(rs.distances.view(torch.int64), rp.distances.view(torch.int64))
    assert torch.equal(rs.counts, rp.counts)
    del rp
    rows = _sample(m, 200_000, 11)
    _announce("C3 oracle nearest_n")
    ref = ot.nearest_n(qh[rows], k, threads=16)
    _check_sample(rs, rows, ref, "C3 nearest_n(16)")
    _announce("C3 within")
    w = tree.within_raw(q, r2, ka.KD_MORTON)
    torch.cuda.synchronize()
    _within_knn_consistent(rs, w.offsets, r2, k, "C3")
    assert _within_prefix_equals_knn(rs, w, k, "C3") > 0.3 * m
    offs = w.offsets.cpu().numpy()
    oo, oi, od = ot.within(qh[rows], r2, threads=16)
    flat, lens = _csr_slices(offs, rows)
    assert_same_bits(lens.astype(np.uint64), np.diff(oo), "C3 within counts")
    assert_same_bits(_rows(w.items, flat), oi, "C3 within items")
    assert_same_bits(_rows(w.distances, flat), od, "C3 within dists")
    total = int(offs[-1])
    assert 0.4 * m < total < 0.8 * m, total  # about 0.56 results per query


def test_c2_mixture_10m_f32_full_size(ka, oracle_mod, cuda):
    """ImmutableKdTree<f32, 3>: 10,000,000 points from 64 Gaussians (sigma 0.01), bucket 32;
    10,000,000 queries from the same mixture; nearest_n(16) on every query; within(r = 0.005
    -> 2.5e-5) unsorted: counts of every query, results of a sample as multisets."""
    torch = cuda
    from kiddo_amd.datasets import gaussian_mixture
    _announce("C2 data")
    n = m = 10_000_000
    k, r2 = 16, 0.005 ** 2
    centres = np.random.default_rng(1 ^ 0x5EED).random((64, 3))
    pts = gaussian_mixture(n, 64, 0.01, 3, seed=1, centres=centres)
    qh = gaussian_mixture(m, 64, 0.01, 3, seed=2, centres=centres)
    lay = ka.HostLayout(pts, 32)
    _announce("C2 oracle build")
    ot = oracle_mod.Tree.build(pts, 32)
    _layout_equal(lay, ot, "C2")
    tree = ka.ImmutableKdTree(lay.desc)
    del lay
    q = torch.from_numpy(qh).cuda()
    _announce("C2 nearest_n")
    tree.autotune_nearest_n(q, k, math.inf, ka.KD_MORTON)
    r = tree.nearest_n_raw(q, k, math.inf, ka.KD_MORTON)
    rows = _sample(m, 200_000, 12)
    ref = ot.nearest_n(qh[rows], k, threads=16)
    _check_sample(r, rows, ref, "C2 nearest_n(16)")
    _announce("C2 within count")
    offs = torch.empty(m + 1, dtype=torch.int64, device="cuda")
    wflags = ka.KD_UNSORTED | ka.KD_MORTON
    ka._check(ka.lib().kd_within_count(tree.handle, q.data_ptr(), m, r2, wflags, offs.data_ptr(),
                                       ka._stream_ptr(None)), "kd_within_count")
    torch.cuda.synchronize()
    _within_knn_consistent(r, offs, r2, k, "C2")
    offs_h = offs.cpu().numpy()
    oc = ot.within_count(qh[rows], r2, threads=16)
    assert_same_bits(np.diff(offs_h)[rows].astype(np.uint64), oc, "C2 within counts")
    assert offs_h[-1] > 1000 * m  # about 1.8e3 results per query
    _announce("C2 within fill (sample)")
    sub = rows[::10]
    qs = qh[sub]
    w = tree.within_raw(qs, r2, wflags)
    wo = w.offsets.cpu().numpy()
    assert_same_bits(np.diff(wo).astype(np.uint64), oc[::10], "C2 within sample counts")
    oo, oi, od = ot.within(qs, r2, threads=16)
    gi, gd = w.items.cpu().numpy(), w.distances.cpu().numpy()
    qid = np.repeat(np.arange(len(sub)), np.diff(wo))
    key = np.lexsort((gi, gd.view(np.uint32), qid))  # distances are non-negative
    assert_same_bits(gi[key], oi, "C2 within items (multiset per query)")
    assert_same_bits(gd[key], od, "C2 within dists (multiset per query)")


def test_c4_forest_4096_trees_full_size(ka, oracle_mod, cuda):
    """4096 f32 trees, sizes log-uniform in [1000, 262144] (noisy unit-sphere clouds), each
    with a query set of log-uniform size in [256, 65536]; nearest_n(8) and within(r = 0.02
    -> 4e-4) in one persistent launch (hipGraph replay); every 64th tree against the oracle."""
    torch = cuda
    from concurrent.futures import ThreadPoolExecutor
    from kiddo_amd.datasets import noisy_sphere
    _announce("C4 data")
    T, k, r2 = 4096, 8, 0.02 ** 2
    rng = np.random.default_rng(1)
    sizes = np.exp(rng.uniform(np.log(1000), np.log(262144), T)).astype(np.int64)
    nq = np.exp(rng.uniform(np.log(256), np.log(65536), T)).astype(np.int64)
    with ThreadPoolExecutor(16) as ex:
        pts = list(ex.map(lambda i: noisy_sphere(int(sizes[i]), seed=10_000 + i), range(T)))
        qs = list(ex.map(lambda i: noisy_sphere(int(nq[i]), seed=50_000 + i), range(T)))
        _announce("C4 host build")
        lays = list(ex.map(lambda p: ka.HostLayout(p, 32, None, 1), pts))
    forest = ka.Forest(lays)
    del lays
    offs = np.zeros(T + 1, np.uint64)
    offs[1:] = np.cumsum(nq)
    m = int(offs[-1])
    q = torch.from_numpy(np.concatenate(qs)).cuda()
    forest.plan(q, offs, chunk=256, morton=True)
    _announce("C4 nearest_n")
    forest.autotune_nearest_n(q, k)
    out = forest._knn_out(k)
    for _ in range(3):  # graph capture, then replays; each replay is checked
        forest.nearest_n(q, k, flags=ka.KD_GRAPH, out=out)
        torch.cuda.synchronize()
        items, dists, counts = out.items.cpu().numpy(), out.distances.cpu().numpy(), out.counts.cpu().numpy()
        if _ == 0:
            first = (items.copy(), dists.copy(), counts.copy())
        else:
            for a, b in zip((items, dists, counts), first):
                assert_same_bits(a, b, "C4 graph replay")
    _announce("C4 within")
    w = forest.within(q, r2, ka.KD_UNSORTED)
    wo = w.offsets.cpu().numpy()
    wi, wd = w.items.cpu().numpy(), w.distances.cpu().numpy()
    _within_knn_consistent(out, w.offsets, r2, k, "C4")
    _announce("C4 oracle (every 64th tree)")
    for i in range(0, T, 64):
        ot = oracle_mod.Tree.build(pts[i], 32)
        a, b = int(offs[i]), int(offs[i + 1])
        ref = ot.nearest_n(qs[i], k, threads=16)
        for x, y, nm in zip((items[a:b], dists[a:b], counts[a:b]), ref, ("items", "dists", "counts")):
            assert_same_bits(x, y, f"C4 tree {i} nearest_n(8) {nm}")
        oo, oi, od = ot.within(qs[i], r2, threads=16)
        s, e = int(wo[a]), int(wo[b])
        assert_same_bits((wo[a:b + 1] - wo[a]).astype(np.uint64), oo, f"C4 tree {i} within offsets")
        gi, gd = wi[s:e], wd[s:e]
        qid = np.repeat(np.arange(b - a), np.diff(wo[a:b + 1]))
        key = np.lexsort((gi, gd.view(np.uint32), qid))
        assert_same_bits(gi[key], oi, f"C4 tree {i} within items (multiset)")
        assert_same_bits(gd[key], od, f"C4 tree {i} within dists (multiset)")


def test_c5_serving_12m_sphere_and_palette_full_size(ka, oracle_mod, cuda):
    """Resident ImmutableKdTree<f64, 3> of 12,000,000 unit vectors (70% in 200 lat/lon
    clusters of sigma 0.5 deg, 30% uniform; items = rank of a seeded per-item f32 weight)
    and a u16 4-D palette of 65,536 colours, served by the continuous-batching loop:
    nearest_one, nearest_n(5) and best_n_within(r, 5) requests of log-uniform sizes in
    [1, 65536] submitted concurrently; every request's results equal the same request
    served alone and the oracle's."""
    from kiddo_amd.datasets import latlon_clusters, palette_u16, weight_rank_items
    _announce("C5 data")
    pts = latlon_clusters(12_000_000, seed=1)
    items = weight_rank_items(len(pts), seed=2)
    pal = palette_u16(65536, seed=3)
    sphere = ka.ImmutableKdTree.new_from_entries(pts, items, 32)
    palette = ka.ImmutableKdTree.new_from_slice(pal, 32)
    _announce("C5 oracle build")
    refs = [oracle_mod.Tree.build(pts, 32, items), oracle_mod.Tree.build(pal, 32)]
    rng = np.random.default_rng(4)
    fams = [ka.FAMILY_NEAREST_ONE, ka.FAMILY_NEAREST_N, ka.FAMILY_BEST_N_WITHIN]
    reqs = []
    for i in range(160):
        tree = int(rng.random() < 0.25)
        fam = fams[i % 3]
        mq = int(np.exp(rng.uniform(0, np.log(65536))))
        if tree == 0:
            qq = latlon_clusters(mq, seed=100_000 + i)
            radius = 0.001 ** 2
        else:
            qq = palette_u16(mq, seed=200_000 + i)
            radius = float(4 * 2000 ** 2)
        flags = ka.KD_MORTON if mq >= 4096 else 0
        reqs.append((tree, fam, qq, 5, radius if fam == ka.FAMILY_BEST_N_WITHIN else np.inf, flags))
    _announce("C5 alone", sum(r[2].shape[0] for r in reqs), "queries")
    srv1 = ka.Server([sphere, palette], arena_bytes=1 << 30, slots=1, max_coalesce=1)
    alone = [srv1.wait(srv1.submit(*r)) for r in reqs]
    srv1.close()
    _announce("C5 mixed")
    srv = ka.Server([sphere, palette], arena_bytes=1 << 30, slots=4, max_coalesce=1 << 18)
    tickets = [None] * len(reqs)
    order = np.random.default_rng(5).permutation(len(reqs))

    def client(idx):
        for j in idx:
            tickets[j] = srv.submit(*reqs[j])

    th = [threading.Thread(target=client, args=(order[h::2],)) for h in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mixed = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    srv.close()
    assert st["done"] == len(reqs)
    _announce("C5 oracle", st)
    for i, ((tree, fam, qq, kk, radius, _), a, b) in enumerate(zip(reqs, alone, mixed)):
        for x, y, nm in zip(a[:3], b[:3], ("items", "dists", "counts")):
            if x is not None:
                assert_same_bits(x, y, f"C5 request {i} mixed vs alone {nm}")
        o = refs[tree]
        if fam == ka.FAMILY_NEAREST_ONE:
            ref = o.nearest_one(qq, threads=16)
        elif fam == ka.FAMILY_NEAREST_N:
            ref = o.nearest_n(qq, kk, threads=16)
        else:
            ref = o.best_n_within(qq, radius, kk, threads=16)
        for x, y, nm in zip(a[:3], ref, ("items", "dists", "counts")):
            if x is not None:
                assert_same_bits(x, y, f"C5 request {i} vs oracle {nm}")
