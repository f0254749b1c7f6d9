"""Serving loop on the GPU: requests of mixed families, trees and sizes submitted
concurrently are coalesced and overlapped, and every request's results equal the results of
the same request run alone, and the oracle's, bit for bit."""
import threading

import numpy as np
import pytest

from engines import assert_same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def serving_case(ka, oracle_mod, cuda):
    from kiddo_amd.datasets import latlon_clusters, palette_u16, weight_rank_items
    print("case: serving build", flush=True)
    pts = latlon_clusters(200_000, seed=3)
    items = weight_rank_items(len(pts), seed=4)
    pal = palette_u16(65536, seed=5)
    t0 = ka.ImmutableKdTree.new_from_entries(pts, items, 32)
    t1 = ka.ImmutableKdTree.new_from_slice(pal, 32)
    o0 = oracle_mod.Tree.build(pts, 32, items)
    o1 = oracle_mod.Tree.build(pal, 32)
    rng = np.random.default_rng(11)
    reqs = []
    for i in range(90):
        tree = int(rng.integers(0, 2))
        fam = [ka.FAMILY_NEAREST_ONE, ka.FAMILY_NEAREST_N, ka.FAMILY_BEST_N_WITHIN][i % 3]
        m = int(np.exp(rng.uniform(0, np.log(5000))))
        if tree == 0:
            q = latlon_clusters(m, seed=1000 + i)
            radius = 1e-4
        else:
            q = palette_u16(m, seed=2000 + i)
            radius = 4.0e6
        if fam == ka.FAMILY_NEAREST_N and i % 2 == 0:
            radius = np.inf  # plain nearest_n; odd ones are nearest_n().within(radius)
        reqs.append((tree, fam, q, 5, radius))
    return [t0, t1], [o0, o1], reqs


def _run_alone(ka, srv, reqs):
    out = []
    for tree, fam, q, k, r in reqs:
        tk = srv.submit(tree, fam, q, k, r)
        out.append(srv.wait(tk))
    return out


def _run_mixed(ka, srv, reqs):
    tickets = [None] * len(reqs)
    order = np.random.default_rng(7).permutation(len(reqs))
    half = len(order) // 2

    def client(idx):
        for i in idx:
            tree, fam, q, k, r = reqs[i]
            tickets[i] = srv.submit(tree, fam, q, k, r)

    th = [threading.Thread(target=client, args=(order[:half],)),
          threading.Thread(target=client, args=(order[half:],))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return [srv.wait(tk) for tk in tickets]


def _same(a, b, tag):
    for x, y, nm in zip(a[:3], b[:3], ("items", "dists", "counts")):
        if x is None and y is None:
            continue
        assert_same_bits(x, y, f"{tag} {nm}")


def test_serving_mix_equals_alone_and_oracle(ka, serving_case):
    trees, refs, reqs = serving_case
    srv_alone = ka.Server(trees, arena_bytes=64 << 20, slots=1, max_coalesce=1)
    alone = _run_alone(ka, srv_alone, reqs)
    srv_alone.close()
    srv = ka.Server(trees, arena_bytes=64 << 20, slots=4, max_coalesce=1 << 16)
    for rep in range(2):
        print(f"case: serving mix rep {rep}", flush=True)
        mixed = _run_mixed(ka, srv, reqs)
        for i, (a, b) in enumerate(zip(alone, mixed)):
            _same(a, b, f"request {i}")
    st = srv.stats()
    srv.close()
    assert st["done"] == 2 * len(reqs)
    assert st["coalesced"] > 0, st
    # oracle on every request
    for i, ((tree, fam, q, k, r), res) in enumerate(zip(reqs, alone)):
        o = refs[tree]
        if fam == ka.FAMILY_NEAREST_ONE:
            oi, od = o.nearest_one(q)
            assert_same_bits(res[0], oi, f"req {i} nearest_one items")
            assert_same_bits(res[1], od, f"req {i} nearest_one dists")
        elif fam == ka.FAMILY_NEAREST_N:
            oi, od, oc = o.nearest_n(q, k, r)
            _same(res, (oi, od, oc), f"req {i} nearest_n")
        else:
            oi, od, oc = o.best_n_within(q, r, k)
            _same(res, (oi, od, oc), f"req {i} best_n_within")


def test_serving_admission_by_memory(ka, serving_case):
    """A small arena forces admissions to wait for memory; results are unchanged."""
    from kiddo_amd.datasets import latlon_clusters
    trees, refs, reqs = serving_case
    # each request needs ~0.8 MiB of queries + results, so a 1 MiB arena holds one at a time
    big = [(0, ka.FAMILY_NEAREST_N, latlon_clusters(9000, seed=300 + i), 5, np.inf) for i in range(8)]
    srv = ka.Server(trees, arena_bytes=1 << 20, slots=4, max_coalesce=1 << 16)
    mixed = _run_mixed(ka, srv, big)
    st = srv.stats()
    srv.close()
    alone_srv = ka.Server(trees, arena_bytes=64 << 20, slots=1, max_coalesce=1)
    alone = _run_alone(ka, alone_srv, big)
    alone_srv.close()
    for i, (a, b) in enumerate(zip(alone, mixed)):
        _same(a, b, f"request {i}")
    assert st["peak_arena_bytes"] <= 1 << 20
    assert st["memory_waits"] > 0, st


def test_serving_rejects_oversized_request(ka, serving_case):
    trees, refs, reqs = serving_case
    srv = ka.Server(trees, arena_bytes=1 << 20, slots=2)
    from kiddo_amd.datasets import latlon_clusters
    with pytest.raises(ka.KiddoAmdError):
        srv.submit(0, ka.FAMILY_NEAREST_N, latlon_clusters(100_000, seed=1), 16)
    srv.close()


def test_serving_parts_equal_alone_and_oracle(ka, serving_case):
    """Requests cut into parts of 700 queries (and a request of exactly one part, and one of
    a part plus one query) are answered exactly as whole requests alone, and as the oracle;
    the server launches more parts than it receives requests."""
    from kiddo_amd.datasets import latlon_clusters, palette_u16
    trees, refs, _ = serving_case
    reqs = []
    for i, m in enumerate((5000, 700, 701, 1, 2800, 6123)):
        reqs.append((0, ka.FAMILY_NEAREST_N, latlon_clusters(m, seed=400 + i), 5, np.inf))
        reqs.append((1, ka.FAMILY_BEST_N_WITHIN, palette_u16(m, seed=500 + i), 5, 4.0e6))
        reqs.append((0, ka.FAMILY_NEAREST_ONE, latlon_clusters(m, seed=600 + i), 1, np.inf))
    srv1 = ka.Server(trees, arena_bytes=64 << 20, slots=1, max_coalesce=1, max_part=0)
    alone = _run_alone(ka, srv1, reqs)
    srv1.close()
    srv = ka.Server(trees, arena_bytes=64 << 20, slots=4, max_coalesce=4000, max_part=700)
    mixed = _run_mixed(ka, srv, reqs)
    st = srv.stats()
    srv.close()
    parts = sum(-(-r[2].shape[0] // 700) for r in reqs)
    assert st["done"] == len(reqs) and st["launches"] + st["coalesced"] == parts, st
    for i, ((tree, fam, q, k, r), a, b) in enumerate(zip(reqs, alone, mixed)):
        _same(a, b, f"request {i} parts vs whole")
        o = refs[tree]
        if fam == ka.FAMILY_NEAREST_ONE:
            oi, od = o.nearest_one(q)
            assert_same_bits(b[0], oi, f"req {i} nearest_one items")
            assert_same_bits(b[1], od, f"req {i} nearest_one dists")
        elif fam == ka.FAMILY_NEAREST_N:
            _same(b, o.nearest_n(q, k, r), f"req {i} nearest_n")
        else:
            _same(b, o.best_n_within(q, r, k), f"req {i} best_n_within")


def test_serving_slot_workers_stress(ka, serving_case):
    """8 slots (one worker thread each), parts of 250 queries, 240 requests from 4 client
    threads: parts of one request finish on different workers in any order, and every
    request equals the same request on a one-slot server, bit for bit."""
    from kiddo_amd.datasets import latlon_clusters, palette_u16
    trees, _, _ = serving_case
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(240):
        tree = int(rng.integers(0, 2))
        fam = [ka.FAMILY_NEAREST_ONE, ka.FAMILY_NEAREST_N, ka.FAMILY_BEST_N_WITHIN][i % 3]
        m = int(rng.integers(1, 3000))
        q = latlon_clusters(m, seed=3000 + i) if tree == 0 else palette_u16(m, seed=4000 + i)
        r = (1e-4 if tree == 0 else 4.0e6) if fam == ka.FAMILY_BEST_N_WITHIN else np.inf
        reqs.append((tree, fam, q, 5, r))
    srv1 = ka.Server(trees, arena_bytes=64 << 20, slots=1, max_coalesce=1, max_part=0)
    alone = _run_alone(ka, srv1, reqs)
    srv1.close()
    srv = ka.Server(trees, arena_bytes=64 << 20, slots=8, max_coalesce=2000, max_part=250)
    tickets = [None] * len(reqs)

    def client(idx):
        for i in idx:
            tree, fam, q, k, r = reqs[i]
            tickets[i] = srv.submit(tree, fam, q, k, r)

    order = rng.permutation(len(reqs))
    th = [threading.Thread(target=client, args=(order[j::4],)) for j in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    mixed = [srv.wait(tk) for tk in tickets]
    st = srv.stats()
    srv.close()
    assert st["done"] == len(reqs), st
    for i, (a, b) in enumerate(zip(alone, mixed)):
        _same(a, b, f"request {i} on 8 slot workers vs alone")
        assert b[3] > 0.0
