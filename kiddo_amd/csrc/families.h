// families.h — the thread-per-query kernels of every query family, templated on the
// distance metric M (device.h: SqEuclidean, Manhattan, Chebyshev, Minkowski).
//   nearest_one      kiddo query/nearest_one.rs:111-143, :321-352
//   approx           query/approx_nearest_one.rs:49-71 (descent to one leaf)
//   nearest_n        query/nearest_n.rs:21-38, nearest_n_within.rs:127-179, :441-479
//   best_n_within    query/best_n_within.rs:163-196, :290-306
//   within           query/within.rs:18-41, within_unsorted.rs:133-156, :267-286
// The metric is a kernel argument (Minkowski carries its exponent); a plain widened query
// means SquaredEuclidean, so those instantiations are the original kernels.
#pragma once

#include "device.h"

namespace kd {
namespace dev {

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_nearest_one(DevTree<A, K> t, const A* __restrict__ queries,
                                                        const uint32_t* __restrict__ order, uint64_t m, M met,
                                                        uint32_t* __restrict__ out_items,
                                                        typename Ax<A>::O* __restrict__ out_d,
                                                        unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    NearestOneVis<O> v{oinf<O>(), 0u};
    traverse(t, q, mq, v, cnt);
    out_items[qi] = v.item;
    out_d[qi] = v.best;
  }
  flush_counters(cnt, stats);
}

// nearest_one, G lanes per query (SquaredEuclidean; GroupNearestOneVis): for batches too
// small to fill the device one thread per query. Lane 0 of the group writes the result.
template <class A, int K, int G>
__global__ __launch_bounds__(kBlock) void k_nearest_one_group(DevTree<A, K> t, const A* __restrict__ queries,
                                                              const uint32_t* __restrict__ order, uint64_t m,
                                                              uint32_t* __restrict__ out_items,
                                                              typename Ax<A>::O* __restrict__ out_d,
                                                              unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  pin_stems(t);
  const uint64_t gid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / G;
  const int gl = int(threadIdx.x) & (G - 1);
  const bool active = gid < m;  // uniform within a group
  Counters c;
  if (active) {
    const uint64_t qi = order ? order[gid] : gid;
    Small<A, K> q;
    Small<O, K> qw;
    load_query<A, K>(queries, qi, q, qw);
    GroupNearestOneVis<O, G> v{oinf<O>(), 0u, gl};
    group_traverse<A, K, G>(t, q, qw, v, c, gl);
    if (gl == 0) {
      out_items[qi] = v.item;
      out_d[qi] = v.best;
    }
    if (gl != 0) c = Counters{};  // counted once per group
  }
  flush_counters(c, stats);
}

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_approx_nearest_one(DevTree<A, K> t, const A* __restrict__ queries,
                                                               const uint32_t* __restrict__ order, uint64_t m,
                                                               M met, uint32_t* __restrict__ out_items,
                                                               A* __restrict__ out_d,
                                                               unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE  // O == A for float trees
  if (active) {
    const auto mq = wrap_qw(qw, met);
    NearestOneVis<O> v{oinf<O>(), 0u};
    if (t.n > 0) {
      // straight descent (orchestrator/mod.rs:133-151, traits/stem_strategy.rs:182-200)
      uint32_t idx = 1;
      int dim = 0;
      for (int level = 0; level < t.nlev; ++level) {
        const A pivot = stem_at(t, idx);
        idx = (idx << 1) | (q[dim] >= pivot ? 1u : 0u);
        dim = (dim + 1 == K) ? 0 : dim + 1;
      }
      const uint32_t leaf = idx - (1u << t.nlev);
      cnt.stems += uint32_t(t.nlev);
      cnt.leaves += 1;
      cnt.points += t.leaf_start[leaf + 1] - t.leaf_start[leaf];
      v.leaf(t, mq, leaf, 1u);
    }
    out_items[qi] = v.item;
    out_d[qi] = A(v.best);
  }
  flush_counters(cnt, stats);
}

template <class A, int K, int MAXK, class M>
__global__ __launch_bounds__(kBlock) void k_nearest_n(DevTree<A, K> t, const A* __restrict__ queries,
                                                      const uint32_t* __restrict__ order, uint64_t m, M met,
                                                      int k, typename Ax<A>::O radius, int excl,
                                                      uint32_t* __restrict__ out_items,
                                                      typename Ax<A>::O* __restrict__ out_d,
                                                      uint32_t* __restrict__ out_counts,
                                                      unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    KnnVis<O, MAXK> v;
    v.init(k, radius, excl != 0);
    traverse(t, q, mq, v, cnt);
    uint32_t* ri = out_items + qi * uint64_t(k);
    O* rdd = out_d + qi * uint64_t(k);
#pragma unroll
    for (int j = 0; j < MAXK; ++j)
      if (j < k) {
        ri[j] = j < v.cnt ? v.top.it[j] : 0xFFFFFFFFu;
        rdd[j] = j < v.cnt ? v.top.d[j] : oinf<O>();
      }
    out_counts[qi] = uint32_t(v.cnt);
  }
  flush_counters(cnt, stats);
}

template <class A, int K, int MAXK, class M>
__global__ __launch_bounds__(kBlock) void k_best_n_within(DevTree<A, K> t, const A* __restrict__ queries,
                                                          const uint32_t* __restrict__ order, uint64_t m, M met,
                                                          int n, typename Ax<A>::O radius, int excl,
                                                          uint32_t* __restrict__ out_items,
                                                          typename Ax<A>::O* __restrict__ out_d,
                                                          uint32_t* __restrict__ out_counts,
                                                          unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    BestNVis<O, MAXK> v;
    v.init(n, radius, excl != 0);
    traverse(t, q, mq, v, cnt);
    uint32_t* ri = out_items + qi * uint64_t(n);
    O* rdd = out_d + qi * uint64_t(n);
#pragma unroll
    for (int j = 0; j < MAXK; ++j)
      if (j < n) {
        ri[j] = j < v.cnt ? v.it[j] : 0xFFFFFFFFu;
        rdd[j] = j < v.cnt ? v.d[j] : oinf<O>();
      }
    out_counts[qi] = uint32_t(v.cnt);
  }
  flush_counters(cnt, stats);
}

// best_n_within, G lanes per query (n <= G; SquaredEuclidean): more lanes for batches too
// small to fill the device one thread per query, and each leaf's entries are loaded G at a
// time. Lane j < n of the group writes slot j of the row.
template <class A, int K, int G>
__global__ __launch_bounds__(kBlock) void k_best_n_within_group(DevTree<A, K> t, const A* __restrict__ queries,
                                                                const uint32_t* __restrict__ order, uint64_t m,
                                                                int n, typename Ax<A>::O radius, int excl,
                                                                uint32_t* __restrict__ out_items,
                                                                typename Ax<A>::O* __restrict__ out_d,
                                                                uint32_t* __restrict__ out_counts,
                                                                unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  pin_stems(t);
  const uint64_t gid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / G;
  const int gl = int(threadIdx.x) & (G - 1);
  const bool active = gid < m;  // uniform within a group
  Counters c;
  if (active) {
    const uint64_t qi = order ? order[gid] : gid;
    Small<A, K> q;
    Small<O, K> qw;
    load_query<A, K>(queries, qi, q, qw);
    GroupBestNVis<O, G> v;
    v.init(n, radius, excl != 0, gl);
    traverse(t, q, qw, v, c);
    if (gl < n) {
      out_items[qi * uint64_t(n) + gl] = gl < v.cnt ? v.si : 0xFFFFFFFFu;
      out_d[qi * uint64_t(n) + gl] = gl < v.cnt ? v.sd : oinf<O>();
    }
    if (gl == 0) out_counts[qi] = uint32_t(v.cnt);
    if (gl != 0) c = Counters{};  // counted once per group
  }
  flush_counters(c, stats);
}

// nearest_n / nearest_n().within for k above the register forms (k > 64): the top-k is a
// max-heap in the query's output row (device.h RowHeap), sorted in place at the end.
// SCREEN: f64 trees with the FP32 shadow (SquaredEuclidean only).
template <class A, int K, class M, bool SCREEN>
__global__ __launch_bounds__(kBlock) void k_nearest_n_heap(DevTree<A, K> t, const A* __restrict__ queries,
                                                           const uint32_t* __restrict__ order, uint64_t m, M met,
                                                           int k, typename Ax<A>::O radius, int excl,
                                                           const double* __restrict__ max_abs,
                                                           uint32_t* __restrict__ out_items,
                                                           typename Ax<A>::O* __restrict__ out_d,
                                                           uint32_t* __restrict__ out_counts,
                                                           unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    HeapKnnVis<O, SCREEN> v;
    v.h.init(out_items + qi * uint64_t(k), out_d + qi * uint64_t(k), uint32_t(k));
    v.radius = radius;
    v.excl = excl != 0;
    v.rescored = 0;
    v.thr = __builtin_huge_valf();
    if constexpr (SCREEN) {
      static_assert(std::is_same<A, double>::value, "the FP32 screen is for f64 trees");
      screen_setup<K>(q, max_abs, v.qf, v.eps_k);
      v.thr = screen_threshold(double(v.bound()), v.eps_k);
    }
    traverse(t, q, mq, v, cnt);
    v.h.finish();
    out_counts[qi] = v.h.cnt;
    cnt.rescored = v.rescored;
  }
  flush_counters(cnt, stats);
}

// best_n_within for n above the register form (n > 64): the n smallest items in a max-heap
// by item in the query's output row, sorted by item at the end.
template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_best_n_within_heap(DevTree<A, K> t, const A* __restrict__ queries,
                                                               const uint32_t* __restrict__ order, uint64_t m,
                                                               M met, int n, typename Ax<A>::O radius, int excl,
                                                               uint32_t* __restrict__ out_items,
                                                               typename Ax<A>::O* __restrict__ out_d,
                                                               uint32_t* __restrict__ out_counts,
                                                               unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    HeapBestNVis<O> v;
    v.h.init(out_items + qi * uint64_t(n), out_d + qi * uint64_t(n), uint32_t(n));
    v.radius = radius;
    v.excl = excl != 0;
    traverse(t, q, mq, v, cnt);
    v.h.finish();
    out_counts[qi] = v.h.cnt;
  }
  flush_counters(cnt, stats);
}

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_within_count(DevTree<A, K> t, const A* __restrict__ queries,
                                                         const uint32_t* __restrict__ order, uint64_t m, M met,
                                                         typename Ax<A>::O radius, int excl,
                                                         uint64_t* __restrict__ counts,
                                                         unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    WithinVis<O> v{radius, excl != 0, 0, nullptr, nullptr, 0};
    traverse(t, q, mq, v, cnt);
    counts[qi] = v.count;
  }
  flush_counters(cnt, stats);
}

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_within_fill(DevTree<A, K> t, const A* __restrict__ queries,
                                                        const uint32_t* __restrict__ order, uint64_t m, M met,
                                                        typename Ax<A>::O radius, int excl,
                                                        const uint64_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out_items,
                                                        typename Ax<A>::O* __restrict__ out_d,
                                                        unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const auto mq = wrap_qw(qw, met);
    const uint64_t o = offsets[qi], cap = offsets[qi + 1] - o;
    WithinVis<O> v{radius, excl != 0, 0, out_items + o, out_d + o, cap};
    traverse(t, q, mq, v, cnt);
  }
  flush_counters(cnt, stats);
}

// Sorts each query's segment of a within result by (distance, item), one thread per query
// (sort_segment_any: short segments in registers). A pass of its own after the fill, so the
// fill kernels keep their register budget (with a 16-slot network inlined the f64 fill took
// 130 VGPRs, 3 waves per SIMD, instead of 76 and 6).
template <class O>
__global__ __launch_bounds__(kBlock) void k_sort_segments(const uint64_t* __restrict__ offsets, uint64_t m,
                                                          uint32_t* __restrict__ items, O* __restrict__ d) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= m) return;
  const uint64_t o = offsets[i];
  sort_segment_any(items + o, d + o, offsets[i + 1] - o);
}

// within count, G lanes per query (SquaredEuclidean): the group form of k_within_count, as
// k_within_fill_group below without writes; lane 0 of each group stores the count.
template <class A, int K, int G>
__global__ __launch_bounds__(kBlock) void k_within_count_group(DevTree<A, K> t, const A* __restrict__ queries,
                                                               const uint32_t* __restrict__ order, uint64_t m,
                                                               typename Ax<A>::O radius, int excl,
                                                               uint64_t* __restrict__ counts,
                                                               unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  pin_stems(t);
  const uint64_t gid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / G;
  const int gl = int(threadIdx.x) & (G - 1);
  const bool active = gid < m;  // uniform within a group
  Counters c;
  if (active) {
    const uint64_t qi = order ? order[gid] : gid;
    Small<A, K> q;
    Small<O, K> qw;
    load_query<A, K>(queries, qi, q, qw);
    GroupWithinVis<O, G> v{radius, excl != 0, gl, 0, nullptr, nullptr, 0};
    traverse(t, q, qw, v, c);
    if (gl == 0) counts[qi] = v.count;
    if (gl != 0) c = Counters{};  // counted once per group
  }
  flush_counters(c, stats);
}

// within fill, G lanes per query (dense outputs; SquaredEuclidean). Every lane of a group
// runs the thread traversal of its query: with the same query and the radius as the only
// bound, the lanes take the same path in lockstep, and the leaf visit is shared (the lanes
// split its entries and write the hits consecutively). The count comes from
// kd_within_count as for the thread form.
template <class A, int K, int G>
__global__ __launch_bounds__(kBlock) void k_within_fill_group(DevTree<A, K> t, const A* __restrict__ queries,
                                                              const uint32_t* __restrict__ order, uint64_t m,
                                                              typename Ax<A>::O radius, int excl,
                                                              const uint64_t* __restrict__ offsets,
                                                              uint32_t* __restrict__ out_items,
                                                              typename Ax<A>::O* __restrict__ out_d) {
  using O = typename Ax<A>::O;
  using V = GroupWithinStageVis<O, G>;
  __shared__ uint32_t ring_it[(kBlock / G) * V::kRing];
  __shared__ O ring_d[(kBlock / G) * V::kRing];
  pin_stems(t);
  const uint64_t gid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / G;
  const int gl = int(threadIdx.x) & (G - 1);
  if (gid >= m) return;  // uniform within a group; pin_stems synchronised the block already
  const uint64_t qi = order ? order[gid] : gid;
  Small<A, K> q;
  Small<O, K> qw;
  load_query<A, K>(queries, qi, q, qw);
  const uint64_t o = offsets[qi], cap = offsets[qi + 1] - o;
  const uint32_t ring = (threadIdx.x / G) * V::kRing;
  V v{radius, excl != 0, gl, 0, 0, o, out_items + o, out_d + o, cap, ring_it + ring, ring_d + ring};
  Counters c;
  traverse(t, q, qw, v, c);
  v.flush(true);
}

}  // namespace dev
}  // namespace kd
