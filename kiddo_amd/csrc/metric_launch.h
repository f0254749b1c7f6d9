// metric_launch.h — launchers of the thread-per-query kernels for one distance metric M.
// The SquaredEuclidean instances are made where the families' default launchers live
// (k_families.hip, k_periodic.hip); every other metric is instantiated in a translation
// unit of its own (k_metric_*.hip) so the builds run in parallel.
//
// The metric travels in the query flags (include/kiddo_amd.h KD_METRIC_*): bits 8-11 the
// kind (0 SquaredEuclidean, 1 Manhattan, 2 Chebyshev, 3 Minkowski), bits 12-19 the
// Minkowski exponent.
#pragma once

#include "families.h"
#include "periodic.h"
#include "tree.h"

namespace kd {

inline int metric_kind(uint32_t flags) { return int((flags >> 8) & 0xFu); }
inline int metric_p(uint32_t flags) { return int((flags >> 12) & 0xFFu); }

template <class A, int K> inline size_t lds_bytes_for(const dev::DevTree<A, K>& dt) {
  return size_t(dt.lds_n) * sizeof(A);
}

template <class A> struct PBox {
  A v[4];
};
template <class A, int K> inline PBox<A> pbox_of(const void* box) {
  PBox<A> b{};
  const A* v = static_cast<const A*>(box);
  for (int d = 0; d < K; ++d) b.v[d] = v[d];
  return b;
}

namespace dev {

#define KD_PERIODIC_PROLOGUE                                          \
  KD_QUERY_PROLOGUE                                                   \
  PeriodicQW<O, K, M> pq;                                             \
  pq.q = qw;                                                          \
  pq.met = met;                                                       \
  pq.box.fill(O(0));                                                  \
  _Pragma("unroll") for (int d = 0; d < K; ++d) pq.box.set(d, O(box.v[d]));

// periodic nearest_one (query/periodic.rs:196-244). Images are the 3^K shifts (each axis
// -L, 0 or +L; index = sum of (shift+1) * 3^(K-1-axis), the reference's enumeration order)
// other than the all-zero one. Their lower bound combines dist1(o_a) from 0 in axis order,
// o_a being how far the shifted coordinate lies outside [0, L_a] (:27-37, :84-108). They
// are visited in ascending (lower bound, index) while the bound does not exceed the best
// distance so far (the canonical, inclusive form of :211-233), each by a fresh Cartesian
// nearest_one of the shifted query; a result replaces the best when it is smaller in
// (distance, item) order.
template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_periodic_nearest_one(DevTree<A, K> t, const A* __restrict__ queries,
                                                                 const uint32_t* __restrict__ order, uint64_t m,
                                                                 M met, PBox<A> box,
                                                                 uint32_t* __restrict__ out_items,
                                                                 A* __restrict__ out_d,
                                                                 unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    // one traversal call site for the home query and every image (a second inlined DFS and
    // the image tables held across it took 127 VGPRs); the tables are recomputed from the
    // query and the box (kernel arguments) at each choice of the next image
    auto shifted = [&](int s, int a) -> A {
      const A L = box.v[a], x = q[a];
      return s == 0 ? A(x - L) : s == 1 ? x : A(x + L);
    };
    auto off1 = [&](int s, int a) -> O {
      const A L = box.v[a], w = shifted(s, a);
      const A o = w < A(0) ? A(fabs(w - A(0))) : (w > L ? A(fabs(w - L)) : A(0));
      return met.d1(O(o));
    };
    constexpr int kImages = K == 2 ? 9 : K == 3 ? 27 : 81;
    constexpr int kHome = kImages / 2;  // all shifts zero
    O best = oinf<O>();
    uint32_t item = 0u;
    O prev_lb = O(0);
    int prev = -1;
    Small<A, K> wq = q;
    Small<O, K> wqw = qw;
    for (;;) {
      NearestOneVis<O> img{oinf<O>(), 0u};
      traverse(t, wq, wrap_qw(wqw, met), img, cnt);
      if (img.best < best || (img.best == best && img.item < item)) {
        best = img.best;
        item = img.item;
      }
      // the next image after (prev_lb, prev) in (lower bound, index) order
      O nlb = oinf<O>();
      int next = -1;
#pragma unroll 1
      for (int i = 0; i < kImages; ++i) {
        if (i == kHome) continue;
        int r = i;
        int dig[K];
#pragma unroll
        for (int a = K - 1; a >= 0; --a) {
          dig[a] = r % 3;
          r /= 3;
        }
        O lb = O(0);
#pragma unroll
        for (int a = 0; a < K; ++a) lb = met.combine(lb, off1(dig[a], a));
        const bool after = prev < 0 || lb > prev_lb || (lb == prev_lb && i > prev);
        const bool before = lb < nlb || (lb == nlb && (next < 0 || i < next));
        if (after && before && !(lb > best)) {
          nlb = lb;
          next = i;
        }
      }
      if (next < 0) break;
      int r = next;
#pragma unroll
      for (int a = K - 1; a >= 0; --a) {
        const A w = shifted(r % 3, a);
        r /= 3;
        wq.set(a, w);
        wqw.set(a, O(w));
      }
      prev_lb = nlb;
      prev = next;
    }
    out_items[qi] = item;
    out_d[qi] = A(best);
  }
  flush_counters(cnt, stats);
}

template <class A, int K, int MAXK, class M>
__global__ __launch_bounds__(kBlock) void k_periodic_nearest_n(DevTree<A, K> t, const A* __restrict__ queries,
                                                               const uint32_t* __restrict__ order, uint64_t m,
                                                               M met, PBox<A> box, int k,
                                                               typename Ax<A>::O radius, int excl,
                                                               uint32_t* __restrict__ out_items,
                                                               typename Ax<A>::O* __restrict__ out_d,
                                                               uint32_t* __restrict__ out_counts,
                                                               unsigned long long* __restrict__ stats) {
  KD_PERIODIC_PROLOGUE
  if (active) {
    KnnVis<O, MAXK> v;
    v.init(k, radius, excl != 0);
    periodic_traverse(t, q, pq, v, cnt);
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

// periodic nearest_n for k above the register forms: the top-k in a max-heap in the
// query's output row (device.h RowHeap), sorted in place at the end
template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_periodic_nearest_n_heap(DevTree<A, K> t, const A* __restrict__ queries,
                                                                    const uint32_t* __restrict__ order, uint64_t m,
                                                                    M met, PBox<A> box, int k,
                                                                    typename Ax<A>::O radius, int excl,
                                                                    uint32_t* __restrict__ out_items,
                                                                    typename Ax<A>::O* __restrict__ out_d,
                                                                    uint32_t* __restrict__ out_counts,
                                                                    unsigned long long* __restrict__ stats) {
  KD_PERIODIC_PROLOGUE
  if (active) {
    HeapKnnVis<O, false> v;
    v.h.init(out_items + qi * uint64_t(k), out_d + qi * uint64_t(k), uint32_t(k));
    v.radius = radius;
    v.excl = excl != 0;
    v.rescored = 0;
    v.thr = __builtin_huge_valf();
    periodic_traverse(t, q, pq, v, cnt);
    v.h.finish();
    out_counts[qi] = v.h.cnt;
  }
  flush_counters(cnt, stats);
}

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_periodic_within_count(DevTree<A, K> t, const A* __restrict__ queries,
                                                                  const uint32_t* __restrict__ order, uint64_t m,
                                                                  M met, PBox<A> box, typename Ax<A>::O radius,
                                                                  int excl, uint64_t* __restrict__ counts,
                                                                  unsigned long long* __restrict__ stats) {
  KD_PERIODIC_PROLOGUE
  if (active) {
    WithinVis<O> v{radius, excl != 0, 0, nullptr, nullptr, 0};
    periodic_traverse(t, q, pq, v, cnt);
    counts[qi] = v.count;
  }
  flush_counters(cnt, stats);
}

template <class A, int K, class M>
__global__ __launch_bounds__(kBlock) void k_periodic_within_fill(DevTree<A, K> t, const A* __restrict__ queries,
                                                                 const uint32_t* __restrict__ order, uint64_t m,
                                                                 M met, PBox<A> box, typename Ax<A>::O radius,
                                                                 int excl,
                                                                 const uint64_t* __restrict__ offsets,
                                                                 uint32_t* __restrict__ out_items,
                                                                 typename Ax<A>::O* __restrict__ out_d,
                                                                 unsigned long long* __restrict__ stats) {
  KD_PERIODIC_PROLOGUE
  if (active) {
    const uint64_t o = offsets[qi], cap = offsets[qi + 1] - o;
    WithinVis<O> v{radius, excl != 0, 0, out_items + o, out_d + o, cap};
    periodic_traverse(t, q, pq, v, cnt);
  }
  flush_counters(cnt, stats);
}

}  // namespace dev

// ------------------------------------------------------------------ launchers --------

// Batches too small to keep the device busy one thread per query: the kernel time is then
// one query's chain of dependent loads, and nearest_one and nearest_n (k <= 16) take their
// group forms when the tree's k-NN form is automatic (knn_variant 0). Measured
// (nearest_n(5), 12M-point f64 tree: 8 lanes 0.12 / 0.38 / 0.69 ms against 0.28 / 0.67 /
// 1.18 ms one thread per query at 1k / 64k / 256k queries, 2.17 against 1.77 ms at 1M;
// nearest_n(8) on a 100k-point tree, which stays in the L2, at 100k queries: 0.22 against
// 0.18 ms): below 2^16 queries, or 2^19 on trees of at least 2^22 entries.
inline bool small_batch(const kd_tree* t, uint64_t m) {
  return m < (uint64_t(1) << 16) || (m < (uint64_t(1) << 19) && t->size >= (uint64_t(1) << 22));
}

template <class M> int mlaunch_nearest_one(const QueryArgs& a, M met, uint32_t* items, void* dists) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    if constexpr (std::is_same<M, dev::SqEuclidean>::value) {
      if (a.t->knn_variant == 0 && small_batch(a.t, a.m)) {
        constexpr int G = 8;
        hipLaunchKernelGGL((dev::k_nearest_one_group<A, KK, G>), dim3(grid_for(a.m * G)), dim3(dev::kBlock),
                           lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, items,
                           static_cast<O*>(dists), a.t->counters);
        KD_HIP(hipGetLastError());
        return KD_OK;
      }
    }
    hipLaunchKernelGGL((dev::k_nearest_one<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes_for(dt),
                       a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, items, static_cast<O*>(dists),
                       a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

template <class M> int mlaunch_approx_nearest_one(const QueryArgs& a, M met, uint32_t* items, void* dists) {
  KD_DISPATCH_FK(a.t->dtype, a.t->dims, {
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_approx_nearest_one<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                       lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, items,
                       static_cast<A*>(dists), a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

template <class A, int K, int MAXK, class M>
int mlaunch_knn_typed(const QueryArgs& a, M met, int k, double radius, uint32_t* items, void* dists,
                      uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  auto dt = dev_tree<A, K>(a.t);
  reset_counters(a.t, a.s);
  KernelTimer kt(a.t, a.s);
  hipLaunchKernelGGL((dev::k_nearest_n<A, K, MAXK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes_for(dt),
                     a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, k, O(radius),
                     (a.flags & KD_EXCLUSIVE) ? 1 : 0, items, static_cast<O*>(dists), counts, a.t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

// nearest_n for k above the register forms (any metric; SCREEN: SquaredEuclidean f64 trees
// with the FP32 shadow)
template <class A, int K, class M, bool SCREEN>
int mlaunch_knn_heap(const QueryArgs& a, M met, int k, double radius, uint32_t* items, void* dists,
                     uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  auto dt = dev_tree<A, K>(a.t);
  reset_counters(a.t, a.s);
  KernelTimer kt(a.t, a.s);
  hipLaunchKernelGGL((dev::k_nearest_n_heap<A, K, M, SCREEN>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                     lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, k, O(radius),
                     (a.flags & KD_EXCLUSIVE) ? 1 : 0, a.t->d_max_abs, items, static_cast<O*>(dists), counts,
                     a.t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

// nearest_n with a metric other than SquaredEuclidean: the thread-per-query kernel
template <class M>
int mlaunch_nearest_n(const QueryArgs& a, M met, int k, double radius, uint32_t* items, void* dists,
                      uint32_t* counts) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    if (k > kMaxK) return mlaunch_knn_heap<A, KK, M, false>(a, met, k, radius, items, dists, counts);
    if (k <= 1) return mlaunch_knn_typed<A, KK, 1>(a, met, k, radius, items, dists, counts);
    if (k <= 8) return mlaunch_knn_typed<A, KK, 8>(a, met, k, radius, items, dists, counts);
    if (k <= 16) return mlaunch_knn_typed<A, KK, 16>(a, met, k, radius, items, dists, counts);
    return mlaunch_knn_typed<A, KK, 64>(a, met, k, radius, items, dists, counts);
  });
  return KD_OK;
}

// Lanes per query of best_n_within (SquaredEuclidean): the tree's setting
// (kd_tree_tune_best_n_within; a group smaller than n means one thread), else
// automatically 16 lanes below 2^17 queries, 8 (n <= 8) or 16 below 2^22, one thread above.
// The family scans every entry in range, which the lanes split. Measured (best_n_within(5),
// r = 0.001, clustered 12M-point f64 tree, ~480 entries in range per query): 16 / 8 lanes /
// one thread 0.34 / 0.45 / 1.56 ms at 1k queries, 0.64 / 0.76 / 2.03 at 64k, 1.39 / 1.17 /
// 4.00 at 256k, 4.45 / 3.16 / 6.69 at 1M, 15.0 / 10.6 / 13.2 at 4M.
inline int best_lanes(const kd_tree* t, uint64_t m, int n) {
  if (t->best_group == 1) return 1;
  if (t->best_group > 1) return t->best_group >= n ? t->best_group : 1;
  if (n > 16 || m >= (uint64_t(1) << 22)) return 1;
  return (n > 8 || m < (uint64_t(1) << 17)) ? 16 : 8;
}

template <class A, int K, int MAXK, class M>
int mlaunch_best_typed(const QueryArgs& a, M met, int n, double radius, uint32_t* items, void* dists,
                       uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  auto dt = dev_tree<A, K>(a.t);
  reset_counters(a.t, a.s);
  KernelTimer kt(a.t, a.s);
  if constexpr (std::is_same<M, dev::SqEuclidean>::value) {
    auto group = [&](auto gc) -> int {
      constexpr int G = decltype(gc)::value;
      hipLaunchKernelGGL((dev::k_best_n_within_group<A, K, G>), dim3(grid_for(a.m * G)), dim3(dev::kBlock),
                         lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, n, O(radius),
                         (a.flags & KD_EXCLUSIVE) ? 1 : 0, items, static_cast<O*>(dists), counts, a.t->counters);
      KD_HIP(hipGetLastError());
      return KD_OK;
    };
    switch (best_lanes(a.t, a.m, n)) {
      case 4: return group(std::integral_constant<int, 4>{});
      case 8: return group(std::integral_constant<int, 8>{});
      case 16: return group(std::integral_constant<int, 16>{});
      case 32: return group(std::integral_constant<int, 32>{});
      default: break;
    }
  }
  hipLaunchKernelGGL((dev::k_best_n_within<A, K, MAXK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                     lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, n, O(radius),
                     (a.flags & KD_EXCLUSIVE) ? 1 : 0, items, static_cast<O*>(dists), counts, a.t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class M>
int mlaunch_best_n_within(const QueryArgs& a, M met, int n, double radius, uint32_t* items, void* dists,
                          uint32_t* counts) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    if (n > kMaxK) {  // the n smallest items in a heap in the output row
      using O = typename dev::Ax<A>::O;
      auto dt = dev_tree<A, KK>(a.t);
      reset_counters(a.t, a.s);
      KernelTimer kt(a.t, a.s);
      hipLaunchKernelGGL((dev::k_best_n_within_heap<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                         lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, n,
                         O(radius), (a.flags & KD_EXCLUSIVE) ? 1 : 0, items, static_cast<O*>(dists), counts,
                         a.t->counters);
      KD_HIP(hipGetLastError());
      return KD_OK;
    }
    if (n <= 4) return mlaunch_best_typed<A, KK, 4>(a, met, n, radius, items, dists, counts);
    if (n <= 8) return mlaunch_best_typed<A, KK, 8>(a, met, n, radius, items, dists, counts);
    if (n <= 16) return mlaunch_best_typed<A, KK, 16>(a, met, n, radius, items, dists, counts);
    return mlaunch_best_typed<A, KK, 64>(a, met, n, radius, items, dists, counts);
  });
  return KD_OK;
}

template <class M> int mlaunch_within_count(const QueryArgs& a, M met, double radius, uint64_t* counts) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    if constexpr (std::is_same<M, dev::SqEuclidean>::value) {
      // G lanes per query as tuned (kd_tree_tune_within); one thread per query by default
      auto group = [&](auto gc) -> int {
        constexpr int G = decltype(gc)::value;
        hipLaunchKernelGGL((dev::k_within_count_group<A, KK, G>), dim3(grid_for(a.m * G)), dim3(dev::kBlock),
                           lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, O(radius),
                           (a.flags & KD_EXCLUSIVE) ? 1 : 0, counts, a.t->counters);
        KD_HIP(hipGetLastError());
        return KD_OK;
      };
      switch (a.t->count_group) {
        case 4: return group(std::integral_constant<int, 4>{});
        case 8: return group(std::integral_constant<int, 8>{});
        case 16: return group(std::integral_constant<int, 16>{});
        case 32: return group(std::integral_constant<int, 32>{});
        default: break;  // one thread per query
      }
    }
    hipLaunchKernelGGL((dev::k_within_count<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes_for(dt),
                       a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, O(radius),
                       (a.flags & KD_EXCLUSIVE) ? 1 : 0, counts, a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

template <class M>
int mlaunch_within_fill(const QueryArgs& a, M met, double radius, const uint64_t* offsets, uint32_t* items,
                        void* dists) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    KernelTimer kt(a.t, a.s);
    bool launched = false;
    if constexpr (std::is_same<M, dev::SqEuclidean>::value) {
      // G lanes per query with coalesced result writes: by default for dense outputs (16
      // lanes), or as tuned (kd_tree_tune_within)
      const int g = a.t->fill_group ? a.t->fill_group : (a.dense ? 16 : 1);
      auto group = [&](auto gc) {
        constexpr int G = decltype(gc)::value;
        // the LDS stems are cut so that they and the result ring fit 20 KiB per block, 8
        // blocks per CU: 1.2M queries of the 10M-point f32 mixture, 8 lanes, 2.1e9 results
        // filled in 32.9 ms with 16 KiB of stems (5 blocks per CU), 21.6 ms with 4 KiB and
        // 21.5 ms with none
        constexpr size_t ring = size_t(dev::kBlock / G) * dev::GroupWithinStageVis<O, G>::kRing *
                                (sizeof(uint32_t) + sizeof(O));
        constexpr size_t kBudget = 20480;
        auto dg = dt;
        uint32_t cap = 0;
        if (ring < kBudget)
          while (size_t(2 * (cap ? cap : 1)) * sizeof(A) <= kBudget - ring) cap = cap ? 2 * cap : 2;
        dg.lds_n = dg.lds_n < cap ? dg.lds_n : cap;
        hipLaunchKernelGGL((dev::k_within_fill_group<A, KK, G>), dim3(grid_for(a.m * G)), dim3(dev::kBlock),
                           lds_bytes_for(dg), a.s, dg, static_cast<const A*>(a.queries), a.order, a.m, O(radius),
                           (a.flags & KD_EXCLUSIVE) ? 1 : 0, offsets, items, static_cast<O*>(dists));
        launched = true;
      };
      switch (g) {
        case 4: group(std::integral_constant<int, 4>{}); break;
        case 8: group(std::integral_constant<int, 8>{}); break;
        case 16: group(std::integral_constant<int, 16>{}); break;
        case 32: group(std::integral_constant<int, 32>{}); break;
        default: break;  // 1: one thread per query
      }
    }
    if (!launched)
      hipLaunchKernelGGL((dev::k_within_fill<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes_for(dt),
                         a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met, O(radius),
                         (a.flags & KD_EXCLUSIVE) ? 1 : 0, offsets, items, static_cast<O*>(dists), nullptr);
    KD_HIP(hipGetLastError());
    if (!(a.flags & KD_UNSORTED) && a.m) {
      hipLaunchKernelGGL((dev::k_sort_segments<O>), dim3(grid_for(a.m)), dim3(dev::kBlock), 0, a.s, offsets, a.m,
                         items, static_cast<O*>(dists));
      KD_HIP(hipGetLastError());
    }
  });
  return KD_OK;
}

template <class M>
int mlaunch_periodic_nearest_one(const QueryArgs& a, M met, const void* box, uint32_t* items, void* dists) {
  KD_DISPATCH_FK(a.t->dtype, a.t->dims, {
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_periodic_nearest_one<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                       lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met,
                       pbox_of<A, KK>(box), items, static_cast<A*>(dists), a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

template <class A, int K, int MAXK, class M>
int mlaunch_pknn_typed(const QueryArgs& a, M met, const void* box, int k, double radius, uint32_t* items,
                       void* dists, uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  auto dt = dev_tree<A, K>(a.t);
  reset_counters(a.t, a.s);
  KernelTimer kt(a.t, a.s);
  hipLaunchKernelGGL((dev::k_periodic_nearest_n<A, K, MAXK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                     lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met,
                     pbox_of<A, K>(box), k, O(radius), (a.flags & KD_EXCLUSIVE) ? 1 : 0, items,
                     static_cast<O*>(dists), counts, a.t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class M>
int mlaunch_periodic_nearest_n(const QueryArgs& a, M met, const void* box, int k, double radius, uint32_t* items,
                               void* dists, uint32_t* counts) {
  KD_DISPATCH_FK(a.t->dtype, a.t->dims, {
    if (k > kMaxK) {  // the top-k in a heap in the output row
      using O = typename dev::Ax<A>::O;
      auto dt = dev_tree<A, KK>(a.t);
      reset_counters(a.t, a.s);
      KernelTimer kt(a.t, a.s);
      hipLaunchKernelGGL((dev::k_periodic_nearest_n_heap<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                         lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met,
                         pbox_of<A, KK>(box), k, O(radius), (a.flags & KD_EXCLUSIVE) ? 1 : 0, items,
                         static_cast<O*>(dists), counts, a.t->counters);
      KD_HIP(hipGetLastError());
      return KD_OK;
    }
    if (k <= 1) return mlaunch_pknn_typed<A, KK, 1>(a, met, box, k, radius, items, dists, counts);
    if (k <= 4) return mlaunch_pknn_typed<A, KK, 4>(a, met, box, k, radius, items, dists, counts);
    if (k <= 8) return mlaunch_pknn_typed<A, KK, 8>(a, met, box, k, radius, items, dists, counts);
    if (k <= 16) return mlaunch_pknn_typed<A, KK, 16>(a, met, box, k, radius, items, dists, counts);
    if (k <= 32) return mlaunch_pknn_typed<A, KK, 32>(a, met, box, k, radius, items, dists, counts);
    return mlaunch_pknn_typed<A, KK, 64>(a, met, box, k, radius, items, dists, counts);
  });
  return KD_OK;
}

template <class M>
int mlaunch_periodic_within_count(const QueryArgs& a, M met, const void* box, double radius, uint64_t* counts) {
  KD_DISPATCH_FK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_periodic_within_count<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                       lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met,
                       pbox_of<A, KK>(box), O(radius), (a.flags & KD_EXCLUSIVE) ? 1 : 0, counts, a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

template <class M>
int mlaunch_periodic_within_fill(const QueryArgs& a, M met, const void* box, double radius,
                                 const uint64_t* offsets, uint32_t* items, void* dists) {
  KD_DISPATCH_FK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_periodic_within_fill<A, KK, M>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                       lds_bytes_for(dt), a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, met,
                       pbox_of<A, KK>(box), O(radius), (a.flags & KD_EXCLUSIVE) ? 1 : 0, offsets, items,
                       static_cast<O*>(dists), nullptr);
    KD_HIP(hipGetLastError());
    if (!(a.flags & KD_UNSORTED) && a.m) {
      hipLaunchKernelGGL((dev::k_sort_segments<O>), dim3(grid_for(a.m)), dim3(dev::kBlock), 0, a.s, offsets, a.m,
                         items, static_cast<O*>(dists));
      KD_HIP(hipGetLastError());
    }
  });
  return KD_OK;
}

// Explicit instantiations per metric, one translation unit each (k_metric_*.hip).
#define KD_METRIC_LAUNCHERS(EXT, M)                                                                          \
  EXT template int mlaunch_nearest_one<M>(const QueryArgs&, M, uint32_t*, void*);                           \
  EXT template int mlaunch_approx_nearest_one<M>(const QueryArgs&, M, uint32_t*, void*);                    \
  EXT template int mlaunch_nearest_n<M>(const QueryArgs&, M, int, double, uint32_t*, void*, uint32_t*);      \
  EXT template int mlaunch_best_n_within<M>(const QueryArgs&, M, int, double, uint32_t*, void*, uint32_t*);  \
  EXT template int mlaunch_within_count<M>(const QueryArgs&, M, double, uint64_t*);                         \
  EXT template int mlaunch_within_fill<M>(const QueryArgs&, M, double, const uint64_t*, uint32_t*, void*);  \
  EXT template int mlaunch_periodic_nearest_one<M>(const QueryArgs&, M, const void*, uint32_t*, void*);     \
  EXT template int mlaunch_periodic_nearest_n<M>(const QueryArgs&, M, const void*, int, double, uint32_t*,  \
                                                 void*, uint32_t*);                                          \
  EXT template int mlaunch_periodic_within_count<M>(const QueryArgs&, M, const void*, double, uint64_t*);   \
  EXT template int mlaunch_periodic_within_fill<M>(const QueryArgs&, M, const void*, double, const uint64_t*, \
                                                   uint32_t*, void*);

KD_METRIC_LAUNCHERS(extern, dev::Manhattan)
KD_METRIC_LAUNCHERS(extern, dev::Chebyshev)
KD_METRIC_LAUNCHERS(extern, dev::Minkowski)

// Calls CALL with `met` bound to the metric the flags name.
#define KD_WITH_METRIC(flags_, CALL)                                                   \
  switch (::kd::metric_kind(flags_)) {                                                 \
    case 0: { const dev::SqEuclidean met{}; return CALL; }                             \
    case 1: { const dev::Manhattan met{}; return CALL; }                               \
    case 2: { const dev::Chebyshev met{}; return CALL; }                               \
    case 3: { const dev::Minkowski met{::kd::metric_p(flags_)}; return CALL; }         \
    default: return ::kd::set_error(KD_ERR_INVALID, "unknown metric %d", ::kd::metric_kind(flags_)); \
  }

}  // namespace kd
