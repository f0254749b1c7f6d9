// k_families.hip — nearest_one, nearest_one().approx(), best_n_within and the two
// passes of within (count, then fill), one thread per query.
//   nearest_one      kiddo query/nearest_one.rs:111-143, :321-352
//   approx           query/approx_nearest_one.rs:49-71 (descent to one leaf)
//   best_n_within    query/best_n_within.rs:163-196, :290-306
//   within           query/within.rs:18-41, within_unsorted.rs:133-156, :267-286

#include "tree.h"

namespace kd {
namespace dev {

template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_nearest_one(DevTree<A, K> t, const A* __restrict__ queries,
                                                        const uint32_t* __restrict__ order, uint64_t m,
                                                        uint32_t* __restrict__ out_items,
                                                        typename Ax<A>::O* __restrict__ out_d,
                                                        unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    NearestOneVis<O> v{oinf<O>(), 0u};
    traverse(t, q, qw, v, cnt);
    out_items[qi] = v.item;
    out_d[qi] = v.best;
  }
  flush_counters(cnt, stats);
}

template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_approx_nearest_one(DevTree<A, K> t, const A* __restrict__ queries,
                                                               const uint32_t* __restrict__ order, uint64_t m,
                                                               uint32_t* __restrict__ out_items,
                                                               A* __restrict__ out_d,
                                                               unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE  // O == A for float trees
  if (active) {
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
      v.leaf(t, qw, leaf);
    }
    out_items[qi] = v.item;
    out_d[qi] = A(v.best);
  }
  flush_counters(cnt, stats);
}

template <class A, int K, int MAXK>
__global__ __launch_bounds__(kBlock) void k_best_n_within(DevTree<A, K> t, const A* __restrict__ queries,
                                                          const uint32_t* __restrict__ order, uint64_t m,
                                                          int n, typename Ax<A>::O radius, int excl,
                                                          uint32_t* __restrict__ out_items,
                                                          typename Ax<A>::O* __restrict__ out_d,
                                                          uint32_t* __restrict__ out_counts,
                                                          unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    BestNVis<O, MAXK> v;
    v.init(n, radius, excl != 0);
    traverse(t, q, qw, v, cnt);
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

template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_within_count(DevTree<A, K> t, const A* __restrict__ queries,
                                                         const uint32_t* __restrict__ order, uint64_t m,
                                                         typename Ax<A>::O radius, int excl,
                                                         uint64_t* __restrict__ counts,
                                                         unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    WithinVis<O> v{radius, excl != 0, 0, nullptr, nullptr, 0};
    traverse(t, q, qw, v, cnt);
    counts[qi] = v.count;
  }
  flush_counters(cnt, stats);
}

template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_within_fill(DevTree<A, K> t, const A* __restrict__ queries,
                                                        const uint32_t* __restrict__ order, uint64_t m,
                                                        typename Ax<A>::O radius, int excl, int sorted,
                                                        const uint64_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ out_items,
                                                        typename Ax<A>::O* __restrict__ out_d,
                                                        unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    const uint64_t o = offsets[qi], cap = offsets[qi + 1] - o;
    WithinVis<O> v{radius, excl != 0, 0, out_items + o, out_d + o, cap};
    traverse(t, q, qw, v, cnt);
    if (sorted) sort_segment(out_items + o, out_d + o, v.count < cap ? v.count : cap);
  }
  flush_counters(cnt, stats);
}

}  // namespace dev

namespace {

template <class A, int K> size_t lds_bytes(const dev::DevTree<A, K>& dt) { return size_t(dt.lds_n) * sizeof(A); }

template <class A, int K, int MAXK>
int launch_best(const QueryArgs& a, int n, double radius, uint32_t* items, void* dists, uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  auto dt = dev_tree<A, K>(a.t);
  reset_counters(a.t, a.s);
  KernelTimer kt(a.t, a.s);
  hipLaunchKernelGGL((dev::k_best_n_within<A, K, MAXK>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes(dt),
                     a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, n, O(radius),
                     (a.flags & KD_EXCLUSIVE) ? 1 : 0, items, static_cast<O*>(dists), counts, a.t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

}  // namespace

int launch_nearest_one(const QueryArgs& a, uint32_t* items, void* dists) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_nearest_one<A, KK>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes(dt), a.s, dt,
                       static_cast<const A*>(a.queries), a.order, a.m, items, static_cast<O*>(dists),
                       a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

int launch_approx_nearest_one(const QueryArgs& a, uint32_t* items, void* dists) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    if constexpr (std::is_floating_point<A>::value) {
      auto dt = dev_tree<A, KK>(a.t);
      reset_counters(a.t, a.s);
      KernelTimer kt(a.t, a.s);
      hipLaunchKernelGGL((dev::k_approx_nearest_one<A, KK>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes(dt),
                         a.s, dt, static_cast<const A*>(a.queries), a.order, a.m, items, static_cast<A*>(dists),
                         a.t->counters);
      KD_HIP(hipGetLastError());
    } else {
      return set_error(KD_ERR_UNSUPPORTED, "approx nearest_one needs a float tree (Output == A)");
    }
  });
  return KD_OK;
}

int launch_best_n_within(const QueryArgs& a, int n, double radius, uint32_t* items, void* dists,
                         uint32_t* counts) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    if (n <= 4) return launch_best<A, KK, 4>(a, n, radius, items, dists, counts);
    if (n <= 8) return launch_best<A, KK, 8>(a, n, radius, items, dists, counts);
    if (n <= 16) return launch_best<A, KK, 16>(a, n, radius, items, dists, counts);
    return launch_best<A, KK, 64>(a, n, radius, items, dists, counts);
  });
  return KD_OK;
}

int launch_within_count(const QueryArgs& a, double radius, uint64_t* counts) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    reset_counters(a.t, a.s);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_within_count<A, KK>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes(dt), a.s,
                       dt, static_cast<const A*>(a.queries), a.order, a.m, O(radius),
                       (a.flags & KD_EXCLUSIVE) ? 1 : 0, counts, a.t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

int launch_within_fill(const QueryArgs& a, double radius, const uint64_t* offsets, uint32_t* items,
                       void* dists) {
  KD_DISPATCH_AK(a.t->dtype, a.t->dims, {
    using O = typename dev::Ax<A>::O;
    auto dt = dev_tree<A, KK>(a.t);
    KernelTimer kt(a.t, a.s);
    hipLaunchKernelGGL((dev::k_within_fill<A, KK>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds_bytes(dt), a.s,
                       dt, static_cast<const A*>(a.queries), a.order, a.m, O(radius),
                       (a.flags & KD_EXCLUSIVE) ? 1 : 0, (a.flags & KD_UNSORTED) ? 0 : 1, offsets, items,
                       static_cast<O*>(dists), nullptr);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

}  // namespace kd
