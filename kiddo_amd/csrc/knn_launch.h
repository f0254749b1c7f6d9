// knn_launch.h — k-NN kernels: nearest_n / nearest_n().within
// (kiddo query/nearest_n.rs:21-38, nearest_n_within.rs:127-179, :441-479).
//
// Two kernel forms, chosen per call (kd_tree_tune, or automatically):
//   * thread per query: one lane walks the DFS and keeps the top-k in registers;
//   * group per query: G lanes share one query, split each leaf's entries, and keep the
//     top-k spread one slot per lane (device.h group_traverse / GroupKnnVis).
// Each form has an FP32-screened variant for f64 trees (KD_SCREEN_F32).

#pragma once

#include "knn_screen.h"
#include "metric_launch.h"

namespace kd {
namespace dev {

// Certified upper bound on the k-th smallest FP64 distance from q, found before the
// traversal: the FP32 distances of the entries of the query's own block of 2^j leaves
// (the smallest block holding 4k entries, 64 at least for k >= 8; leaves under one stem
// are contiguous) give a k-th smallest S32, and every one of those k entries has an FP64
// distance of at most
// (sqrt((S32 + 1e-44) (1+5u)) + sqrt(K) eps)^2 (1 + 2e-15): the FP32 sum of K <= 4
// non-negative squares is at least ||delta||^2 (1-u)^K - K 2^-150 (a product that
// underflows into the subnormal range loses up to 2^-150 absolute, a sum with a subnormal
// result is exact; 1e-44 > 4 * 2^-150, and 1 / (1-u)^4 < 1 + 5u), each FP32 coordinate
// difference is within eps of the exact one,
// and the FP64 evaluation is within (1 + 1e-15) of the exact square. Returns
// +inf when the tree holds fewer than k entries. `up` receives the block's height in
// levels (2^up leaves) and [b, e) its entries; 0 and an empty range with an infinite cap.
template <int K, int MAXK>
__device__ double prescan_cap(const DevTree<double, K>& t, const Small<double, K>& q, const float (&qf)[4],
                              double eps_k, int k, int& up_out, uint32_t& b_out, uint32_t& e_out) {
  up_out = 0;
  b_out = e_out = 0;
  if (t.n < uint32_t(k)) return __builtin_huge_val();
  const int D = t.nlev;
  uint32_t idx = 1;
  int dim = 0, level = 0;
  auto step = [&](const double pivot) {
    idx = (idx << 1) | (q[dim] >= pivot ? 1u : 0u);
    dim = (dim + 1 == K) ? 0 : dim + 1;
    ++level;
  };
  // the levels held in LDS, then the rest (one kind of memory per loop)
  const int lds_levels = t.lds_n ? 31 - __clz(t.lds_n) : 0;
  const int mid = D < lds_levels ? D : lds_levels;
  while (level < mid) step(t.lds[idx]);
  while (level < D) step(t.stems[idx]);
  const uint32_t leaf = idx - (1u << D);
  // 4k entries, and at least 64 from k = 8 up (measured: a 64-entry block is best at k = 8
  // and 16; for k <= 4 the smaller block is)
  const uint32_t want = k >= 8 && 4 * k < 64 ? 64u : uint32_t(4 * k);
  uint32_t b = 0, e = 0;
  int up = 0;
  for (;; ++up) {
    const uint32_t first = (leaf >> up) << up;
    b = t.leaf_start[first];
    e = t.leaf_start[first + (1u << up)];
    if (e - b >= want || up >= D) break;
  }
  if (e - b < uint32_t(k)) return __builtin_huge_val();
  float top[MAXK];
#pragma unroll
  for (int j = 0; j < MAXK; ++j) top[j] = __builtin_huge_valf();
  constexpr uint32_t kBatch = 8;
  for (uint32_t p0 = b; p0 < e; p0 += kBatch) {
    float d32[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      const float4 sh = t.shadow[p0 + u];  // up to 7 past e: the shadow's padding
      const float acc = shadow_sq<K>(sh, qf);
      d32[u] = p0 + u < e ? acc : __builtin_huge_valf();
    }
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      const float c = d32[u];
#pragma unroll
      for (int j = MAXK - 1; j > 0; --j) top[j] = med3(top[j - 1], c, top[j]);  // sorted insert
      top[0] = fminf(top[0], c);
    }
  }
  float kth = top[0];
#pragma unroll
  for (int j = 1; j < MAXK; ++j) kth = (j == k - 1) ? top[j] : kth;
  if (!(kth < __builtin_huge_valf())) return __builtin_huge_val();
  up_out = up;
  b_out = b;
  e_out = e;
  const double r = sqrt((double(kth) + 1e-44) * (1.0 + 5.0 * 0x1p-24)) * (1.0 + 1e-15) + eps_k;
  return r * r * (1.0 + 2e-15);
}

template <int K, int MAXK, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_nearest_n_screen(
    DevTree<double, K> t, const double* __restrict__ queries, const uint32_t* __restrict__ order,
    uint64_t m, int k, double radius, int excl, const double* __restrict__ max_abs,
    uint32_t* __restrict__ out_items, double* __restrict__ out_d, uint32_t* __restrict__ out_counts,
    unsigned long long* __restrict__ stats) {
  using A = double;
  KD_QUERY_PROLOGUE
  if (active) {
    ScreenKnnVis<MAXK> v;
    v.v.init(k, radius, excl != 0);
    screen_setup<K>(q, max_abs, v.qf, v.eps_k);
    // the pre-scan's block is then the first visit: its entries within the cap are the first
    // candidates of every lane (sorted in together before the traversal), and its leaves are
    // not visited again
    int up = 0;
    uint32_t bb = 0, be = 0;
    v.v.cap = prescan_cap<K, MAXK>(t, q, v.qf, v.eps_k, k, up, bb, be);
    v.thr = screen_threshold(v.v.bound(), v.eps_k);
    v.rescored = 0;
    v.visits = 0;
    v.qhead = 0;
    v.qn = 0;
    if (t.n != 0) v.first_block(t, qw, bb, be);
    traverse(t, q, qw, v, cnt, up);
    uint32_t* ri = out_items + qi * uint64_t(k);
    double* rdd = out_d + qi * uint64_t(k);
#pragma unroll
    for (int j = 0; j < MAXK; ++j)
      if (j < k) {
        ri[j] = j < v.v.cnt ? v.v.top.it[j] : 0xFFFFFFFFu;
        rdd[j] = j < v.v.cnt ? v.v.top.d[j] : __builtin_huge_val();
      }
    out_counts[qi] = uint32_t(v.v.cnt);
    cnt.rescored = v.rescored;
  }
  flush_counters(cnt, stats);
}

template <class A, int K, int G, bool SCREEN>
__global__ __launch_bounds__(kBlock) void k_knn_group(DevTree<A, K> t, const A* __restrict__ queries,
                                                      const uint32_t* __restrict__ order, uint64_t m,
                                                      int k, typename Ax<A>::O radius, int excl,
                                                      const double* __restrict__ max_abs,
                                                      uint32_t* __restrict__ out_items,
                                                      typename Ax<A>::O* __restrict__ out_d,
                                                      uint32_t* __restrict__ out_counts,
                                                      unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  pin_stems(t);
  const uint64_t gid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / G;
  const int gl = int(threadIdx.x) & (G - 1);
  const bool active = gid < m;  // uniform within a group
  const uint64_t qi = active ? (order ? order[gid] : gid) : 0;
  Small<A, K> q;
  Small<O, K> qw;
  Counters cnt;
  if (active) {
    load_query<A, K>(queries, qi, q, qw);
    GroupKnnVis<A, K, G, SCREEN> v;
    v.init(k, radius, excl != 0, gl);
    if constexpr (SCREEN) {
      if constexpr (std::is_same<A, double>::value) {
        screen_setup<K>(q, max_abs, v.qf, v.eps_k);
        v.thr = screen_threshold(double(v.bound()), v.eps_k);
      }
    }
    group_traverse<A, K, G>(t, q, qw, v, cnt, gl);
    if (gl < k) {
      out_items[qi * uint64_t(k) + gl] = gl < v.cnt ? v.si : 0xFFFFFFFFu;
      out_d[qi * uint64_t(k) + gl] = gl < v.cnt ? v.sd : oinf<O>();
    }
    if (gl == 0) out_counts[qi] = uint32_t(v.cnt);
    cnt.rescored = v.rescored;
    if (gl != 0) cnt.leaves = cnt.points = cnt.stems = 0;  // counted once per group
  }
  flush_counters(cnt, stats);
}

}  // namespace dev

namespace knn {

template <class A, int K, int MAXK, int WAVES>
int launch_thread_w(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists, uint32_t* counts);

template <class A, int K, int MAXK>
int launch_thread(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists, uint32_t* counts) {
  // register/occupancy trade-off of the screened kernel (chosen by the autotuner)
  switch (a.t->waves) {
    case 4: return launch_thread_w<A, K, MAXK, 4>(a, k, radius, items, dists, counts);
    case 5: return launch_thread_w<A, K, MAXK, 5>(a, k, radius, items, dists, counts);
    case 6: return launch_thread_w<A, K, MAXK, 6>(a, k, radius, items, dists, counts);
    default: return launch_thread_w<A, K, MAXK, 3>(a, k, radius, items, dists, counts);
  }
}

template <class A, int K, int MAXK, int WAVES>
int launch_thread_w(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists, uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  kd_tree* t = a.t;
  const int excl = (a.flags & KD_EXCLUSIVE) ? 1 : 0;
  const A* q = static_cast<const A*>(a.queries);
  auto dt = dev_tree<A, K>(t);
  const size_t lds = size_t(dt.lds_n) * sizeof(A);
  reset_counters(t, a.s);
  KernelTimer kt(t, a.s);
  constexpr bool kFloatTree = std::is_same<A, double>::value || std::is_same<A, float>::value;
  if constexpr (kFloatTree) {
  if ((a.flags & KD_SCREEN_F32) && MAXK <= 16) {
    // the FP32 collector (f64: screened, f32: exact). LDS: the candidate list plus the stems
    // within 40 KiB per block, so that 4 blocks (16 waves) fit a CU's 160 KiB: the stems take
    // the largest power of two that fits
    constexpr int MK = MAXK <= 16 ? MAXK : 16;
    constexpr size_t kList = size_t(dev::Screen32Vis<MK, A>::kCand) * 4 * dev::kBlock;
    auto st = dt;
    while (st.lds_n && kList + size_t(st.lds_n) * sizeof(A) > 40960) st.lds_n >>= 1;
    hipLaunchKernelGGL((dev::k_nearest_n_screen32<A, K, MK, WAVES>), dim3(grid_for(a.m)), dim3(dev::kBlock),
                       size_t(st.lds_n) * sizeof(A), a.s, st, q, a.order, a.m, k, O(radius), excl, t->d_max_abs,
                       items, static_cast<O*>(dists), counts, t->counters);
    KD_HIP(hipGetLastError());
    // the queries whose candidate list overflowed, answered exactly; global stems only
    auto rt = dt;
    rt.lds_n = 0;
    const uint32_t rg = std::min<uint32_t>(grid_for(a.m), 1024u);
    hipLaunchKernelGGL((dev::k_nearest_n_redo<A, K, MK>), dim3(rg), dim3(dev::kBlock), 0, a.s, rt, q, a.m, k,
                       O(radius), excl, items, static_cast<O*>(dists), counts, t->counters);
    KD_HIP(hipGetLastError());
    return KD_OK;
  }
  }
  if constexpr (std::is_same<A, double>::value) {
    if (a.flags & KD_SCREEN_F32) {
      hipLaunchKernelGGL((dev::k_nearest_n_screen<K, MAXK, WAVES>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds, a.s,
                         dt, q, a.order, a.m, k, radius, excl, t->d_max_abs, items,
                         static_cast<double*>(dists), counts, t->counters);
      KD_HIP(hipGetLastError());
      return KD_OK;
    }
  }
  hipLaunchKernelGGL((dev::k_nearest_n<A, K, MAXK, dev::SqEuclidean>), dim3(grid_for(a.m)), dim3(dev::kBlock), lds,
                     a.s, dt, q, a.order, a.m, dev::SqEuclidean{}, k, O(radius), excl, items, static_cast<O*>(dists),
                     counts, t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class A, int K, int G>
int launch_group(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists, uint32_t* counts) {
  using O = typename dev::Ax<A>::O;
  kd_tree* t = a.t;
  const int excl = (a.flags & KD_EXCLUSIVE) ? 1 : 0;
  const A* q = static_cast<const A*>(a.queries);
  auto dt = dev_tree<A, K>(t);
  const size_t lds = size_t(dt.lds_n) * sizeof(A);
  const uint32_t grid = grid_for(a.m * G);
  reset_counters(t, a.s);
  KernelTimer kt(t, a.s);
  // the group form screens f64 trees only; f32 trees take its plain kernel (the same results)
  if ((a.flags & KD_SCREEN_F32) && std::is_same<A, double>::value) {
    if constexpr (std::is_same<A, double>::value) {
      hipLaunchKernelGGL((dev::k_knn_group<A, K, G, true>), dim3(grid), dim3(dev::kBlock), lds, a.s, dt, q,
                         a.order, a.m, k, O(radius), excl, t->d_max_abs, items, static_cast<O*>(dists), counts,
                         t->counters);
    }
  } else {
    hipLaunchKernelGGL((dev::k_knn_group<A, K, G, false>), dim3(grid), dim3(dev::kBlock), lds, a.s, dt, q,
                       a.order, a.m, k, O(radius), excl, t->d_max_abs, items, static_cast<O*>(dists), counts,
                       t->counters);
  }
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class A, int K>
int launch_typed(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists, uint32_t* counts) {
  kd_tree* t = a.t;
  if (a.flags & KD_SCREEN_F32) {
    if (!(std::is_same<A, double>::value || std::is_same<A, float>::value))
      return set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 applies to f64 and f32 trees");
    if (!t->shadow)
      return set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 needs a tree created with KD_TREE_SHADOW_F32");
  }
  // above the register forms: the top-k in a heap in each query's output row
  if (k > kMaxK) {
    if constexpr (std::is_same<A, double>::value)
      if (a.flags & KD_SCREEN_F32) return mlaunch_knn_heap<A, K, dev::SqEuclidean, true>(a, {}, k, radius, items, dists, counts);
    return mlaunch_knn_heap<A, K, dev::SqEuclidean, false>(a, {}, k, radius, items, dists, counts);
  }
  // group form: k must fit in one slot per lane. Automatic (knn_variant 0): small batches
  // (small_batch) with k <= 16 take 8 lanes per query (k <= 8) or 16.
  int g = t->group_size;
  const bool small = t->knn_variant == 0 && k <= 16 && small_batch(t, a.m);
  if (g == 0) g = k <= 4 ? (small ? 8 : 4) : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
  const bool group = (t->knn_variant == 2 || small) && g >= k;
  if (group) {
    switch (g) {
      case 4: return launch_group<A, K, 4>(a, k, radius, items, dists, counts);
      case 8: return launch_group<A, K, 8>(a, k, radius, items, dists, counts);
      case 16: return launch_group<A, K, 16>(a, k, radius, items, dists, counts);
      case 32: return launch_group<A, K, 32>(a, k, radius, items, dists, counts);
      case 64: return launch_group<A, K, 64>(a, k, radius, items, dists, counts);
      default: return set_error(KD_ERR_INVALID, "group size %d is not one of 4, 8, 16, 32, 64", g);
    }
  }
  if (k <= 1) return launch_thread<A, K, 1>(a, k, radius, items, dists, counts);
  if (k <= 4) return launch_thread<A, K, 4>(a, k, radius, items, dists, counts);
  if (k <= 8) return launch_thread<A, K, 8>(a, k, radius, items, dists, counts);
  if (k <= 16) return launch_thread<A, K, 16>(a, k, radius, items, dists, counts);
  if (k <= 32) return launch_thread<A, K, 32>(a, k, radius, items, dists, counts);
  return launch_thread<A, K, 64>(a, k, radius, items, dists, counts);
}

}  // namespace knn

}  // namespace kd
