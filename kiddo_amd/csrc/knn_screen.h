// knn_screen.h — the k-NN kernel of the FP32 collector (device.h Screen32Vis) for k <= 16:
// FP32-screened on f64 trees, exact on f32 trees (both need the tree's FP32 shadow), and the
// exact redo pass for the queries whose candidate list overflowed (kiddo
// query/nearest_n.rs:21-38, nearest_n_within.rs:127-179, :441-479).
#pragma once

#include "families.h"

namespace kd {
namespace dev {

// Pre-scan of the FP32 collector: the query descends to its leaf (the traversal's rule: a
// sentinel stem sends it left) and takes the smallest block of 2^j sibling leaves holding 4k
// entries (on f64 trees 64 at least from k = 8 up; leaves under one stem are contiguous), whose FP32
// distances of certified in-radius entries fill top[] (all lanes of a wave do this in step).
// The block is then the traversal's first visit. `up` receives the block's height in levels,
// [b, e) its entries. The tree must not be empty.
template <class A, int K, int MAXK>
__device__ __forceinline__ void prescan_top(const DevTree<A, K>& t, const Small<A, K>& q, Screen32Vis<MAXK, A>& v,
                            int& up_out, uint32_t& b_out, uint32_t& e_out) {
  const int D = t.nlev;
  uint32_t idx = 1;
  int dim = 0, level = 0;
  auto step = [&](const A pivot) {
    idx = (idx << 1) | (((pivot < Ax<A>::maxv()) & (q[dim] >= pivot)) ? 1u : 0u);
    dim = (dim + 1 == K) ? 0 : dim + 1;
    ++level;
  };
  const int lds_levels = t.lds_n ? 31 - __clz(t.lds_n) : 0;
  const int mid = D < lds_levels ? D : lds_levels;
  while (level < mid) step(t.lds[idx]);
  while (level < D) step(t.stems[idx]);
  const uint32_t leaf = idx - (1u << D);
  // f64 trees: 64 entries at least from k = 8 up (measured best at k = 8 and 16 on a 16.8M-point
  // tree); f32 trees: 4k (nearest_n(8) over 4096 trees of 1k..262k points: 17.4 ms with 32
  // entries, 26.1 ms with 64)
  const uint32_t want = !Screen32Vis<MAXK, A>::kExact && v.k >= 8 && 4 * v.k < 64 ? 64u : uint32_t(4 * v.k);
  uint32_t b = 0, e = 0;
  int up = 0;
  for (;; ++up) {
    const uint32_t first = (leaf >> up) << up;
    b = t.leaf_start[first];
    e = t.leaf_start[first + (1u << up)];
    if (e - b >= want || up >= D) break;
  }
  constexpr uint32_t kBatch = 8;
  for (uint32_t p0 = b; p0 < e; p0 += kBatch) {
    float d32[kBatch];
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) {
      const float acc = shadow_sq<K>(t.shadow[p0 + u], v.qf);  // up to 7 past e: the shadow's padding
      d32[u] = (p0 + u < e) & (acc <= v.rlow) ? acc : __builtin_huge_valf();
    }
#pragma unroll
    for (uint32_t u = 0; u < kBatch; ++u) v.insert32(d32[u]);
  }
  up_out = up;
  b_out = b;
  e_out = e;
}

// nearest_n with k <= 16 with the FP32 collector (Screen32Vis): FP32-screened on f64 trees,
// exact on f32 trees. Queries whose candidate list overflowed get count kRedo and are
// answered by k_nearest_n_redo, launched after this kernel on the same stream.
constexpr uint32_t kRedo = 0xFFFFFFFFu;

template <class A, int K, int MAXK, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_nearest_n_screen32(
    DevTree<A, K> t, const A* __restrict__ queries, const uint32_t* __restrict__ order,
    uint64_t m, int k, typename Ax<A>::O radius, int excl, const double* __restrict__ max_abs,
    uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d, uint32_t* __restrict__ out_counts,
    unsigned long long* __restrict__ stats) {
  __shared__ uint32_t cand[Screen32Vis<MAXK, A>::kCand * kBlock];
  KD_QUERY_PROLOGUE
  if (active) {
    Screen32Vis<MAXK, A> v;
    screen_setup<K>(q, max_abs, v.qf, v.eps_k);
    v.init(k, radius, excl != 0, cand);
    int up = 0;
    uint32_t bb = 0, be = 0;
    if (t.n != 0) {
      prescan_top<A, K, MAXK>(t, q, v, up, bb, be);
      v.pre = true;
    }
    v.refresh();
    traverse(t, q, qw, v, cnt, up);
    uint32_t* ri = out_items + qi * uint64_t(k);
    O* rdd = out_d + qi * uint64_t(k);
    if (v.ovf) {
      out_counts[qi] = kRedo;
      atomicAdd(stats + kRedoWord, 1ull);
    } else {
      TopK<O, MAXK> res;
      int rc = 0;
      v.finish(t, qw, bb, be, res, rc);
#pragma unroll
      for (int j = 0; j < MAXK; ++j)
        if (j < k) {
          ri[j] = j < rc ? res.it[j] : 0xFFFFFFFFu;
          rdd[j] = j < rc ? res.d[j] : oinf<O>();
        }
      out_counts[qi] = uint32_t(rc);
    }
    cnt.rescored = v.rescored;
  }
  flush_counters(cnt, stats);
}

// Exact redo of the queries k_nearest_n_screen32 marked (count kRedo): the plain FP64
// traversal and collector, one thread per marked query; the grid strides over this call's
// counts in input order. The decision reads only this call's own counts: the tree's redo
// counter is shared by every call on the tree (and cleared by each), so a call running
// concurrently on another stream could clear it between this call's kernels.
template <class A, int K, int MAXK>
__global__ __launch_bounds__(kBlock) void k_nearest_n_redo(DevTree<A, K> t, const A* __restrict__ queries,
                                                           uint64_t m, int k, typename Ax<A>::O radius, int excl,
                                                           uint32_t* __restrict__ out_items,
                                                           typename Ax<A>::O* __restrict__ out_d,
                                                           uint32_t* __restrict__ out_counts,
                                                           unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  Counters cnt;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < m; i += stride) {
    if (out_counts[i] != kRedo) continue;
    Small<A, K> q;
    Small<O, K> qw;
    load_query<A, K>(queries, i, q, qw);
    KnnVis<O, MAXK> v;
    v.init(k, radius, excl != 0);
    traverse(t, q, qw, v, cnt);
    uint32_t* ri = out_items + i * uint64_t(k);
    O* rdd = out_d + i * uint64_t(k);
#pragma unroll
    for (int j = 0; j < MAXK; ++j)
      if (j < k) {
        ri[j] = j < v.cnt ? v.top.it[j] : 0xFFFFFFFFu;
        rdd[j] = j < v.cnt ? v.top.d[j] : oinf<O>();
      }
    out_counts[i] = uint32_t(v.cnt);
  }
  flush_counters(cnt, stats);
}

}  // namespace dev
}  // namespace kd
