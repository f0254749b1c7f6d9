// k_forest.hip — batched form: thousands of independent trees of uneven sizes, each with
// its own query set, answered by one persistent launch fed by a device work queue.
//
// All trees live in pooled HBM buffers (stems, leaf offsets, coordinates, items); a tree is
// a small descriptor into the pools. A plan cuts every tree's query range into work items of
// up to `chunk` queries, orders them by estimated cost (entries x queries, largest first, so
// the tail of the launch is made of small items), and optionally Morton-orders the queries
// inside each tree's range. Each block of the persistent kernel takes the next work item
// from an atomic counter until the list is exhausted, so every block exits on its own. The
// k-NN launch (queue reset + persistent kernel) can be captured once in a hipGraph and
// replayed. Per-query results are those of the single-tree kernels (same traversal and
// collectors), written at each query's index in the concatenated query array.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "knn_screen.h"
#include "tree.h"

namespace kd {
namespace dev {

struct ForestTree {
  uint64_t stems_off;  // into the stems pool (A units)
  uint64_t leaf_off;   // into the leaf_start pool (u32 units); entries are pool indices
  uint32_t stems_len;
  uint32_t n;
  int32_t nlev;
  int32_t pad;
};

struct WorkItem {
  uint32_t tree;
  uint32_t begin;  // first position in the (possibly permuted) query order
  uint32_t count;
  uint32_t pad;
};

template <class A, int K> struct ForestView {
  const A* stems;
  const uint32_t* leaf_start;
  const A* coords;
  const uint32_t* items;
  const float4* shadow;  // pooled FP32 shadow (forests created with KD_TREE_SHADOW_F32) or null
  uint64_t plane;
  const ForestTree* trees;

  __device__ DevTree<A, K> tree(uint32_t i) const {
    const ForestTree ft = trees[i];
    DevTree<A, K> t;
    t.stems = stems + ft.stems_off;
    t.leaf_start = leaf_start + ft.leaf_off;
    t.coords = coords;
    t.items = items;
    t.shadow = shadow;
    t.lds = nullptr;
    t.lds_n = 0;
    t.plane = plane;
    t.n = ft.n;
    t.nlev = ft.nlev;
    return t;
  }
};

// Takes the next work item for the whole block; returns false when the list is done.
__device__ __forceinline__ bool next_item(unsigned int* __restrict__ queue, uint32_t n_work, uint32_t& w) {
  __shared__ uint32_t s_w;
  __syncthreads();  // everyone has finished with the previous item and its LDS stems
  if (threadIdx.x == 0) s_w = atomicAdd(queue, 1u);
  __syncthreads();
  w = s_w;
  return w < n_work;
}

template <class A, int K>
__device__ __forceinline__ void pin_forest_stems(DevTree<A, K>& t, uint32_t stems_len, uint32_t lds_cap) {
  extern __shared__ __align__(16) unsigned char kd_lds_raw[];
  A* lds = reinterpret_cast<A*>(kd_lds_raw);
  const uint32_t n = stems_len < lds_cap ? stems_len : lds_cap;
  for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) lds[j] = t.stems[j];
  __syncthreads();
  t.lds = lds;
  t.lds_n = n;
}

template <class A, int K, int MAXK, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_forest_knn(
    ForestView<A, K> f, const WorkItem* __restrict__ work, uint32_t n_work, unsigned int* __restrict__ queue,
    const A* __restrict__ queries, const uint32_t* __restrict__ order, int k, typename Ax<A>::O radius,
    int excl, uint32_t lds_cap, uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d,
    uint32_t* __restrict__ out_counts, unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  Counters cnt;
  uint32_t w;
  while (next_item(queue, n_work, w)) {
    const WorkItem wi = work[w];
    DevTree<A, K> t = f.tree(wi.tree);
    if (lds_cap) pin_forest_stems(t, f.trees[wi.tree].stems_len, lds_cap);
    for (uint32_t j = threadIdx.x; j < wi.count; j += kBlock) {
      const uint64_t pos = uint64_t(wi.begin) + j;
      const uint64_t qi = order ? order[pos] : pos;
      Small<A, K> q;
      Small<O, K> qw;
      load_query<A, K>(queries, qi, q, qw);
      KnnVis<O, MAXK> v;
      v.init(k, radius, excl != 0);
      traverse(t, q, qw, v, cnt);
      uint32_t* ri = out_items + qi * uint64_t(k);
      O* rdd = out_d + qi * uint64_t(k);
#pragma unroll
      for (int jj = 0; jj < MAXK; ++jj)
        if (jj < k) {
          ri[jj] = jj < v.cnt ? v.top.it[jj] : 0xFFFFFFFFu;
          rdd[jj] = jj < v.cnt ? v.top.d[jj] : oinf<O>();
        }
      out_counts[qi] = uint32_t(v.cnt);
    }
  }
  flush_counters(cnt, stats);
}

// The FP32 collector (device.h Screen32Vis, exact on f32 trees) in the persistent form:
// each lane answers a query of the item with FP32 distances only and resolves items at
// the end. Queries whose candidate list overflowed get count kRedo and are answered by
// k_forest_knn_redo, launched after this kernel on the same stream. The candidate list is
// static LDS; the item's top stems take what is left of 40 KiB per block (lds_cap).
template <class A, int K, int MAXK, int WAVES>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(WAVES, 8))) void k_forest_knn_s32(
    ForestView<A, K> f, const WorkItem* __restrict__ work, uint32_t n_work, unsigned int* __restrict__ queue,
    const A* __restrict__ queries, const uint32_t* __restrict__ order, int k, typename Ax<A>::O radius,
    int excl, uint32_t lds_cap, uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d,
    uint32_t* __restrict__ out_counts, unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  __shared__ uint32_t cand[Screen32Vis<MAXK, A>::kCand * kBlock];
  Counters cnt;
  uint32_t w;
  while (next_item(queue, n_work, w)) {
    const WorkItem wi = work[w];
    DevTree<A, K> t = f.tree(wi.tree);
    if (lds_cap) pin_forest_stems(t, f.trees[wi.tree].stems_len, lds_cap);
    for (uint32_t j = threadIdx.x; j < wi.count; j += kBlock) {
      const uint64_t pos = uint64_t(wi.begin) + j;
      const uint64_t qi = order ? order[pos] : pos;
      Small<A, K> q;
      Small<O, K> qw;
      load_query<A, K>(queries, qi, q, qw);
      Screen32Vis<MAXK, A> v;
      screen_setup<K>(q, nullptr, v.qf, v.eps_k);
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
        for (int jj = 0; jj < MAXK; ++jj)
          if (jj < k) {
            ri[jj] = jj < rc ? res.it[jj] : 0xFFFFFFFFu;
            rdd[jj] = jj < rc ? res.d[jj] : oinf<O>();
          }
        out_counts[qi] = uint32_t(rc);
      }
      cnt.rescored += v.rescored;
    }
  }
  flush_counters(cnt, stats);
}

// The queries of k_forest_knn_s32 whose candidate list overflowed (count kRedo), answered
// by the plain collector on their tree: blocks stride over the work items, each lane over
// the item's queries; global stems only. The collector counts its overflows in the
// forest's counters (cleared by k_forest_reset before it, on the same stream): with none,
// the pass returns at once instead of reading every query's count (one forest runs one
// launch at a time: its work queue is shared as well).
template <class A, int K, int MAXK>
__global__ __launch_bounds__(kBlock) void k_forest_knn_redo(
    ForestView<A, K> f, const WorkItem* __restrict__ work, uint32_t n_work, const A* __restrict__ queries,
    const uint32_t* __restrict__ order, int k, typename Ax<A>::O radius, int excl,
    uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d, uint32_t* __restrict__ out_counts,
    unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  if (stats[kRedoWord] == 0ull) return;
  Counters cnt;
  for (uint32_t w = blockIdx.x; w < n_work; w += gridDim.x) {
    const WorkItem wi = work[w];
    for (uint32_t j = threadIdx.x; j < wi.count; j += kBlock) {
      const uint64_t pos = uint64_t(wi.begin) + j;
      const uint64_t qi = order ? order[pos] : pos;
      if (out_counts[qi] != kRedo) continue;
      const DevTree<A, K> t = f.tree(wi.tree);
      Small<A, K> q;
      Small<O, K> qw;
      load_query<A, K>(queries, qi, q, qw);
      KnnVis<O, MAXK> v;
      v.init(k, radius, excl != 0);
      traverse(t, q, qw, v, cnt);
      uint32_t* ri = out_items + qi * uint64_t(k);
      O* rdd = out_d + qi * uint64_t(k);
#pragma unroll
      for (int jj = 0; jj < MAXK; ++jj)
        if (jj < k) {
          ri[jj] = jj < v.cnt ? v.top.it[jj] : 0xFFFFFFFFu;
          rdd[jj] = jj < v.cnt ? v.top.d[jj] : oinf<O>();
        }
      out_counts[qi] = uint32_t(v.cnt);
    }
  }
  flush_counters(cnt, stats);
}

// Group-per-query form: G lanes per query (k <= G); a block answers kBlock / G queries of
// the item at a time.
template <class A, int K, int G>
__global__ __launch_bounds__(kBlock) void k_forest_knn_group(
    ForestView<A, K> f, const WorkItem* __restrict__ work, uint32_t n_work, unsigned int* __restrict__ queue,
    const A* __restrict__ queries, const uint32_t* __restrict__ order, int k, typename Ax<A>::O radius,
    int excl, uint32_t lds_cap, uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d,
    uint32_t* __restrict__ out_counts, unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  Counters cnt;
  const int gl = int(threadIdx.x) & (G - 1);
  const uint32_t slot = threadIdx.x / G;
  constexpr uint32_t per_block = kBlock / G;
  uint32_t w;
  while (next_item(queue, n_work, w)) {
    const WorkItem wi = work[w];
    DevTree<A, K> t = f.tree(wi.tree);
    if (lds_cap) pin_forest_stems(t, f.trees[wi.tree].stems_len, lds_cap);
    for (uint32_t j = slot; j < wi.count; j += per_block) {  // uniform within a group
      const uint64_t pos = uint64_t(wi.begin) + j;
      const uint64_t qi = order ? order[pos] : pos;
      Small<A, K> q;
      Small<O, K> qw;
      load_query<A, K>(queries, qi, q, qw);
      GroupKnnVis<A, K, G, false> v;
      v.init(k, radius, excl != 0, gl);
      Counters c1;
      group_traverse<A, K, G>(t, q, qw, v, c1, gl);
      if (gl == 0) {
        cnt.leaves += c1.leaves;
        cnt.points += c1.points;
        cnt.stems += c1.stems;
      }
      if (gl < k) {
        out_items[qi * uint64_t(k) + gl] = gl < v.cnt ? v.si : 0xFFFFFFFFu;
        out_d[qi * uint64_t(k) + gl] = gl < v.cnt ? v.sd : oinf<O>();
      }
      if (gl == 0) out_counts[qi] = uint32_t(v.cnt);
    }
  }
  flush_counters(cnt, stats);
}

template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_forest_within(
    ForestView<A, K> f, const WorkItem* __restrict__ work, uint32_t n_work, unsigned int* __restrict__ queue,
    const A* __restrict__ queries, const uint32_t* __restrict__ order, typename Ax<A>::O radius, int excl,
    int sorted, uint32_t lds_cap, uint64_t* __restrict__ counts, const uint64_t* __restrict__ offsets,
    uint32_t* __restrict__ out_items, typename Ax<A>::O* __restrict__ out_d, unsigned long long* __restrict__ stats) {
  using O = typename Ax<A>::O;
  Counters cnt;
  uint32_t w;
  while (next_item(queue, n_work, w)) {
    const WorkItem wi = work[w];
    DevTree<A, K> t = f.tree(wi.tree);
    if (lds_cap) pin_forest_stems(t, f.trees[wi.tree].stems_len, lds_cap);
    for (uint32_t j = threadIdx.x; j < wi.count; j += kBlock) {
      const uint64_t pos = uint64_t(wi.begin) + j;
      const uint64_t qi = order ? order[pos] : pos;
      Small<A, K> q;
      Small<O, K> qw;
      load_query<A, K>(queries, qi, q, qw);
      if (offsets) {
        const uint64_t o = offsets[qi], cap = offsets[qi + 1] - o;
        WithinVis<O> v{radius, excl != 0, 0, out_items + o, out_d + o, cap};
        traverse(t, q, qw, v, cnt);
        if (sorted) sort_segment(out_items + o, out_d + o, v.count < cap ? v.count : cap);
      } else {
        WithinVis<O> v{radius, excl != 0, 0, nullptr, nullptr, 0};
        traverse(t, q, qw, v, cnt);
        counts[qi] = v.count;
      }
    }
  }
  flush_counters(cnt, stats);
}

__global__ void k_forest_reset(unsigned int* queue, unsigned long long* stats) {
  if (threadIdx.x == 0) *queue = 0u;
  for (uint32_t i = threadIdx.x; i < uint32_t(kCounterWords); i += blockDim.x) stats[i] = 0ull;
}

// Morton keys tagged with the query's tree: sorting (tree, key) pairs reorders queries only
// inside each tree's range.
template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_forest_morton(const A* __restrict__ queries, uint64_t m,
                                                          const uint32_t* __restrict__ q_tree,
                                                          const double* __restrict__ bbox,  // 8 per tree
                                                          uint64_t* __restrict__ keys, uint32_t* __restrict__ ids) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= m) return;
  constexpr int bits = 32 / K;
  const uint32_t cap = (1u << bits) - 1u;
  const uint32_t tr = q_tree[i];
  const double* lo = bbox + 8 * uint64_t(tr);
  const double* sc = lo + 4;
  uint32_t key = 0, c[K];
#pragma unroll
  for (int d = 0; d < K; ++d) {
    double x = (double(queries[i * K + d]) - lo[d]) * sc[d];
    x = x > 0.0 ? x : 0.0;
    c[d] = x >= double(cap) ? cap : uint32_t(x);
  }
#pragma unroll
  for (int b = bits - 1; b >= 0; --b)
#pragma unroll
    for (int d = 0; d < K; ++d) key = (key << 1) | ((c[d] >> b) & 1u);
  keys[i] = (uint64_t(tr) << 32) | key;
  ids[i] = uint32_t(i);
}

}  // namespace dev
}  // namespace kd

// ========================================================================= host ========

struct kd_forest {
  int dtype = 0, dims = 0;
  uint32_t ntrees = 0;
  uint64_t total_points = 0, stems_total = 0, leaf_total = 0;
  void* stems = nullptr;
  uint32_t* leaf_start = nullptr;
  void* coords = nullptr;
  uint64_t plane = 0;
  uint32_t* items = nullptr;
  float4* shadow = nullptr;  // FP32 shadow of the pooled entries (KD_TREE_SHADOW_F32)
  kd::dev::ForestTree* trees = nullptr;
  std::vector<kd::dev::ForestTree> h_trees;
  std::vector<double> h_bbox;  // per tree lo[4], scale[4]
  double* bbox = nullptr;
  unsigned long long* counters = nullptr;
  unsigned int* queue = nullptr;
  uint64_t device_bytes = 0;
  // plan
  uint64_t m = 0;
  std::vector<uint64_t> q_off;
  kd::dev::WorkItem* work = nullptr;
  uint32_t n_work = 0, work_cap = 0;
  uint32_t* order = nullptr;
  uint64_t order_cap = 0;
  bool ordered = false;
  // tuning
  int variant = 1, group_size = 0, lds_levels = 0, waves = 0;
  uint32_t grid_blocks = 0;
  // graph
  hipStream_t gstream = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<uint64_t> gkey;
  // timing
  bool timing = false;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool ev_pending = false;
};

namespace kd {
namespace {

void free_forest(kd_forest* f) {
  if (!f) return;
  if (f->gexec) (void)hipGraphExecDestroy(f->gexec);
  if (f->graph) (void)hipGraphDestroy(f->graph);
  if (f->gstream) (void)hipStreamDestroy(f->gstream);
  for (hipEvent_t e : {f->ev_in, f->ev_out, f->ev_start, f->ev_stop})
    if (e) (void)hipEventDestroy(e);
  for (void* p : {f->stems, (void*)f->leaf_start, f->coords, (void*)f->items, (void*)f->shadow, (void*)f->trees, (void*)f->bbox,
                  (void*)f->counters, (void*)f->queue, (void*)f->work, (void*)f->order})
    if (p) (void)hipFree(p);
  delete f;
}

int fmalloc(void** p, size_t bytes, uint64_t& total) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  total += bytes;
  return KD_OK;
}

template <class F> void par_for(uint64_t count, F&& f) {
  int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (count < 64) threads = 1;
  std::vector<std::thread> th;
  std::atomic<uint64_t> next{0};
  for (int i = 0; i < threads; ++i)
    th.emplace_back([&] {
      for (;;) {
        uint64_t j = next.fetch_add(1);
        if (j >= count) break;
        f(j);
      }
    });
  for (auto& x : th) x.join();
}

template <class A, int K>
int forest_create_typed(const kd_tree_desc* descs, uint32_t ntrees, bool with_shadow, kd_forest** out) {
  kd_forest* f = new kd_forest;
  f->dtype = descs[0].dtype;
  f->dims = K;
  f->ntrees = ntrees;
  f->h_trees.resize(ntrees);
  f->h_bbox.assign(size_t(ntrees) * 8, 0.0);
  std::vector<uint64_t> pt_off(ntrees + 1, 0);
  const size_t entry_bytes = K * sizeof(A) + 4;
  uint64_t stems_total = 0, leaf_total = 0;
  for (uint32_t i = 0; i < ntrees; ++i) {
    const kd_tree_desc& d = descs[i];
    if (d.dtype != f->dtype || d.dims != K) {
      free_forest(f);
      return set_error(KD_ERR_INVALID, "tree %u has a different dtype or dims", i);
    }
    const int msl = d.max_stem_level;
    if (msl >= kMaxStemLevels || msl < -1) {
      free_forest(f);
      return set_error(KD_ERR_TOO_DEEP, "tree %u: max_stem_level %d unsupported", i, msl);
    }
    const uint64_t want = msl >= 0 ? (uint64_t(1) << (msl + 1)) : (d.size > 0 ? 1 : 0);
    if (d.leaf_count != want && !(msl < 0 && d.leaf_count <= 1)) {
      free_forest(f);
      return set_error(KD_ERR_INVALID, "tree %u: leaf_count does not match max_stem_level", i);
    }
    if (msl >= 0 && d.stems_len < (uint64_t(1) << (msl + 1))) {
      free_forest(f);
      return set_error(KD_ERR_INVALID, "tree %u: stems_len too short", i);
    }
    uint64_t tot = 0;
    for (uint64_t l = 0; l < d.leaf_count; ++l) {
      const uint64_t off = d.leaf_extents[2 * l], len = d.leaf_extents[2 * l + 1];
      if (len > uint64_t(kMaxLeafLen) || off + len * entry_bytes > d.leaf_bytes_len) {
        free_forest(f);
        return set_error(KD_ERR_INVALID, "tree %u: leaf %llu is invalid", i, (unsigned long long)l);
      }
      tot += len;
    }
    if (tot != d.size) {
      free_forest(f);
      return set_error(KD_ERR_INVALID, "tree %u: leaf lengths do not sum to its size", i);
    }
    auto& ft = f->h_trees[i];
    ft.stems_off = stems_total;
    ft.leaf_off = leaf_total;
    ft.stems_len = uint32_t(msl >= 0 ? (uint64_t(1) << (msl + 1)) : 0);
    ft.n = uint32_t(d.size);
    ft.nlev = msl + 1;
    stems_total += std::max<uint64_t>(d.stems_len, 1);
    leaf_total += d.leaf_count + 1;
    pt_off[i + 1] = pt_off[i] + d.size;
  }
  const uint64_t n = pt_off[ntrees];
  if (n > 0xFFFFFFFFull) {
    free_forest(f);
    return set_error(KD_ERR_INVALID, "a forest holds at most 2^32-1 entries in total");
  }
  f->total_points = n;
  f->stems_total = stems_total;
  f->leaf_total = leaf_total;
  const uint64_t plane = (n + 63) / 64 * 64;
  f->plane = plane;
  std::vector<A> stems(stems_total, A(0));
  std::vector<uint32_t> leaf_start(leaf_total, 0);
  std::vector<A> coords(plane * K);
  std::vector<uint32_t> items(std::max<uint64_t>(n, 1));
  std::atomic<bool> bad{false};
  par_for(ntrees, [&](uint64_t i) {
    const kd_tree_desc& d = descs[i];
    const auto& ft = f->h_trees[i];
    if (d.stems_len) std::memcpy(&stems[ft.stems_off], d.stems, d.stems_len * sizeof(A));
    uint64_t dst = pt_off[i];
    double lo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, hi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (uint64_t l = 0; l < d.leaf_count; ++l) {
      leaf_start[ft.leaf_off + l] = uint32_t(dst);
      const uint8_t* p = d.leaf_bytes + d.leaf_extents[2 * l];
      const uint64_t len = d.leaf_extents[2 * l + 1];
      uint64_t base = 0;
      while (base < len) {
        const uint64_t rem = len - base;
        const int tl = rem >= 32 ? 32 : rem >= 8 ? 8 : rem >= 4 ? 4 : rem >= 2 ? 2 : 1;
        for (int dd = 0; dd < K; ++dd)
          for (int j = 0; j < tl; ++j) {
            A v;
            std::memcpy(&v, p + (size_t(dd) * tl + j) * sizeof(A), sizeof(A));
            coords[dd * plane + dst + j] = v;
            const double x = double(v);
            if (std::isnan(x) || v == AxisHost<A>::max_value()) bad = true;
            lo[dd] = std::min(lo[dd], x);
            hi[dd] = std::max(hi[dd], x);
          }
        for (int j = 0; j < tl; ++j) std::memcpy(&items[dst + j], p + size_t(K) * tl * sizeof(A) + 4 * j, 4);
        p += size_t(K) * tl * sizeof(A) + 4 * size_t(tl);
        dst += tl;
        base += tl;
      }
    }
    leaf_start[ft.leaf_off + d.leaf_count] = uint32_t(dst);
    for (int dd = 0; dd < K; ++dd) {
      const double ext = d.size ? hi[dd] - lo[dd] : 0.0;
      f->h_bbox[8 * i + dd] = d.size ? lo[dd] : 0.0;
      f->h_bbox[8 * i + 4 + dd] = ext > 0 ? double((1u << (32 / K)) - 1u) / ext : 0.0;
    }
  });
  if (bad) {
    free_forest(f);
    return set_error(KD_ERR_INVALID, "a tree holds a NaN coordinate or the reserved maximum value");
  }
  uint64_t& tot = f->device_bytes;
  int rc = KD_OK;
  auto alloc_up = [&](void** dstp, const void* src, size_t bytes) {
    if (rc) return;
    rc = fmalloc(dstp, bytes, tot);
    if (rc == KD_OK && bytes && src && hipMemcpy(*dstp, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      rc = set_error(KD_ERR_DEVICE, "forest upload failed");
  };
  alloc_up(&f->stems, stems.data(), stems.size() * sizeof(A));
  alloc_up((void**)&f->leaf_start, leaf_start.data(), leaf_start.size() * 4);
  alloc_up(&f->coords, coords.data(), coords.size() * sizeof(A));
  alloc_up((void**)&f->items, items.data(), items.size() * 4);
  if (with_shadow) {
    // {x, y, z, item bits} for K <= 3, {x, y, z, w} for K = 4, zeroed padding past the end
    std::vector<float4> sh(n + dev::kShadowPad, float4{0.0f, 0.0f, 0.0f, 0.0f});
    par_for(ntrees, [&](uint64_t i) {
      for (uint64_t e = pt_off[i]; e < pt_off[i + 1]; ++e) {
        float4 v;
        v.x = float(coords[0 * plane + e]);
        v.y = float(coords[1 * plane + e]);
        v.z = K > 2 ? float(coords[2 * plane + e]) : 0.0f;
        if (K > 3)
          v.w = float(coords[3 * plane + e]);
        else
          std::memcpy(&v.w, &items[e], 4);
        sh[e] = v;
      }
    });
    alloc_up((void**)&f->shadow, sh.data(), sh.size() * sizeof(float4));
  }
  alloc_up((void**)&f->trees, f->h_trees.data(), f->h_trees.size() * sizeof(dev::ForestTree));
  alloc_up((void**)&f->bbox, f->h_bbox.data(), f->h_bbox.size() * sizeof(double));
  alloc_up((void**)&f->counters, nullptr, kd::dev::kCounterWords * sizeof(unsigned long long));
  alloc_up((void**)&f->queue, nullptr, 64);
  if (rc) {
    free_forest(f);
    return rc;
  }
  *out = f;
  return KD_OK;
}

template <class A, int K> dev::ForestView<A, K> view_of(const kd_forest* f) {
  dev::ForestView<A, K> v;
  v.stems = static_cast<const A*>(f->stems);
  v.leaf_start = f->leaf_start;
  v.coords = static_cast<const A*>(f->coords);
  v.items = f->items;
  v.shadow = f->shadow;
  v.plane = f->plane;
  v.trees = f->trees;
  return v;
}

uint32_t forest_lds_cap(const kd_forest* f, size_t a_size) {
  if (f->lds_levels <= 0) return 0;
  uint64_t cap = uint64_t(1) << f->lds_levels;
  while (cap * a_size > 16384) cap >>= 1;
  return uint32_t(cap);
}

template <class KernelFn>
uint32_t persistent_grid(KernelFn kernel, size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, dev::kBlock, lds);
  if (per_cu < 1) per_cu = 1;
  if (cus < 1) cus = 1;
  return uint32_t(cus * per_cu);
}

// Launches queue reset + the persistent k-NN kernel on stream s.
template <class A, int K, int MAXK>
int launch_forest_knn(kd_forest* f, const A* q, int k, double radius, uint32_t flags, uint32_t* items,
                      void* dists, uint32_t* counts, hipStream_t s) {
  using O = typename dev::Ax<A>::O;
  const uint32_t lds_cap = forest_lds_cap(f, sizeof(A));
  const size_t lds = size_t(lds_cap) * sizeof(A);
  const int excl = (flags & KD_EXCLUSIVE) ? 1 : 0;
  const uint32_t* order = f->ordered ? f->order : nullptr;
  hipLaunchKernelGGL(dev::k_forest_reset, dim3(1), dim3(64), 0, s, f->queue, f->counters);
  int g = f->group_size;
  if (g == 0) g = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
  if (f->variant == 2 && g >= k) {
#define KD_FOREST_GROUP(GG)                                                                           \
  case GG: {                                                                                          \
    auto kern = dev::k_forest_knn_group<A, K, GG>;                                                   \
    hipLaunchKernelGGL(kern, dim3(persistent_grid(kern, lds)), dim3(dev::kBlock), lds, s, view_of<A, K>(f), \
                       f->work, f->n_work, f->queue, q, order, k, O(radius), excl, lds_cap, items,     \
                       static_cast<O*>(dists), counts, f->counters);                                  \
    break;                                                                                            \
  }
    switch (g) {
      KD_FOREST_GROUP(4)
      KD_FOREST_GROUP(8)
      KD_FOREST_GROUP(16)
      KD_FOREST_GROUP(32)
      KD_FOREST_GROUP(64)
      default: return set_error(KD_ERR_INVALID, "bad group size %d", g);
    }
#undef KD_FOREST_GROUP
  } else if ((flags & KD_SCREEN_F32) && MAXK <= 16) {
    if constexpr (std::is_same<A, float>::value) {
      // the FP32 collector: its candidate list is static LDS, the stems take the rest of 40 KiB
      constexpr int MK = MAXK <= 16 ? MAXK : 16;
      constexpr size_t kList = size_t(dev::Screen32Vis<MK, A>::kCand) * 4 * dev::kBlock;
      uint32_t cap = lds_cap;
      while (cap && kList + size_t(cap) * sizeof(A) > 40960) cap >>= 1;
      const size_t lds_s = size_t(cap) * sizeof(A);
#define KD_FOREST_S32(W)                                                                              \
  {                                                                                                   \
    auto kern = dev::k_forest_knn_s32<A, K, MK, W>;                                                  \
    hipLaunchKernelGGL(kern, dim3(persistent_grid(kern, lds_s)), dim3(dev::kBlock), lds_s, s, view_of<A, K>(f), \
                       f->work, f->n_work, f->queue, q, order, k, O(radius), excl, cap, items,         \
                       static_cast<O*>(dists), counts, f->counters);                                  \
  }
      if (f->waves == 4) KD_FOREST_S32(4)
      else if (f->waves == 5) KD_FOREST_S32(5)
      else KD_FOREST_S32(3)
#undef KD_FOREST_S32
      KD_HIP(hipGetLastError());
      const uint32_t rg = std::min<uint32_t>(std::max<uint32_t>(f->n_work, 1u), 2048u);
      hipLaunchKernelGGL((dev::k_forest_knn_redo<A, K, MK>), dim3(rg), dim3(dev::kBlock), 0, s, view_of<A, K>(f),
                         f->work, f->n_work, q, order, k, O(radius), excl, items, static_cast<O*>(dists), counts,
                         f->counters);
    }
  } else {
#define KD_FOREST_THREAD(W)                                                                           \
  {                                                                                                   \
    auto kern = dev::k_forest_knn<A, K, MAXK, W>;                                                    \
    hipLaunchKernelGGL(kern, dim3(persistent_grid(kern, lds)), dim3(dev::kBlock), lds, s, view_of<A, K>(f), \
                       f->work, f->n_work, f->queue, q, order, k, O(radius), excl, lds_cap, items,     \
                       static_cast<O*>(dists), counts, f->counters);                                  \
  }
    if (f->waves == 4) KD_FOREST_THREAD(4)
    else if (f->waves == 5) KD_FOREST_THREAD(5)
    else KD_FOREST_THREAD(3)
#undef KD_FOREST_THREAD
  }
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class A, int K>
int forest_knn_typed(kd_forest* f, const A* q, int k, double radius, uint32_t flags, uint32_t* items,
                     void* dists, uint32_t* counts, hipStream_t s) {
  if (k <= 1) return launch_forest_knn<A, K, 1>(f, q, k, radius, flags, items, dists, counts, s);
  if (k <= 4) return launch_forest_knn<A, K, 4>(f, q, k, radius, flags, items, dists, counts, s);
  if (k <= 8) return launch_forest_knn<A, K, 8>(f, q, k, radius, flags, items, dists, counts, s);
  if (k <= 16) return launch_forest_knn<A, K, 16>(f, q, k, radius, flags, items, dists, counts, s);
  if (k <= 32) return launch_forest_knn<A, K, 32>(f, q, k, radius, flags, items, dists, counts, s);
  return launch_forest_knn<A, K, 64>(f, q, k, radius, flags, items, dists, counts, s);
}

int forest_knn_dispatch(kd_forest* f, const void* q, int k, double radius, uint32_t flags, uint32_t* items,
                        void* dists, uint32_t* counts, hipStream_t s) {
  KD_DISPATCH_AK(f->dtype, f->dims,
                 return (forest_knn_typed<A, KK>(f, static_cast<const A*>(q), k, radius, flags, items, dists, counts, s)));
  return KD_ERR_UNSUPPORTED;
}

struct ForestTimer {
  kd_forest* f;
  hipStream_t s;
  ForestTimer(kd_forest* f_, hipStream_t s_) : f(f_), s(s_) {
    if (f->timing) (void)hipEventRecord(f->ev_start, s);
  }
  ~ForestTimer() {
    if (f->timing) {
      (void)hipEventRecord(f->ev_stop, s);
      f->ev_pending = true;
    }
  }
};

}  // namespace
}  // namespace kd

extern "C" {

int kd_forest_create(const kd_tree_desc* descs, uint32_t ntrees, uint32_t tree_flags, kd_forest** out) {
  if (!descs || !out || ntrees == 0) return kd::set_error(KD_ERR_INVALID, "null argument or no trees");
  if (tree_flags & ~uint32_t(KD_TREE_SHADOW_F32)) return kd::set_error(KD_ERR_UNSUPPORTED, "unknown forest tree flags");
  const bool with_shadow = (tree_flags & KD_TREE_SHADOW_F32) != 0;
  if (with_shadow && descs[0].dtype != KD_F32)
    return kd::set_error(KD_ERR_UNSUPPORTED, "forests keep an FP32 shadow for f32 trees only");
  try {
    KD_DISPATCH_AK(descs[0].dtype, descs[0].dims,
                   return (kd::forest_create_typed<A, KK>(descs, ntrees, with_shadow, out)));
  } catch (const std::bad_alloc&) {
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed while uploading the forest");
  }
  return KD_ERR_INVALID;
}

int kd_forest_destroy(kd_forest* f) {
  if (!f) return KD_OK;
  (void)hipDeviceSynchronize();
  kd::free_forest(f);
  return KD_OK;
}

int kd_forest_plan(kd_forest* f, const void* queries, const uint64_t* query_offsets, uint32_t chunk, uint32_t flags,
                   void* stream) {
  if (!f || !query_offsets) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (chunk == 0) chunk = 256;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t m = query_offsets[f->ntrees];
  if (m > 0xFFFFFFFFull) return kd::set_error(KD_ERR_INVALID, "a plan holds at most 2^32-1 queries");
  for (uint32_t i = 0; i < f->ntrees; ++i)
    if (query_offsets[i + 1] < query_offsets[i]) return kd::set_error(KD_ERR_INVALID, "query_offsets decrease");
  f->m = m;
  f->q_off.assign(query_offsets, query_offsets + f->ntrees + 1);
  // work items, costliest first (entries of the tree x queries of the item)
  std::vector<kd::dev::WorkItem> w;
  std::vector<double> cost;
  for (uint32_t i = 0; i < f->ntrees; ++i)
    for (uint64_t b = query_offsets[i]; b < query_offsets[i + 1]; b += chunk) {
      const uint32_t c = uint32_t(std::min<uint64_t>(chunk, query_offsets[i + 1] - b));
      w.push_back({i, uint32_t(b), c, 0});
      cost.push_back(std::log2(double(f->h_trees[i].n) + 2.0) * c);
    }
  std::vector<uint32_t> idx(w.size());
  for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
  std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  std::vector<kd::dev::WorkItem> ws(w.size());
  for (size_t i = 0; i < idx.size(); ++i) ws[i] = w[idx[i]];
  if (ws.size() > f->work_cap) {
    if (f->work) (void)hipFree(f->work);
    f->work = nullptr;
    KD_HIP(hipMalloc(&f->work, std::max<size_t>(ws.size(), 1) * sizeof(kd::dev::WorkItem)));
    f->work_cap = uint32_t(ws.size());
  }
  if (!ws.empty()) KD_HIP(hipMemcpy(f->work, ws.data(), ws.size() * sizeof(kd::dev::WorkItem), hipMemcpyHostToDevice));
  f->n_work = uint32_t(ws.size());
  f->ordered = false;
  if ((flags & KD_MORTON) && m > 1 && queries) {
    if (m > f->order_cap) {
      if (f->order) (void)hipFree(f->order);
      f->order = nullptr;
      KD_HIP(hipMalloc(&f->order, m * sizeof(uint32_t)));
      f->order_cap = m;
    }
    std::vector<uint32_t> q_tree(m);
    for (uint32_t i = 0; i < f->ntrees; ++i)
      for (uint64_t b = query_offsets[i]; b < query_offsets[i + 1]; ++b) q_tree[b] = i;
    void* mem = nullptr;
    size_t temp = 0;
    uint64_t *k0 = nullptr, *k1 = nullptr;
    uint32_t *v0 = nullptr, *qt = nullptr;
    KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, k0, k1, v0, f->order, int(m), 0, 64, s));
    const size_t a8 = (m * 8 + 255) / 256 * 256, a4 = (m * 4 + 255) / 256 * 256;
    KD_HIP(hipMalloc(&mem, 2 * a8 + 2 * a4 + temp));
    uint8_t* base = static_cast<uint8_t*>(mem);
    k0 = reinterpret_cast<uint64_t*>(base);
    k1 = reinterpret_cast<uint64_t*>(base + a8);
    v0 = reinterpret_cast<uint32_t*>(base + 2 * a8);
    qt = reinterpret_cast<uint32_t*>(base + 2 * a8 + a4);
    int rc = KD_OK;
    if (hipMemcpy(qt, q_tree.data(), m * 4, hipMemcpyHostToDevice) != hipSuccess)
      rc = kd::set_error(KD_ERR_DEVICE, "upload of query trees failed");
    if (rc == KD_OK) {
      KD_DISPATCH_AK(f->dtype, f->dims, {
        hipLaunchKernelGGL((kd::dev::k_forest_morton<A, KK>), dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0, s,
                           static_cast<const A*>(queries), m, qt, f->bbox, k0, v0);
      });
      const int bits = 32 + (f->ntrees > 1 ? 32 - __builtin_clz(f->ntrees - 1) : 1);
      if (hipcub::DeviceRadixSort::SortPairs(base + 2 * a8 + 2 * a4, temp, k0, k1, v0, f->order, int(m), 0, bits,
                                             s) != hipSuccess)
        rc = kd::set_error(KD_ERR_DEVICE, "query sort failed");
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(mem);
    if (rc) return rc;
    f->ordered = true;
  }
  // a new plan invalidates a captured graph
  if (f->gexec) {
    (void)hipGraphExecDestroy(f->gexec);
    f->gexec = nullptr;
  }
  if (f->graph) {
    (void)hipGraphDestroy(f->graph);
    f->graph = nullptr;
  }
  return KD_OK;
}

int kd_forest_tune(kd_forest* f, int32_t variant, int32_t group_size, int32_t lds_levels, int32_t waves) {
  if (!f) return kd::set_error(KD_ERR_INVALID, "forest is null");
  if (variant < 0 || variant > 2) return kd::set_error(KD_ERR_INVALID, "variant must be 0, 1 or 2");
  if (group_size != 0 && group_size != 4 && group_size != 8 && group_size != 16 && group_size != 32 &&
      group_size != 64)
    return kd::set_error(KD_ERR_INVALID, "bad group size");
  if (lds_levels < 0 || lds_levels > 14) return kd::set_error(KD_ERR_INVALID, "lds_levels must be in [0, 14]");
  if (waves != 0 && (waves < 3 || waves > 5)) return kd::set_error(KD_ERR_INVALID, "waves must be 0 or 3..5");
  f->variant = variant == 0 ? 1 : variant;
  f->group_size = group_size;
  f->lds_levels = lds_levels;
  f->waves = waves;
  if (f->gexec) {
    (void)hipGraphExecDestroy(f->gexec);
    f->gexec = nullptr;
  }
  if (f->graph) {
    (void)hipGraphDestroy(f->graph);
    f->graph = nullptr;
  }
  return KD_OK;
}

int kd_forest_nearest_n(kd_forest* f, const void* queries, uint32_t k, double radius, uint32_t flags,
                        uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (!f) return kd::set_error(KD_ERR_INVALID, "forest is null");
  if (flags & 0x000FFF00u) return kd::set_error(KD_ERR_UNSUPPORTED, "batched forests answer SquaredEuclidean only");
  if (k == 0 || k > uint32_t(kd::kMaxK)) return kd::set_error(KD_ERR_INVALID, "k must be in [1, %d]", kd::kMaxK);
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  if (f->q_off.empty()) return kd::set_error(KD_ERR_INVALID, "call kd_forest_plan first");
  if (f->m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  if ((flags & KD_SCREEN_F32) && !f->shadow)
    return kd::set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 needs a forest created with KD_TREE_SHADOW_F32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (!(flags & KD_GRAPH)) {
    kd::ForestTimer ft(f, s);
    return kd::forest_knn_dispatch(f, queries, int(k), radius, flags, items, dists, counts, s);
  }
  // Graph mode: capture queue reset + persistent kernel once on the forest's own stream,
  // then replay it ordered after the caller's stream.
  if (!f->gstream) {
    KD_HIP(hipStreamCreateWithFlags(&f->gstream, hipStreamNonBlocking));
    KD_HIP(hipEventCreateWithFlags(&f->ev_in, hipEventDisableTiming));
    KD_HIP(hipEventCreateWithFlags(&f->ev_out, hipEventDisableTiming));
  }
  uint64_t rbits;
  std::memcpy(&rbits, &radius, 8);
  const std::vector<uint64_t> key = {uint64_t(queries), k, rbits, flags & ~uint32_t(KD_GRAPH), uint64_t(items),
                                     uint64_t(dists), uint64_t(counts)};
  if (!f->gexec || key != f->gkey) {
    if (f->gexec) (void)hipGraphExecDestroy(f->gexec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    f->gexec = nullptr;
    f->graph = nullptr;
    KD_HIP(hipStreamBeginCapture(f->gstream, hipStreamCaptureModeThreadLocal));
    int rc = kd::forest_knn_dispatch(f, queries, int(k), radius, flags, items, dists, counts, f->gstream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(f->gstream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return kd::set_error(KD_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    f->graph = g;
    KD_HIP(hipGraphInstantiate(&f->gexec, f->graph, nullptr, nullptr, 0));
    f->gkey = key;
  }
  KD_HIP(hipEventRecord(f->ev_in, s));
  KD_HIP(hipStreamWaitEvent(f->gstream, f->ev_in, 0));
  {
    kd::ForestTimer ft(f, f->gstream);
    KD_HIP(hipGraphLaunch(f->gexec, f->gstream));
  }
  KD_HIP(hipEventRecord(f->ev_out, f->gstream));
  KD_HIP(hipStreamWaitEvent(s, f->ev_out, 0));
  return KD_OK;
}

int kd_forest_within(kd_forest* f, const void* queries, double radius, uint32_t flags, uint64_t* offsets,
                     uint32_t* items, void* dists, void* stream) {
  // offsets[m+1] must be filled by a count call (items == NULL) before the fill call.
  if (!f) return kd::set_error(KD_ERR_INVALID, "forest is null");
  if (flags & 0x000FFF00u) return kd::set_error(KD_ERR_UNSUPPORTED, "batched forests answer SquaredEuclidean only");
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  if (f->q_off.empty()) return kd::set_error(KD_ERR_INVALID, "call kd_forest_plan first");
  if (!offsets) return kd::set_error(KD_ERR_INVALID, "offsets is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t m = f->m;
  const int excl = (flags & KD_EXCLUSIVE) ? 1 : 0;
  const int sorted = (flags & KD_UNSORTED) ? 0 : 1;
  const uint32_t* order = f->ordered ? f->order : nullptr;
  if (m >= 0x7FFFFFFFull) return kd::set_error(KD_ERR_UNSUPPORTED, "within plans hold < 2^31 queries");
  int rc = KD_OK;
  KD_DISPATCH_AK(f->dtype, f->dims, {
    using O = typename kd::dev::Ax<A>::O;
    const uint32_t lds_cap = kd::forest_lds_cap(f, sizeof(A));
    const size_t lds = size_t(lds_cap) * sizeof(A);
    auto kern = kd::dev::k_forest_within<A, KK>;
    const uint32_t grid = kd::persistent_grid(kern, lds);
    if (!items) {
      void* mem = nullptr;
      size_t temp = 0;
      uint64_t* cnt = nullptr;
      KD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, cnt, offsets, int(m + 1), s));
      const size_t arr = ((m + 1) * 8 + 255) / 256 * 256;
      KD_HIP(hipMallocAsync(&mem, arr + temp, s));
      cnt = static_cast<uint64_t*>(mem);
      hipLaunchKernelGGL(kd::dev::k_forest_reset, dim3(1), dim3(64), 0, s, f->queue, f->counters);
      {
        kd::ForestTimer ft(f, s);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(kd::dev::kBlock), lds, s, kd::view_of<A, KK>(f), f->work, f->n_work,
                           f->queue, static_cast<const A*>(queries), order, O(radius), excl, sorted, lds_cap, cnt,
                           nullptr, nullptr, nullptr, f->counters);
      }
      hipError_t e = hipcub::DeviceScan::ExclusiveSum(static_cast<uint8_t*>(mem) + arr, temp, cnt, offsets,
                                                      int(m + 1), s);
      (void)hipFreeAsync(mem, s);
      if (e != hipSuccess) rc = kd::set_error(KD_ERR_DEVICE, "scan failed: %s", hipGetErrorString(e));
    } else {
      hipLaunchKernelGGL(kd::dev::k_forest_reset, dim3(1), dim3(64), 0, s, f->queue, f->counters);
      kd::ForestTimer ft(f, s);
      hipLaunchKernelGGL(kern, dim3(grid), dim3(kd::dev::kBlock), lds, s, kd::view_of<A, KK>(f), f->work, f->n_work,
                         f->queue, static_cast<const A*>(queries), order, O(radius), excl, sorted, lds_cap, nullptr,
                         offsets, items, static_cast<O*>(dists), f->counters);
    }
    KD_HIP(hipGetLastError());
  });
  return rc;
}

int kd_forest_set_timing(kd_forest* f, int32_t enable) {
  if (!f) return kd::set_error(KD_ERR_INVALID, "forest is null");
  if (enable && !f->ev_start) {
    KD_HIP(hipEventCreate(&f->ev_start));
    KD_HIP(hipEventCreate(&f->ev_stop));
  }
  f->timing = enable != 0;
  f->ev_pending = false;
  return KD_OK;
}

int kd_forest_last_kernel_ms(kd_forest* f, float* ms) {
  if (!f || !ms) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (!f->timing || !f->ev_pending) return kd::set_error(KD_ERR_INVALID, "no timed call recorded");
  KD_HIP(hipEventSynchronize(f->ev_stop));
  KD_HIP(hipEventElapsedTime(ms, f->ev_start, f->ev_stop));
  return KD_OK;
}

int kd_forest_query_stats(kd_forest* f, uint64_t* out4) {
  if (!f || !out4) return kd::set_error(KD_ERR_INVALID, "null argument");
  std::vector<unsigned long long> h(4 * kd::dev::kStatSlots);
  KD_HIP(hipDeviceSynchronize());
  KD_HIP(hipMemcpy(h.data(), f->counters, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost));
  for (int j = 0; j < 4; ++j) out4[j] = 0;
  for (size_t i = 0; i < h.size(); ++i) out4[i % 4] += h[i];
  return KD_OK;
}

int kd_forest_autotune_nearest_n(kd_forest* f, const void* queries, uint32_t k, double radius, uint32_t flags,
                                 uint32_t* items, void* dists, uint32_t* counts, int32_t* chosen4, void* stream) {
  if (!f) return kd::set_error(KD_ERR_INVALID, "forest is null");
  struct Cand { int variant, group, lds, waves; };
  int g0 = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
  const std::vector<Cand> cands = {{1, 0, 0, 0}, {1, 0, 8, 0}, {1, 0, 11, 0}, {1, 0, 8, 4}, {2, g0, 8, 0}};
  const bool save_timing = f->timing;
  int rc = kd_forest_set_timing(f, 1);
  float best_ms = 1e30f;
  Cand best = cands[0];
  for (const Cand& c : cands) {
    if (rc) break;
    rc = kd_forest_tune(f, c.variant, c.group, c.lds, c.waves);
    float ms = 0.0f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      rc = kd_forest_nearest_n(f, queries, k, radius, flags & ~uint32_t(KD_GRAPH), items, dists, counts, stream);
      if (rc == KD_OK) rc = kd_forest_last_kernel_ms(f, &ms);
    }
    if (rc == KD_OK && ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  f->timing = save_timing;
  if (rc) return rc;
  rc = kd_forest_tune(f, best.variant, best.group, best.lds, best.waves);
  if (chosen4) {
    chosen4[0] = best.variant;
    chosen4[1] = best.group;
    chosen4[2] = best.lds;
    chosen4[3] = best.waves;
  }
  return rc;
}

// Bucket size of a batched forest by measurement: the given trees (a representative sample of
// the forest) are built on the host at each candidate bucket (threads over trees), uploaded as
// a forest, their queries planned (chunks of 256, Morton order inside each tree when flags
// hold KD_MORTON), the launch shape autotuned and nearest_n timed; the fastest bucket wins.
int kd_forest_autotune_bucket(int32_t dtype, int32_t dims, uint32_t ntrees, const void* const* points,
                              const uint64_t* sizes, const void* queries, const uint64_t* query_offsets,
                              uint32_t k, double radius, uint32_t flags, const int32_t* buckets,
                              uint32_t nbuckets, int32_t* best_bucket, float* ms_out, void* stream) {
  if (!points || !sizes || !queries || !query_offsets || !buckets || !best_bucket)
    return kd::set_error(KD_ERR_INVALID, "null argument");
  if (ntrees == 0 || nbuckets == 0) return kd::set_error(KD_ERR_INVALID, "no trees or no candidate bucket");
  if (k == 0 || k > uint32_t(kd::kMaxK)) return kd::set_error(KD_ERR_INVALID, "k must be in 1..%d", kd::kMaxK);
  for (uint32_t i = 0; i < nbuckets; ++i)
    if (buckets[i] < 1 || buckets[i] > 1024)
      return kd::set_error(KD_ERR_INVALID, "bucket %d is not in 1..1024", buckets[i]);
  const uint64_t m = query_offsets[ntrees];
  if (m == 0) return kd::set_error(KD_ERR_INVALID, "no queries");
  const size_t osz = m * k * (sizeof(uint32_t) + sizeof(double)) + m * sizeof(uint32_t) + 1024;
  void* mem = nullptr;
  if (hipMalloc(&mem, osz) != hipSuccess) return kd::set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed", osz);
  uint32_t* oitems = static_cast<uint32_t*>(mem);
  void* odists = static_cast<uint8_t*>(mem) + ((m * k * 4 + 255) / 256) * 256;
  uint32_t* ocounts = reinterpret_cast<uint32_t*>(static_cast<uint8_t*>(odists) + m * k * sizeof(double));
  const unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  float best_ms = 1e30f;
  int rc = KD_OK;
  *best_bucket = buckets[0];
  for (uint32_t b = 0; b < nbuckets && rc == KD_OK; ++b) {
    std::vector<kd_layout*> lays(ntrees, nullptr);
    std::vector<int> rcs(ntrees, KD_OK);
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < nth; ++w)
      pool.emplace_back([&, w]() {
        for (uint32_t i = w; i < ntrees; i += nth)
          rcs[i] = kd_build_layout(dtype, dims, buckets[b], sizes[i], points[i], nullptr, 1, &lays[i]);
      });
    for (auto& th : pool) th.join();
    for (uint32_t i = 0; i < ntrees && rc == KD_OK; ++i) rc = rcs[i];
    std::vector<kd_tree_desc> descs(ntrees);
    for (uint32_t i = 0; i < ntrees && rc == KD_OK; ++i) rc = kd_layout_desc(lays[i], &descs[i]);
    kd_forest* f = nullptr;
    // the collector's candidates need the FP32 shadow
    const uint32_t tflags = ((flags & KD_SCREEN_F32) && dtype == KD_F32) ? uint32_t(KD_TREE_SHADOW_F32) : 0u;
    if (rc == KD_OK) rc = kd_forest_create(descs.data(), ntrees, tflags, &f);
    for (kd_layout* l : lays)
      if (l) kd_layout_free(l);
    if (rc) break;
    rc = kd_forest_plan(f, queries, query_offsets, 256, flags & KD_MORTON, stream);
    if (rc == KD_OK)
      rc = kd_forest_autotune_nearest_n(f, queries, k, radius, flags, oitems, odists, ocounts, nullptr, stream);
    if (rc == KD_OK) rc = kd_forest_set_timing(f, 1);
    float tm = 1e30f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      float e = 0.0f;
      rc = kd_forest_nearest_n(f, queries, k, radius, flags & ~uint32_t(KD_GRAPH), oitems, odists, ocounts, stream);
      if (rc == KD_OK) rc = kd_forest_last_kernel_ms(f, &e);
      tm = std::min(tm, e);
    }
    (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
    kd_forest_destroy(f);
    if (rc) break;
    if (ms_out) ms_out[b] = tm;
    if (tm < best_ms) {
      best_ms = tm;
      *best_bucket = buckets[b];
    }
  }
  (void)hipFree(mem);
  return rc;
}

}  // extern "C"
