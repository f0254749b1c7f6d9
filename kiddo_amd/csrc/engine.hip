// engine.hip — MI355X (gfx950) batch query engine for Kiddo's ImmutableKdTree.
//
// One HIP thread answers one query with a depth-first traversal that takes the same
// decisions as Kiddo's arithmetic DFS (kiddo src/kd_tree/orchestrator/mod.rs:388-603):
// near child first, far child queued when its incremental rect distance
// rd' = (rd - old_off^2) + new_off^2 (src/dist/distance_metric_core.rs:89-98) passes the
// current bound, and queued far children re-checked against the bound of the moment
// when they are popped (:472-483). Kiddo keeps a restore stack and a far stack; here the
// stack is one slot per stem level plus two bit masks, because the Eytzinger path of the
// last leaf gives every ancestor index (eytzinger/mod.rs:131-168) and an off[] entry
// only changes when a far child is entered, so only those levels need a saved value.
//
// Distances are SquaredEuclidean in Kiddo's scalar order: ((0 + d0*d0) + d1*d1) + ...,
// each product and sum rounded separately (squared_euclidean/mod.rs:36-39,
// distance_metric_core.rs:38-40); the explicit *_rn intrinsics keep the compiler from
// fusing them, and the build also passes -ffp-contract=off.
//
// Results use the canonical order (distance, then lower item): the collectors accept an
// equal-distance candidate with a lower item, and an equal-distance subtree is only pruned
// by an exclusive radius, so the answer does not depend on within-leaf order.
//
// HBM layout of a tree (built by kd_tree_create from Kiddo's arena layout):
//   stems[stems_len]        A, Eytzinger order exactly as Kiddo's
//   leaf_start[L+1]         u32 prefix offsets of the 2^d leaves into the point arrays
//   coords[K][plane]        A, structure of arrays, leaves contiguous in leaf order
//   items[n]                u32
//   shadow[n]               float4 {x, y, z, item bits} (f64 trees with an FP32 shadow)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "kiddo_amd_internal.h"

// ------------------------------------------------------------------- errors --------
namespace kd {
namespace {
thread_local std::string g_last_error;
}
int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
}  // namespace kd

#define KD_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return kd::set_error(KD_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------ device numerics --------
namespace kd {
namespace dev {

template <class A> struct Ax;
template <> struct Ax<float> {
  using O = float;
  static __device__ __host__ constexpr float maxv() { return __builtin_huge_valf(); }
};
template <> struct Ax<double> {
  using O = double;
  static __device__ __host__ constexpr double maxv() { return __builtin_huge_val(); }
};
template <> struct Ax<uint16_t> {
  using O = double;
  static __device__ __host__ constexpr uint16_t maxv() { return 0xFFFFu; }
};
template <> struct Ax<uint32_t> {
  using O = double;
  static __device__ __host__ constexpr uint32_t maxv() { return 0xFFFFFFFFu; }
};

template <class O> __device__ __host__ constexpr O oinf();
template <> __device__ __host__ constexpr float oinf<float>() { return __builtin_huge_valf(); }
template <> __device__ __host__ constexpr double oinf<double>() { return __builtin_huge_val(); }

__device__ __forceinline__ float add_rn(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float sub_rn(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ float mul_rn(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ double add_rn(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double sub_rn(double a, double b) { return __dsub_rn(a, b); }
__device__ __forceinline__ double mul_rn(double a, double b) { return __dmul_rn(a, b); }

// |a - b| in O (Axis::saturating_dist for floats, axis.rs:124; exact for widened ints)
template <class O> __device__ __forceinline__ O abs_diff(O a, O b) {
  return a >= b ? sub_rn(a, b) : sub_rn(b, a);
}

// Register-resident small arrays indexed by a run-time dimension.
template <class T, int K> __device__ __forceinline__ T pick(const T (&a)[K], int d) {
  T r = a[0];
#pragma unroll
  for (int j = 1; j < K; ++j)
    if (d == j) r = a[j];
  return r;
}
template <class T, int K> __device__ __forceinline__ void put(T (&a)[K], int d, T v) {
#pragma unroll
  for (int j = 0; j < K; ++j)
    if (d == j) a[j] = v;
}

constexpr int kMaxLevels = 32;

template <class A, int K> struct DevTree {
  const A* __restrict__ stems;
  const uint32_t* __restrict__ leaf_start;
  const A* __restrict__ coords;  // K planes of `plane` entries
  const uint32_t* __restrict__ items;
  const float4* __restrict__ shadow;
  uint64_t plane;
  uint32_t n;
  int nlev;  // number of stem levels = max_stem_level + 1
};

// Squared distance of entry p in Kiddo's operation order.
template <class A, int K>
__device__ __forceinline__ typename Ax<A>::O point_dist(const DevTree<A, K>& t, uint32_t p,
                                                        const typename Ax<A>::O (&qw)[K]) {
  using O = typename Ax<A>::O;
  O acc = O(0);
#pragma unroll
  for (int d = 0; d < K; ++d) {
    const O c = O(t.coords[d * t.plane + p]);
    const O df = sub_rn(c, qw[d]);
    acc = add_rn(acc, mul_rn(df, df));
  }
  return acc;
}

// Per-thread traversal counters, summed per wave into the tree's counters: leaves
// visited, entries scanned, stem nodes read, entries re-measured in FP64 (screen only).
struct Counters {
  uint32_t leaves = 0, points = 0, stems = 0, rescored = 0;
};

// ------------------------------------------------------------------ traversal --------
// V supplies bound(), prune(rd), unbounded() and leaf(tree, qw, leaf_index).
template <class A, int K, class V>
__device__ __forceinline__ void traverse(const DevTree<A, K>& t, const A (&q)[K],
                                         const typename Ax<A>::O (&qw)[K], V& vis, Counters& c) {
  using O = typename Ax<A>::O;
  if (t.n == 0) return;  // orchestrator/mod.rs:399-401
  const int D = t.nlev;
  O off[K];
#pragma unroll
  for (int j = 0; j < K; ++j) off[j] = O(0);
  O slot[kMaxLevels];   // far rd while pending, saved off[dim] while inside the far child
  uint32_t pending = 0, taken = 0;
  O rd = O(0);
  uint32_t idx = 1;
  int level = 0, dim = 0;
  bool check = !vis.unbounded();
  for (;;) {
    // descend to a leaf (orchestrator/mod.rs:520-603)
    c.stems += uint32_t(D - level);
    while (level < D) {
      const A pivot = t.stems[idx];
      if (pivot < Ax<A>::maxv()) {
        const bool right = pick(q, dim) >= pivot;
        const O qd = pick(qw, dim);
        const O no = abs_diff(qd, O(pivot));
        const O oo = pick(off, dim);
        const O rdf = add_rn(sub_rn(rd, mul_rn(oo, oo)), mul_rn(no, no));
        if (!check || !(rdf > vis.bound())) {
          slot[level] = rdf;
          pending |= 1u << level;
        }
        idx = (idx << 1) | (right ? 1u : 0u);
      } else {
        idx <<= 1;  // sentinel stem: only the left child holds entries
      }
      ++level;
      dim = (dim + 1 == K) ? 0 : dim + 1;
    }
    {
      const uint32_t leaf = idx - (1u << D);
      c.leaves += 1;
      c.points += t.leaf_start[leaf + 1] - t.leaf_start[leaf];
      vis.leaf(t, qw, leaf);
    }
    // pop frames from the deepest level upwards (orchestrator/mod.rs:450-515)
    bool entered = false;
    while (pending | taken) {
      const int l = 31 - __clz(pending | taken);
      const uint32_t bit = 1u << l;
      const int ld = l % K;
      if (taken & bit) {
        put(off, ld, slot[l]);
        taken ^= bit;
        continue;
      }
      pending ^= bit;
      const O frd = slot[l];
      if (vis.prune(frd)) continue;
      const A pivot = t.stems[idx >> (D - l)];
      c.stems += 1;
      const O no = abs_diff(pick(qw, ld), O(pivot));
      slot[l] = pick(off, ld);
      taken |= bit;
      put(off, ld, no);
      rd = frd;
      idx = (idx >> (D - l - 1)) ^ 1u;  // far child of the level-l ancestor
      level = l + 1;
      dim = level % K;
      check = true;
      entered = true;
      break;
    }
    if (!entered) return;
  }
}

// ------------------------------------------------------------------- visitors --------

// nearest_one: best (distance, item) so far; bound = best distance
// (nearest_one.rs:330-352). Kiddo prunes on equal and keeps the first of equal minima;
// the canonical rule keeps equal-distance subtrees and takes the lower item.
template <class O> struct NearestOneVis {
  O best;
  uint32_t item;
  __device__ O bound() const { return best; }
  __device__ bool prune(O rd) const { return rd > best; }
  __device__ bool unbounded() const { return true; }
  template <class A, int K>
  __device__ void leaf(const DevTree<A, K>& t, const O (&qw)[K], uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p = s; p < e; ++p) {
      const O d = point_dist(t, p, qw);
      if (d <= best) {
        const uint32_t it = t.items[p];
        if (d < best || it < item) {
          best = d;
          item = it;
        }
      }
    }
  }
};

// Sorted top-k in registers, ordered by (distance, item). Unused slots hold (+inf, ~0u).
template <class O, int MAXK> struct TopK {
  O d[MAXK];
  uint32_t it[MAXK];
  __device__ void init() {
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      d[j] = oinf<O>();
      it[j] = 0xFFFFFFFFu;
    }
  }
  static __device__ __forceinline__ bool less(O ad, uint32_t ai, O bd, uint32_t bi) {
    return ad < bd || (ad == bd && ai < bi);
  }
  // inserts c, dropping the last slot
  __device__ __forceinline__ void insert(O cd, uint32_t ci) {
    bool lt[MAXK];
#pragma unroll
    for (int j = 0; j < MAXK; ++j) lt[j] = less(cd, ci, d[j], it[j]);
#pragma unroll
    for (int j = MAXK - 1; j > 0; --j) {
      const bool sh = lt[j - 1];
      const bool here = lt[j];
      d[j] = sh ? d[j - 1] : (here ? cd : d[j]);
      it[j] = sh ? it[j - 1] : (here ? ci : it[j]);
    }
    if (lt[0]) {
      d[0] = cd;
      it[0] = ci;
    }
  }
  __device__ __forceinline__ void kth(int k, O& kd, uint32_t& kit) const {
    if (k == MAXK) {
      kd = d[MAXK - 1];
      kit = it[MAXK - 1];
      return;
    }
    kd = d[0];
    kit = it[0];
#pragma unroll
    for (int j = 1; j < MAXK; ++j)
      if (j == k - 1) {
        kd = d[j];
        kit = it[j];
      }
  }
};

// nearest_n / nearest_n().within (nearest_n_within.rs:441-479): bound = min(k-th distance
// when full, radius); leaf filter `<= radius` (`<` exclusive); a full collector accepts a
// candidate that is smaller in (distance, item) order.
template <class O, int MAXK> struct KnnVis {
  TopK<O, MAXK> top;
  int k, cnt;
  O radius;
  bool excl;
  O kd;
  uint32_t kit;
  __device__ void init(int k_, O r, bool ex) {
    top.init();
    k = k_;
    cnt = 0;
    radius = r;
    excl = ex;
    kd = oinf<O>();
    kit = 0xFFFFFFFFu;
  }
  __device__ O bound() const { return (cnt >= k && kd < radius) ? kd : radius; }
  __device__ bool prune(O rd) const { return rd > bound() || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  __device__ __forceinline__ void offer(O d, uint32_t it) {
    top.insert(d, it);
    cnt = cnt < k ? cnt + 1 : k;
    if (cnt >= k) top.kth(k, kd, kit);
  }
  template <class A, int K>
  __device__ void leaf(const DevTree<A, K>& t, const O (&qw)[K], uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p = s; p < e; ++p) {
      const O d = point_dist(t, p, qw);
      const bool ok = excl ? d < radius : d <= radius;
      if (!ok) continue;
      if (cnt >= k && !(d <= kd)) continue;
      const uint32_t it = t.items[p];
      if (cnt >= k && !TopK<O, MAXK>::less(d, it, kd, kit)) continue;
      offer(d, it);
    }
  }
};

// within (count / fill): bound = radius, prune on equal iff exclusive
// (within_unsorted.rs:267-286).
template <class O> struct WithinVis {
  O radius;
  bool excl;
  uint64_t count;
  uint32_t* out_items;  // null: count only
  O* out_d;
  uint64_t cap;
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  template <class A, int K>
  __device__ void leaf(const DevTree<A, K>& t, const O (&qw)[K], uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p = s; p < e; ++p) {
      const O d = point_dist(t, p, qw);
      const bool ok = excl ? d < radius : d <= radius;
      if (!ok) continue;
      if (out_items && count < cap) {
        out_items[count] = t.items[p];
        out_d[count] = d;
      }
      ++count;
    }
  }
};

// best_n_within (best_n_within.rs:163-196, result_collection.rs:463-561): the n smallest
// items within the radius; registers hold them sorted by item.
template <class O, int MAXK> struct BestNVis {
  O d[MAXK];
  uint32_t it[MAXK];
  int n, cnt;
  O radius;
  bool excl;
  uint32_t kit;
  __device__ void init(int n_, O r, bool ex) {
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      d[j] = oinf<O>();
      it[j] = 0xFFFFFFFFu;
    }
    n = n_;
    cnt = 0;
    radius = r;
    excl = ex;
    kit = 0xFFFFFFFFu;
  }
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  template <class A, int K>
  __device__ void leaf(const DevTree<A, K>& t, const O (&qw)[K], uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p = s; p < e; ++p) {
      const O dd = point_dist(t, p, qw);
      const bool ok = excl ? dd < radius : dd <= radius;
      if (!ok) continue;
      const uint32_t ci = t.items[p];
      if (cnt >= n && !(ci < kit)) continue;
      bool lt[MAXK];
#pragma unroll
      for (int j = 0; j < MAXK; ++j) lt[j] = ci < it[j];
#pragma unroll
      for (int j = MAXK - 1; j > 0; --j) {
        const bool sh = lt[j - 1], here = lt[j];
        d[j] = sh ? d[j - 1] : (here ? dd : d[j]);
        it[j] = sh ? it[j - 1] : (here ? ci : it[j]);
      }
      if (lt[0]) {
        d[0] = dd;
        it[0] = ci;
      }
      cnt = cnt < n ? cnt + 1 : n;
      if (cnt >= n) {
        kit = it[0];
#pragma unroll
        for (int j = 1; j < MAXK; ++j)
          if (j == n - 1) kit = it[j];
      }
    }
  }
};

// FP32-screened k-NN for f64 trees. Every leaf entry is first measured in FP32 from the
// shadow copy; only entries whose FP32 distance is within the certified error bound of the
// current FP64 bound are re-measured in FP64 and offered to the collector, so the result
// is the plain kernel's. Bound derivation, with u = 2^-24 and eta = 2^-150:
//   |fl32(p) - p| <= u|p| + eta, so each FP32 coordinate difference delta_i satisfies
//   |delta_i - (p_i - q_i)| <= u(2+u)(|p_i| + |q_i|) + 3 eta <= eps (per query, from the
//   largest |coordinate| of the tree and the query), hence ||delta|| <= ||t|| + sqrt(K) eps.
//   The FP32 sum of squares of non-negative terms is at most ||delta||^2 (1+u)^K, and the
//   exact distance is at most the FP64 one times (1 + 1e-15). So FP64 dist <= T implies
//   FP32 dist <= (sqrt(T (1+1e-15)) + sqrt(K) eps)^2 (1 + 4u) + 1e-44, rounded up.
__device__ __forceinline__ float screen_threshold(double T, double eps_k) {
  if (!(T < __builtin_huge_val())) return __builtin_huge_valf();
  const double r = sqrt(T * (1.0 + 1e-15)) + eps_k;
  const double tt = r * r * (1.0 + 4.0 * 0x1p-24) + 1e-44;
  float f = (float)tt;
  if ((double)f < tt) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up (f > 0)
  return f;
}

template <int MAXK> struct ScreenKnnVis {
  KnnVis<double, MAXK> v;
  float qf[3];
  double eps_k;
  float thr;
  uint32_t rescored;
  __device__ double bound() const { return v.bound(); }
  __device__ bool prune(double rd) const { return v.prune(rd); }
  __device__ bool unbounded() const { return v.unbounded(); }
  template <int K>
  __device__ void leaf(const DevTree<double, K>& t, const double (&qw)[K], uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p = s; p < e; ++p) {
      const float4 sh = t.shadow[p];
      float acc;
      {
        const float dx = sub_rn(sh.x, qf[0]);
        acc = mul_rn(dx, dx);
        const float dy = sub_rn(sh.y, qf[1]);
        acc = add_rn(acc, mul_rn(dy, dy));
        if (K > 2) {
          const float dz = sub_rn(sh.z, qf[2]);
          acc = add_rn(acc, mul_rn(dz, dz));
        }
      }
      if (acc > thr) continue;
      ++rescored;
      const double d = point_dist(t, p, qw);
      const bool ok = v.excl ? d < v.radius : d <= v.radius;
      if (!ok) continue;
      const uint32_t it = __float_as_uint(sh.w);
      if (v.cnt >= v.k && !TopK<double, MAXK>::less(d, it, v.kd, v.kit)) continue;
      v.offer(d, it);
      thr = screen_threshold(v.bound(), eps_k);
    }
  }
};

// ---------------------------------------------------------------------- kernels --------

// Sums the wave's counters and adds them to the global ones with one atomic each.
// Every lane of the wave must reach this call.
__device__ __forceinline__ void flush_counters(const Counters& c, unsigned long long* __restrict__ out) {
  if (!out) return;
  unsigned long long v[4] = {c.leaves, c.points, c.stems, c.rescored};
#pragma unroll
  for (int j = 0; j < 4; ++j)
    for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_xor(v[j], o);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (v[j]) atomicAdd(&out[j], v[j]);
  }
}

template <class A, int K>
__device__ __forceinline__ void load_query(const A* __restrict__ queries, uint64_t qi, A (&q)[K],
                                           typename Ax<A>::O (&qw)[K]) {
  using O = typename Ax<A>::O;
#pragma unroll
  for (int d = 0; d < K; ++d) {
    q[d] = queries[qi * K + d];
    qw[d] = O(q[d]);
  }
}

constexpr int kBlock = 256;

// Kernel prologue shared by every family: one thread per query, queries optionally
// visited in a permuted (Morton) order; results always land at the query's own index.
#define KD_QUERY_PROLOGUE                                             \
  using O = typename Ax<A>::O;                                        \
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;     \
  const bool active = i < m;                                          \
  const uint64_t qi = active ? (order ? order[i] : i) : 0;            \
  A q[K];                                                             \
  O qw[K];                                                            \
  if (active) load_query<A, K>(queries, qi, q, qw);                   \
  Counters cnt;

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
        const A pivot = t.stems[idx];
        idx = (idx << 1) | (pick(q, dim) >= pivot ? 1u : 0u);
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
__global__ __launch_bounds__(kBlock) void k_nearest_n(DevTree<A, K> t, const A* __restrict__ queries,
                                                      const uint32_t* __restrict__ order, uint64_t m,
                                                      int k, typename Ax<A>::O radius, int excl,
                                                      uint32_t* __restrict__ out_items,
                                                      typename Ax<A>::O* __restrict__ out_d,
                                                      uint32_t* __restrict__ out_counts,
                                                      unsigned long long* __restrict__ stats) {
  KD_QUERY_PROLOGUE
  if (active) {
    KnnVis<O, MAXK> v;
    v.init(k, radius, excl != 0);
    traverse(t, q, qw, v, cnt);
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

template <int K, int MAXK>
__global__ __launch_bounds__(kBlock) void k_nearest_n_screen(
    DevTree<double, K> t, const double* __restrict__ queries, const uint32_t* __restrict__ order,
    uint64_t m, int k, double radius, int excl, const double* __restrict__ max_abs,
    uint32_t* __restrict__ out_items, double* __restrict__ out_d, uint32_t* __restrict__ out_counts,
    unsigned long long* __restrict__ stats) {
  using A = double;
  KD_QUERY_PROLOGUE
  if (active) {
    ScreenKnnVis<MAXK> v;
    v.v.init(k, radius, excl != 0);
    double amax = 0.0;
#pragma unroll
    for (int d = 0; d < 3; ++d) v.qf[d] = 0.0f;
#pragma unroll
    for (int d = 0; d < K; ++d) {
      v.qf[d] = (float)q[d];
      amax = fmax(amax, max_abs[d] + fabs(q[d]));
    }
    const double u = 0x1p-24;
    const double eps = u * (2.0 + u) * amax + 3.0 * 0x1p-150;
    v.eps_k = sqrt((double)K) * eps * (1.0 + 1e-6);
    v.thr = screen_threshold(v.v.bound(), v.eps_k);
    v.rescored = 0;
    traverse(t, q, qw, v, cnt);
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

// Sorts one query's segment by (distance, item): insertion sort for short segments,
// heapsort otherwise (in place, global memory).
template <class O>
__device__ void sort_segment(uint32_t* __restrict__ it, O* __restrict__ d, uint64_t len) {
  auto less = [](O ad, uint32_t ai, O bd, uint32_t bi) { return ad < bd || (ad == bd && ai < bi); };
  if (len <= 48) {
    for (uint64_t a = 1; a < len; ++a) {
      const O kd = d[a];
      const uint32_t ki = it[a];
      uint64_t b = a;
      while (b > 0 && less(kd, ki, d[b - 1], it[b - 1])) {
        d[b] = d[b - 1];
        it[b] = it[b - 1];
        --b;
      }
      d[b] = kd;
      it[b] = ki;
    }
    return;
  }
  auto sift = [&](uint64_t root, uint64_t end) {
    for (;;) {
      uint64_t c = 2 * root + 1;
      if (c >= end) return;
      if (c + 1 < end && less(d[c], it[c], d[c + 1], it[c + 1])) ++c;
      if (!less(d[root], it[root], d[c], it[c])) return;
      const O td = d[root];
      d[root] = d[c];
      d[c] = td;
      const uint32_t ti = it[root];
      it[root] = it[c];
      it[c] = ti;
      root = c;
    }
  };
  for (uint64_t s = len / 2; s-- > 0;) sift(s, len);
  for (uint64_t e = len - 1; e > 0; --e) {
    const O td = d[0];
    d[0] = d[e];
    d[e] = td;
    const uint32_t ti = it[0];
    it[0] = it[e];
    it[e] = ti;
    sift(0, e);
  }
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

// Morton keys: each coordinate scaled into [0, 2^bits) over the tree's bounding box and
// bit-interleaved, so that consecutive queries share traversal paths and leaves.
template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_morton(const A* __restrict__ queries, uint64_t m,
                                                   const double* __restrict__ lo,
                                                   const double* __restrict__ scale,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ ids) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= m) return;
  constexpr int bits = 32 / K;
  const uint32_t cap = (1u << bits) - 1u;
  uint32_t key = 0;
  uint32_t c[K];
#pragma unroll
  for (int d = 0; d < K; ++d) {
    double x = (double(queries[i * K + d]) - lo[d]) * scale[d];
    x = x > 0.0 ? x : 0.0;  // also maps NaN to 0
    c[d] = x >= double(cap) ? cap : uint32_t(x);
  }
#pragma unroll
  for (int b = bits - 1; b >= 0; --b)
#pragma unroll
    for (int d = 0; d < K; ++d) key = (key << 1) | ((c[d] >> b) & 1u);
  keys[i] = key;
  ids[i] = uint32_t(i);
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// Clears n counters (a kernel rather than hipMemsetAsync so that callers may capture it).
__global__ void k_zero_u64(unsigned long long* p, int n) {
  if (int(threadIdx.x) < n) p[threadIdx.x] = 0ull;
}

}  // namespace dev
}  // namespace kd

// ======================================================================= handles ======

struct kd_layout {
  kd::HostLayout L;
};

struct kd_tree {
  int dtype = 0, dims = 0, bucket = 0, max_stem_level = -1;
  uint64_t size = 0, leaf_count = 0, max_leaf_len = 0;
  int device = 0;
  void* stems = nullptr;
  uint64_t stems_len = 0;
  uint32_t* leaf_start = nullptr;
  void* coords = nullptr;
  uint64_t plane = 0;
  uint32_t* items = nullptr;
  float4* shadow = nullptr;
  double* d_max_abs = nullptr;   // per-dimension max |coordinate| (FP32 screen bound)
  double* d_bbox = nullptr;      // lo[4], scale[4] for Morton keys
  unsigned long long* counters = nullptr;
  double bbox_min[4] = {0, 0, 0, 0}, bbox_max[4] = {0, 0, 0, 0};
  uint64_t device_bytes = 0;
  // optional HIP-event timing of the main query kernel of each call
  bool timing = false;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool ev_pending = false;
};

namespace kd {
namespace {

template <class A, int K> dev::DevTree<A, K> dev_tree(const kd_tree* t) {
  dev::DevTree<A, K> d;
  d.stems = static_cast<const A*>(t->stems);
  d.leaf_start = t->leaf_start;
  d.coords = static_cast<const A*>(t->coords);
  d.items = t->items;
  d.shadow = t->shadow;
  d.plane = t->plane;
  d.n = uint32_t(t->size);
  d.nlev = t->max_stem_level + 1;
  return d;
}

void free_tree(kd_tree* t) {
  if (!t) return;
  if (t->ev_start) (void)hipEventDestroy(t->ev_start);
  if (t->ev_stop) (void)hipEventDestroy(t->ev_stop);
  (void)hipFree(t->stems);
  (void)hipFree(t->leaf_start);
  (void)hipFree(t->coords);
  (void)hipFree(t->items);
  (void)hipFree(t->shadow);
  (void)hipFree(t->d_max_abs);
  (void)hipFree(t->d_bbox);
  (void)hipFree(t->counters);
  delete t;
}

template <class T> int dmalloc(T** p, size_t bytes, uint64_t& total) {
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  }
  total += bytes;
  return KD_OK;
}

#define KD_TRY(x)           \
  do {                      \
    int rc_ = (x);          \
    if (rc_ != KD_OK) {     \
      free_tree(t);         \
      return rc_;           \
    }                       \
  } while (0)

template <class F> void host_parallel(uint64_t count, F&& f) {
  int threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (count < 65536) threads = 1;
  std::vector<std::thread> th;
  uint64_t chunk = (count + threads - 1) / threads;
  for (int i = 0; i < threads; ++i) {
    uint64_t s = i * chunk, e = std::min(count, s + chunk);
    if (s >= e) break;
    th.emplace_back([&f, s, e] { f(s, e); });
  }
  for (auto& x : th) x.join();
}

// Decodes Kiddo's leaf arena into SoA + items, validates it, and uploads everything.
template <class A, int K> int create_typed(const kd_tree_desc* desc, uint32_t flags, kd_tree** out) {
  const uint64_t n = desc->size, L = desc->leaf_count;
  const int msl = desc->max_stem_level;
  if (msl >= kMaxStemLevels) return set_error(KD_ERR_TOO_DEEP, "max_stem_level %d too deep", msl);
  if (msl < -1) return set_error(KD_ERR_INVALID, "max_stem_level %d invalid", msl);
  const uint64_t want_leaves = msl >= 0 ? (uint64_t(1) << (msl + 1)) : (n > 0 ? 1 : 0);
  if (L != want_leaves && !(msl < 0 && L <= 1))
    return set_error(KD_ERR_INVALID, "leaf_count %llu does not match max_stem_level %d",
                     (unsigned long long)L, msl);
  if (msl >= 0 && desc->stems_len < (uint64_t(1) << (msl + 1)))
    return set_error(KD_ERR_INVALID, "stems_len %llu too short", (unsigned long long)desc->stems_len);
  if (n > 0xFFFFFFFFull) return set_error(KD_ERR_INVALID, "trees above 2^32-1 entries are not supported");
  if ((flags & KD_TREE_SHADOW_F32) && (desc->dtype != KD_F64 || K > 3))
    return set_error(KD_ERR_UNSUPPORTED, "the FP32 shadow is provided for f64 trees with dims <= 3");
  const size_t entry_bytes = K * sizeof(A) + 4;
  std::vector<uint32_t> leaf_start(L + 1, 0);
  uint64_t total = 0, max_len = 0;
  for (uint64_t l = 0; l < L; ++l) {
    const uint64_t off = desc->leaf_extents[2 * l], len = desc->leaf_extents[2 * l + 1];
    if (len > uint64_t(kMaxLeafLen))
      return set_error(KD_ERR_LEAF_OVERFLOW, "leaf %llu holds %llu entries (> 1024)",
                       (unsigned long long)l, (unsigned long long)len);
    if (off + len * entry_bytes > desc->leaf_bytes_len)
      return set_error(KD_ERR_INVALID, "leaf %llu extends past the arena", (unsigned long long)l);
    leaf_start[l] = uint32_t(total);
    total += len;
    max_len = std::max(max_len, len);
  }
  leaf_start[L] = uint32_t(total);
  if (total != n)
    return set_error(KD_ERR_INVALID, "leaf lengths sum to %llu, size is %llu",
                     (unsigned long long)total, (unsigned long long)n);

  const uint64_t plane = (n + 63) / 64 * 64;
  std::vector<A> coords(plane * K);
  std::vector<uint32_t> items(n);
  std::vector<float4> shadow((flags & KD_TREE_SHADOW_F32) ? n : 0);
  // per-thread partial bounding boxes
  std::mutex mu;
  double lo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, hi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
  double mabs[4] = {0, 0, 0, 0};
  bool bad = false;
  host_parallel(L, [&](uint64_t s, uint64_t e) {
    double tlo[4] = {INFINITY, INFINITY, INFINITY, INFINITY}, thi[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    double tabs[4] = {0, 0, 0, 0};
    bool tbad = false;
    for (uint64_t l = s; l < e; ++l) {
      const uint8_t* p = desc->leaf_bytes + desc->leaf_extents[2 * l];
      const uint64_t len = desc->leaf_extents[2 * l + 1];
      uint64_t dst = leaf_start[l], base = 0;
      while (base < len) {
        const uint64_t rem = len - base;
        const int tl = rem >= 32 ? 32 : rem >= 8 ? 8 : rem >= 4 ? 4 : rem >= 2 ? 2 : 1;
        for (int d = 0; d < K; ++d)
          for (int i = 0; i < tl; ++i) {
            A v;
            std::memcpy(&v, p + (size_t(d) * tl + i) * sizeof(A), sizeof(A));
            coords[d * plane + dst + i] = v;
            const double x = double(v);
            if (std::isnan(x) || v == AxisHost<A>::max_value()) tbad = true;
            tlo[d] = std::min(tlo[d], x);
            thi[d] = std::max(thi[d], x);
            tabs[d] = std::max(tabs[d], std::fabs(x));
          }
        for (int i = 0; i < tl; ++i) std::memcpy(&items[dst + i], p + size_t(K) * tl * sizeof(A) + 4 * i, 4);
        p += size_t(K) * tl * sizeof(A) + 4 * size_t(tl);
        dst += tl;
        base += tl;
      }
    }
    std::lock_guard<std::mutex> g(mu);
    for (int d = 0; d < 4; ++d) {
      lo[d] = std::min(lo[d], tlo[d]);
      hi[d] = std::max(hi[d], thi[d]);
      mabs[d] = std::max(mabs[d], tabs[d]);
    }
    bad |= tbad;
  });
  if (bad) return set_error(KD_ERR_INVALID, "tree holds a NaN coordinate or the reserved maximum value");
  if (!shadow.empty()) {
    for (int d = 0; d < K; ++d)
      if (!(mabs[d] <= 0x1p60))
        return set_error(KD_ERR_UNSUPPORTED, "FP32 shadow needs |coordinates| <= 2^60");
    host_parallel(n, [&](uint64_t s, uint64_t e) {
      for (uint64_t i = s; i < e; ++i) {
        float4 f;
        f.x = float(coords[0 * plane + i]);
        f.y = float(coords[1 * plane + i]);
        f.z = K > 2 ? float(coords[2 * plane + i]) : 0.0f;
        std::memcpy(&f.w, &items[i], 4);
        shadow[i] = f;
      }
    });
  }

  kd_tree* t = new kd_tree;
  t->dtype = desc->dtype;
  t->dims = K;
  t->bucket = desc->bucket;
  t->max_stem_level = msl;
  t->size = n;
  t->leaf_count = L;
  t->max_leaf_len = max_len;
  t->plane = plane;
  t->stems_len = desc->stems_len;
  (void)hipGetDevice(&t->device);
  for (int d = 0; d < K; ++d) {
    t->bbox_min[d] = n ? lo[d] : 0.0;
    t->bbox_max[d] = n ? hi[d] : 0.0;
  }
  uint64_t& tot = t->device_bytes;
  KD_TRY(dmalloc(&t->stems, std::max<uint64_t>(desc->stems_len, 1) * sizeof(A), tot));
  KD_TRY(dmalloc(&t->leaf_start, (L + 1) * sizeof(uint32_t), tot));
  KD_TRY(dmalloc(&t->coords, coords.size() * sizeof(A), tot));
  KD_TRY(dmalloc(&t->items, std::max<uint64_t>(n, 1) * sizeof(uint32_t), tot));
  KD_TRY(dmalloc(&t->d_max_abs, 4 * sizeof(double), tot));
  KD_TRY(dmalloc(&t->d_bbox, 8 * sizeof(double), tot));
  KD_TRY(dmalloc(&t->counters, 4 * sizeof(unsigned long long), tot));
  if (!shadow.empty()) KD_TRY(dmalloc(&t->shadow, shadow.size() * sizeof(float4), tot));
  double bb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < K; ++d) {
    bb[d] = t->bbox_min[d];
    const double ext = t->bbox_max[d] - t->bbox_min[d];
    bb[4 + d] = ext > 0 ? double((1u << (32 / K)) - 1u) / ext : 0.0;
  }
  auto up = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return KD_OK;
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return set_error(KD_ERR_DEVICE, "upload failed: %s", hipGetErrorString(e));
    return KD_OK;
  };
  if (desc->stems_len) KD_TRY(up(t->stems, desc->stems, desc->stems_len * sizeof(A)));
  KD_TRY(up(t->leaf_start, leaf_start.data(), leaf_start.size() * sizeof(uint32_t)));
  KD_TRY(up(t->coords, coords.data(), coords.size() * sizeof(A)));
  KD_TRY(up(t->items, items.data(), items.size() * sizeof(uint32_t)));
  KD_TRY(up(t->d_max_abs, mabs, 4 * sizeof(double)));
  KD_TRY(up(t->d_bbox, bb, 8 * sizeof(double)));
  if (!shadow.empty()) KD_TRY(up(t->shadow, shadow.data(), shadow.size() * sizeof(float4)));
  if (hipMemset(t->counters, 0, 4 * sizeof(unsigned long long)) != hipSuccess) {
    free_tree(t);
    return set_error(KD_ERR_DEVICE, "hipMemset failed");
  }
  *out = t;
  return KD_OK;
}

uint32_t grid_for(uint64_t m) { return uint32_t((m + dev::kBlock - 1) / dev::kBlock); }

void reset_counters(kd_tree* t, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_zero_u64, dim3(1), dim3(64), 0, s, t->counters, 4);
}

// Brackets the main query kernel of a call with HIP events on its stream when the tree
// has timing enabled (kd_tree_set_timing); kd_last_kernel_ms reads the interval.
struct KernelTimer {
  kd_tree* t;
  hipStream_t s;
  KernelTimer(kd_tree* t_, hipStream_t s_) : t(t_), s(s_) {
    if (t->timing) (void)hipEventRecord(t->ev_start, s);
  }
  ~KernelTimer() {
    if (t->timing) {
      (void)hipEventRecord(t->ev_stop, s);
      t->ev_pending = true;
    }
  }
};

// Optional Morton ordering of a batch: returns a device permutation (or null).
struct Order {
  uint32_t* perm = nullptr;
  void* mem = nullptr;
  hipStream_t s = nullptr;
  ~Order() {
    if (mem) (void)hipFreeAsync(mem, s);
  }
};

template <class A, int K>
int make_order(const kd_tree* t, const A* queries, uint64_t m, uint32_t flags, hipStream_t s, Order& o) {
  if (!(flags & KD_MORTON) || m < 2) return KD_OK;
  if (m > 0xFFFFFFFFull) return set_error(KD_ERR_INVALID, "Morton ordering supports < 2^32 queries");
  o.s = s;
  size_t temp = 0;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr;
  KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, k0, k1, v0, v1, int(m), 0, 32, s));
  const size_t arr = ((m * 4 + 255) / 256) * 256;
  const size_t total = 4 * arr + temp;
  KD_HIP(hipMallocAsync(&o.mem, total, s));
  uint8_t* base = static_cast<uint8_t*>(o.mem);
  k0 = reinterpret_cast<uint32_t*>(base);
  k1 = reinterpret_cast<uint32_t*>(base + arr);
  v0 = reinterpret_cast<uint32_t*>(base + 2 * arr);
  v1 = reinterpret_cast<uint32_t*>(base + 3 * arr);
  void* tmp = base + 4 * arr;
  hipLaunchKernelGGL((dev::k_morton<A, K>), dim3(grid_for(m)), dim3(dev::kBlock), 0, s, queries, m,
                     t->d_bbox, t->d_bbox + 4, k0, v0);
  KD_HIP(hipGetLastError());
  KD_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, k0, k1, v0, v1, int(m), 0, 32, s));
  o.perm = v1;
  return KD_OK;
}

int check_tree(const kd_tree* t) {
  if (!t) return set_error(KD_ERR_INVALID, "tree is null");
  return KD_OK;
}

template <class A, int K, int MAXK>
int launch_knn(kd_tree* t, const A* q, uint64_t m, int k, double radius, uint32_t flags,
               uint32_t* items, void* dists, uint32_t* counts, const uint32_t* perm, hipStream_t s) {
  using O = typename dev::Ax<A>::O;
  const int excl = (flags & KD_EXCLUSIVE) ? 1 : 0;
  if constexpr (std::is_same<A, double>::value && K <= 3) {
    if (flags & KD_SCREEN_F32) {
      if (!t->shadow) return set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 needs a tree created with KD_TREE_SHADOW_F32");
      reset_counters(t, s);
      KernelTimer kt(t, s);
      hipLaunchKernelGGL((dev::k_nearest_n_screen<K, MAXK>), dim3(grid_for(m)), dim3(dev::kBlock), 0, s,
                         dev_tree<double, K>(t), q, perm, m, k, radius, excl, t->d_max_abs, items,
                         static_cast<double*>(dists), counts, t->counters);
      KD_HIP(hipGetLastError());
      return KD_OK;
    }
  }
  if (flags & KD_SCREEN_F32) return set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 applies to f64 trees with dims <= 3");
  reset_counters(t, s);
  KernelTimer kt(t, s);
  hipLaunchKernelGGL((dev::k_nearest_n<A, K, MAXK>), dim3(grid_for(m)), dim3(dev::kBlock), 0, s,
                     dev_tree<A, K>(t), q, perm, m, k, O(radius), excl, items, static_cast<O*>(dists), counts,
                     t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

template <class A, int K, int MAXK>
int launch_best(kd_tree* t, const A* q, uint64_t m, int n, double radius, uint32_t flags,
                uint32_t* items, void* dists, uint32_t* counts, const uint32_t* perm, hipStream_t s) {
  using O = typename dev::Ax<A>::O;
  reset_counters(t, s);
  KernelTimer kt(t, s);
  hipLaunchKernelGGL((dev::k_best_n_within<A, K, MAXK>), dim3(grid_for(m)), dim3(dev::kBlock), 0, s,
                     dev_tree<A, K>(t), q, perm, m, n, O(radius), (flags & KD_EXCLUSIVE) ? 1 : 0, items,
                     static_cast<O*>(dists), counts, t->counters);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

constexpr int kMaxK = 64;

}  // namespace
}  // namespace kd

// ====================================================================== C ABI ========

extern "C" {

const char* kd_last_error(void) { return kd::g_last_error.c_str(); }
int kd_version(void) { return 1; }

int kd_build_layout(int32_t dtype, int32_t dims, int32_t bucket, uint64_t n, const void* points,
                    const uint32_t* items, int32_t threads, kd_layout** out) {
  if (!out) return kd::set_error(KD_ERR_INVALID, "out is null");
  try {
    auto* l = new kd_layout;
    int rc = kd::build_layout(l->L, dtype, dims, bucket, n, points, items, threads);
    if (rc != KD_OK) {
      delete l;
      return rc;
    }
    *out = l;
    return KD_OK;
  } catch (const std::bad_alloc&) {
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed while building");
  } catch (...) {
    return kd::set_error(KD_ERR_INVALID, "unexpected failure while building");
  }
}

int kd_layout_desc(const kd_layout* layout, kd_tree_desc* out) {
  if (!layout || !out) return kd::set_error(KD_ERR_INVALID, "null argument");
  const kd::HostLayout& L = layout->L;
  out->dtype = L.dtype;
  out->dims = L.dims;
  out->bucket = L.bucket;
  out->max_stem_level = L.max_stem_level;
  out->size = L.size;
  out->stems_len = L.stems_len;
  out->stems = L.stems.data();
  out->leaf_count = L.extents.size() / 2;
  out->leaf_extents = L.extents.data();
  out->leaf_bytes_len = L.arena.size();
  out->leaf_bytes = L.arena.data();
  return KD_OK;
}

void kd_layout_free(kd_layout* layout) { delete layout; }

int kd_tree_create(const kd_tree_desc* desc, uint32_t tree_flags, kd_tree** out) {
  if (!desc || !out) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (desc->leaf_count && (!desc->leaf_extents || !desc->leaf_bytes))
    return kd::set_error(KD_ERR_INVALID, "leaf buffers are null");
  if (desc->stems_len && !desc->stems) return kd::set_error(KD_ERR_INVALID, "stems is null");
  try {
    KD_DISPATCH_AK(desc->dtype, desc->dims, return (kd::create_typed<A, KK>(desc, tree_flags, out)));
  } catch (const std::bad_alloc&) {
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed while uploading");
  }
  return KD_ERR_INVALID;
}

int kd_tree_destroy(kd_tree* tree) {
  if (!tree) return KD_OK;
  (void)hipDeviceSynchronize();
  kd::free_tree(tree);
  return KD_OK;
}

int kd_tree_get_info(const kd_tree* t, kd_tree_info* out) {
  if (!t || !out) return kd::set_error(KD_ERR_INVALID, "null argument");
  out->dtype = t->dtype;
  out->dims = t->dims;
  out->max_stem_level = t->max_stem_level;
  out->has_shadow_f32 = t->shadow != nullptr;
  out->size = t->size;
  out->leaf_count = t->leaf_count;
  out->max_leaf_len = t->max_leaf_len;
  out->device_bytes = t->device_bytes;
  for (int d = 0; d < 4; ++d) {
    out->bbox_min[d] = t->bbox_min[d];
    out->bbox_max[d] = t->bbox_max[d];
  }
  return KD_OK;
}

int kd_nearest_one(kd_tree* t, const void* queries, uint64_t m, uint32_t flags, uint32_t* items,
                   void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists) return kd::set_error(KD_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  KD_DISPATCH_AK(t->dtype, t->dims, {
    using O = typename kd::dev::Ax<A>::O;
    const A* q = static_cast<const A*>(queries);
    kd::Order o;
    if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
    kd::reset_counters(t, s);
    kd::KernelTimer kt(t, s);
    hipLaunchKernelGGL((kd::dev::k_nearest_one<A, KK>), dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0, s,
                       kd::dev_tree<A, KK>(t), q, o.perm, m, items, static_cast<O*>(dists), t->counters);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

int kd_approx_nearest_one(kd_tree* t, const void* queries, uint64_t m, uint32_t flags, uint32_t* items,
                          void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (t->dtype != KD_F32 && t->dtype != KD_F64)
    return kd::set_error(KD_ERR_UNSUPPORTED, "approx nearest_one needs a float tree (Output == A)");
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists) return kd::set_error(KD_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  KD_DISPATCH_AK(t->dtype, t->dims, {
    if constexpr (std::is_floating_point<A>::value) {
      const A* q = static_cast<const A*>(queries);
      kd::Order o;
      if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
      kd::reset_counters(t, s);
      kd::KernelTimer kt(t, s);
      hipLaunchKernelGGL((kd::dev::k_approx_nearest_one<A, KK>), dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0,
                         s, kd::dev_tree<A, KK>(t), q, o.perm, m, items, static_cast<A*>(dists), t->counters);
      KD_HIP(hipGetLastError());
    }
  });
  return KD_OK;
}

int kd_nearest_n(kd_tree* t, const void* queries, uint64_t m, uint32_t k, double radius, uint32_t flags,
                 uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1 (NonZeroUsize)");
  if (k > uint32_t(kd::kMaxK)) return kd::set_error(KD_ERR_UNSUPPORTED, "k above %d is not provided", kd::kMaxK);
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  KD_DISPATCH_AK(t->dtype, t->dims, {
    const A* q = static_cast<const A*>(queries);
    kd::Order o;
    if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
    int rc;
    if (k <= 1) rc = kd::launch_knn<A, KK, 1>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    else if (k <= 4) rc = kd::launch_knn<A, KK, 4>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    else if (k <= 8) rc = kd::launch_knn<A, KK, 8>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    else if (k <= 16) rc = kd::launch_knn<A, KK, 16>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    else if (k <= 32) rc = kd::launch_knn<A, KK, 32>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    else rc = kd::launch_knn<A, KK, 64>(t, q, m, k, radius, flags, items, dists, counts, o.perm, s);
    if (rc) return rc;
  });
  return KD_OK;
}

int kd_within_count(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t flags,
                    uint64_t* offsets, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!offsets) return kd::set_error(KD_ERR_INVALID, "offsets is null");
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (m == 0) {
    KD_HIP(hipMemsetAsync(offsets, 0, sizeof(uint64_t), s));
    return KD_OK;
  }
  if (!queries) return kd::set_error(KD_ERR_INVALID, "queries is null");
  KD_DISPATCH_AK(t->dtype, t->dims, {
    using O = typename kd::dev::Ax<A>::O;
    const A* q = static_cast<const A*>(queries);
    kd::Order o;
    if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
    // counts into a scratch array, then exclusive-scan m+1 entries into offsets
    void* mem = nullptr;
    size_t temp = 0;
    uint64_t* cnt = nullptr;
    KD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, cnt, offsets, int(m + 1), s));
    const size_t arr = ((m + 1) * 8 + 255) / 256 * 256;
    KD_HIP(hipMallocAsync(&mem, arr + temp, s));
    cnt = static_cast<uint64_t*>(mem);
    kd::reset_counters(t, s);
    kd::KernelTimer kt(t, s);
    hipLaunchKernelGGL((kd::dev::k_within_count<A, KK>), dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0, s,
                       kd::dev_tree<A, KK>(t), q, o.perm, m, O(radius), (flags & KD_EXCLUSIVE) ? 1 : 0, cnt,
                       t->counters);
    hipLaunchKernelGGL(kd::dev::k_set_u64, dim3(1), dim3(1), 0, s, cnt + m, uint64_t(0));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(static_cast<uint8_t*>(mem) + arr, temp, cnt, offsets,
                                                    int(m + 1), s);
    (void)hipFreeAsync(mem, s);
    if (e != hipSuccess) return kd::set_error(KD_ERR_DEVICE, "scan failed: %s", hipGetErrorString(e));
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

int kd_within_fill(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t flags,
                   const uint64_t* offsets, uint32_t* items, void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !offsets) return kd::set_error(KD_ERR_INVALID, "null buffer");
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  hipStream_t s = static_cast<hipStream_t>(stream);
  KD_DISPATCH_AK(t->dtype, t->dims, {
    using O = typename kd::dev::Ax<A>::O;
    const A* q = static_cast<const A*>(queries);
    kd::Order o;
    if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
    kd::KernelTimer kt(t, s);
    hipLaunchKernelGGL((kd::dev::k_within_fill<A, KK>), dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0, s,
                       kd::dev_tree<A, KK>(t), q, o.perm, m, O(radius), (flags & KD_EXCLUSIVE) ? 1 : 0,
                       (flags & KD_UNSORTED) ? 0 : 1, offsets, items, static_cast<O*>(dists), nullptr);
    KD_HIP(hipGetLastError());
  });
  return KD_OK;
}

int kd_best_n_within(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t n, uint32_t flags,
                     uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (n == 0) return kd::set_error(KD_ERR_INVALID, "n must be >= 1 (NonZero)");
  if (n > uint32_t(kd::kMaxK)) return kd::set_error(KD_ERR_UNSUPPORTED, "n above %d is not provided", kd::kMaxK);
  if (std::isnan(radius)) return kd::set_error(KD_ERR_INVALID, "radius is NaN");
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  KD_DISPATCH_AK(t->dtype, t->dims, {
    const A* q = static_cast<const A*>(queries);
    kd::Order o;
    if (int rc = kd::make_order<A, KK>(t, q, m, flags, s, o)) return rc;
    int rc;
    if (n <= 4) rc = kd::launch_best<A, KK, 4>(t, q, m, n, radius, flags, items, dists, counts, o.perm, s);
    else if (n <= 8) rc = kd::launch_best<A, KK, 8>(t, q, m, n, radius, flags, items, dists, counts, o.perm, s);
    else if (n <= 16) rc = kd::launch_best<A, KK, 16>(t, q, m, n, radius, flags, items, dists, counts, o.perm, s);
    else rc = kd::launch_best<A, KK, 64>(t, q, m, n, radius, flags, items, dists, counts, o.perm, s);
    if (rc) return rc;
  });
  return KD_OK;
}

int kd_tree_set_timing(kd_tree* t, int32_t enable) {
  if (int rc = kd::check_tree(t)) return rc;
  if (enable && !t->ev_start) {
    KD_HIP(hipEventCreate(&t->ev_start));
    KD_HIP(hipEventCreate(&t->ev_stop));
  }
  t->timing = enable != 0;
  t->ev_pending = false;
  return KD_OK;
}

int kd_last_kernel_ms(kd_tree* t, float* ms) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!ms) return kd::set_error(KD_ERR_INVALID, "ms is null");
  if (!t->timing || !t->ev_pending) return kd::set_error(KD_ERR_INVALID, "no timed call recorded");
  KD_HIP(hipEventSynchronize(t->ev_stop));
  KD_HIP(hipEventElapsedTime(ms, t->ev_start, t->ev_stop));
  return KD_OK;
}

int kd_query_stats(kd_tree* t, uint64_t* out4) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!out4) return kd::set_error(KD_ERR_INVALID, "out is null");
  unsigned long long h[4] = {0, 0, 0, 0};
  KD_HIP(hipDeviceSynchronize());
  KD_HIP(hipMemcpy(h, t->counters, sizeof(h), hipMemcpyDeviceToHost));
  for (int j = 0; j < 4; ++j) out4[j] = h[j];
  return KD_OK;
}

}  // extern "C"
