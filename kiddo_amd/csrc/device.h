// device.h — device-side building blocks shared by the query kernels: numeric traits,
// the tree view, Kiddo-order traversal (thread-per-query and group-per-query forms) and
// the result collectors.
//
// The traversal takes the same decisions as Kiddo's arithmetic DFS
// (kiddo src/kd_tree/orchestrator/mod.rs:388-603): near child first, far child queued when
// its incremental rect distance rd' = (rd - old_off^2) + new_off^2
// (src/dist/distance_metric_core.rs:89-98) passes the current bound, and queued far children
// re-checked against the bound of the moment when they are popped (:472-483). Kiddo keeps a
// restore stack and a far stack; here the stack is one slot per stem level plus two bit
// masks, because the Eytzinger path of the last leaf gives every ancestor index
// (eytzinger/mod.rs:131-168) and an off[] entry only changes when a far child is entered,
// so only those levels need a saved value.
//
// Distances are SquaredEuclidean in Kiddo's scalar order: ((0 + d0*d0) + d1*d1) + ...,
// each product and sum rounded separately (squared_euclidean/mod.rs:36-39,
// distance_metric_core.rs:38-40); the explicit *_rn intrinsics keep the compiler from
// fusing them, and the build also passes -ffp-contract=off.
//
// Results use the canonical order (distance, then lower item): the collectors accept an
// equal-distance candidate with a lower item, and an equal-distance subtree is only pruned
// by an exclusive radius, so the answer does not depend on within-leaf order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

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

// |a - b| in O: Axis::saturating_dist for floats, (a - b).abs() (axis.rs:124-126); for
// integer axes (larger minus smaller, axis.rs:191-197) the widened difference is exact.
template <class O> __device__ __forceinline__ O abs_diff(O a, O b) { return fabs(sub_rn(a, b)); }

// A K-vector (K <= 4) indexed by a run-time dimension. It is held as one vector value, so
// an element access is an extract/insert on registers that the backend expands into
// compare-and-select; named scalars or an array let the optimiser fold the selects back
// into an address computation on a stack slot, which spills the vector to scratch memory.
template <class T, int K> struct Small {
  static_assert(K >= 2 && K <= 4, "K must be 2..4");
  typedef T vec __attribute__((ext_vector_type(K)));
  vec v;
  __device__ __forceinline__ T operator[](int d) const { return v[d]; }
  __device__ __forceinline__ void set(int d, T x) { v[d] = x; }
  __device__ __forceinline__ void fill(T x) { v = vec(x); }
};

constexpr int kMaxLevels = 32;
constexpr int kBlock = 256;  // threads per block of every query kernel
constexpr int kStatSlots = 256;  // query-statistics counters: groups of 4, summed on read
// The counter array holds the statistics groups, then one word that counts the queries the
// screened k-NN kernel hands to its exact redo pass (k_nearest_n_redo), then padding.
constexpr int kRedoWord = 4 * kStatSlots;
constexpr int kCounterWords = 4 * kStatSlots + 4;
// The FP32 shadow array holds this many (zero) entries past the last point, so that batched
// loads may run past the end of a leaf without clamping their indices.
constexpr uint32_t kShadowPad = 8;

// ------------------------------------------------------------------- metrics ---------
// Distance metrics (kiddo src/dist/*/mod.rs). Each gives dist1 of a coordinate difference
// (or of an axis offset, i.e. dist1(x, 0)), how per-axis terms combine from 0 in axis
// order (distance_metric_core.rs:38-40 and overrides), and the rect distance after one
// axis offset changes (distance_metric_core.rs:89-98 and overrides). Every product and sum
// is rounded separately, as in Kiddo's scalar path.
struct SqEuclidean {  // squared_euclidean/mod.rs:36-39
  template <class O> __device__ __forceinline__ O d1(O x) const { return mul_rn(x, x); }
  template <class O> __device__ __forceinline__ O combine(O acc, O c) const { return add_rn(acc, c); }
  template <class O, int K>
  __device__ __forceinline__ O rect(O rd, const Small<O, K>& off, int dim, O no) const {
    const O oo = off[dim];
    return add_rn(sub_rn(rd, mul_rn(oo, oo)), mul_rn(no, no));
  }
};
struct Manhattan {  // manhattan/mod.rs:36-49: |d|, sum, (rd - old) + new
  template <class O> __device__ __forceinline__ O d1(O x) const { return fabs(x); }
  template <class O> __device__ __forceinline__ O combine(O acc, O c) const { return add_rn(acc, c); }
  template <class O, int K>
  __device__ __forceinline__ O rect(O rd, const Small<O, K>& off, int dim, O no) const {
    return add_rn(sub_rn(rd, off[dim]), no);
  }
};
struct Chebyshev {  // chebyshev/mod.rs:34-61: |d|, max, offsets rescanned
  template <class O> __device__ __forceinline__ O d1(O x) const { return fabs(x); }
  template <class O> __device__ __forceinline__ O combine(O acc, O c) const { return c > acc ? c : acc; }
  template <class O, int K>
  __device__ __forceinline__ O rect(O, const Small<O, K>& off, int dim, O no) const {
    O acc = O(0);
#pragma unroll
    for (int a = 0; a < K; ++a) acc = combine(acc, a == dim ? no : off[a]);
    return acc;
  }
};
struct Minkowski {  // minkowski/mod.rs:55-59: |d|^P, sum, (rd - old^P) + new^P
  int p;
  // powi by binary exponentiation (the expansion of llvm.powi, as compiler-rt's __powidf2)
  template <class O> __device__ __forceinline__ O d1(O x) const {
    O a = fabs(x), r = O(1);
    for (int b = p;;) {
      if (b & 1) r = mul_rn(r, a);
      b >>= 1;
      if (!b) break;
      a = mul_rn(a, a);
    }
    return r;
  }
  template <class O> __device__ __forceinline__ O combine(O acc, O c) const { return add_rn(acc, c); }
  template <class O, int K>
  __device__ __forceinline__ O rect(O rd, const Small<O, K>& off, int dim, O no) const {
    return add_rn(sub_rn(rd, d1(off[dim])), d1(no));
  }
};

// A widened query that carries a metric other than SquaredEuclidean (a plain Small<O, K>
// means SquaredEuclidean).
template <class O, int K, class M> struct MetricQW {
  Small<O, K> q;
  M met;
  __device__ __forceinline__ O operator[](int d) const { return q[d]; }
};
template <class QW> __device__ __forceinline__ SqEuclidean metric_of(const QW&) { return SqEuclidean{}; }
template <class O, int K, class M> __device__ __forceinline__ M metric_of(const MetricQW<O, K, M>& w) {
  return w.met;
}
template <class O, int K> __device__ __forceinline__ Small<O, K> wrap_qw(const Small<O, K>& qw, SqEuclidean) {
  return qw;
}
template <class O, int K, class M> __device__ __forceinline__ MetricQW<O, K, M> wrap_qw(const Small<O, K>& qw, M met) {
  return MetricQW<O, K, M>{qw, met};
}

template <class A, int K> struct DevTree {
  const A* __restrict__ stems;
  const uint32_t* __restrict__ leaf_start;
  const A* __restrict__ coords;  // K planes of `plane` entries
  const uint32_t* __restrict__ items;
  const float4* __restrict__ shadow;  // n + kShadowPad entries
  const A* lds;     // the first lds_n stems, copied into LDS by the kernel prologue
  uint32_t lds_n;   // 0: no LDS copy
  uint64_t plane;
  uint32_t n;
  int nlev;  // number of stem levels = max_stem_level + 1
};

template <class A, int K>
__device__ __forceinline__ A stem_at(const DevTree<A, K>& t, uint32_t idx) {
  return idx < t.lds_n ? t.lds[idx] : t.stems[idx];
}

// Copies the top stem levels into LDS (dynamic shared memory of lds_n entries). Every
// thread of the block must call it.
template <class A, int K> __device__ __forceinline__ void pin_stems(DevTree<A, K>& t) {
  extern __shared__ __align__(16) unsigned char kd_lds_raw[];
  A* lds = reinterpret_cast<A*>(kd_lds_raw);
  for (uint32_t j = threadIdx.x; j < t.lds_n; j += blockDim.x) lds[j] = t.stems[j];
  __syncthreads();
  t.lds = lds;
}

// Squared distance of entry p in Kiddo's operation order.
template <class A, int K, class QW>
__device__ __forceinline__ typename Ax<A>::O point_dist(const DevTree<A, K>& t, uint32_t p, const QW& qw) {
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

// Distance of entry p under a metric other than SquaredEuclidean.
template <class A, int K, class O, class M>
__device__ __forceinline__ O point_dist(const DevTree<A, K>& t, uint32_t p, const MetricQW<O, K, M>& w) {
  O acc = O(0);
#pragma unroll
  for (int d = 0; d < K; ++d) {
    const O c = O(t.coords[d * t.plane + p]);
    acc = w.met.combine(acc, w.met.d1(sub_rn(c, w.q[d])));
  }
  return acc;
}

// Per-thread traversal counters, summed per wave into the tree's counters: leaves
// visited, entries scanned, stem nodes read, entries re-measured in FP64 (screen only).
struct Counters {
  uint32_t leaves = 0, points = 0, stems = 0, rescored = 0;
};

// ------------------------------------------------------------------ traversal --------
// V supplies bound(), prune_at(rd, bound()) (= prune(rd)), unbounded() and leaf(tree, qw,
// first_leaf, n_leaves), which scans the entries of n_leaves consecutive leaves (one,
// except for the first visit below).
// first_up > 0: the first descent stops 2^first_up leaves above the bottom and that whole
// block of leaves (the query's own) is scanned as one visit; its subtree is not entered
// again. Any visit order finds the same canonical result (the bound only prunes what cannot
// reach it); Kiddo's own order is first_up = 0. When the DFS is done, the visitor gets one
// more visit of no leaves (n_leaves = 0), in which it may finish deferred work.
template <class A, int K, class V, class QW>
__device__ __forceinline__ void traverse(const DevTree<A, K>& t, const Small<A, K>& q, const QW& qw, V& vis,
                                         Counters& c, int first_up = 0) {
  // Register-only form of Kiddo's DFS: no per-level stack. A queued far child at level l is
  // a `pending` bit; its rect distance rd' = (rd - off[dim]^2) + new_off^2 is recomputed
  // when the bit is popped, from the same rd and off[] it was computed from at push time
  // (nothing below l has been entered then) and the same pivot, so it is bit-identical.
  // Entering far children is recorded in `taken`; rd and off[] are a function of the taken
  // levels along the current path, so popping a taken level rebuilds them by replaying the
  // remaining taken levels top-down, with the exact operations used when they were entered.
  using O = typename Ax<A>::O;
  if (t.n == 0) return;  // orchestrator/mod.rs:399-401
  const auto met = metric_of(qw);
  const int D = t.nlev;
  Small<O, K> off;
  off.fill(O(0));
  uint32_t pending = 0, taken = 0;
  O rd = O(0);
  uint32_t idx = 1;
  int level = 0, dim = 0;
  bool check = !vis.unbounded();
  int lim = D - (first_up < D ? first_up : D);
  bool final = false;
  // levels whose stems are all in LDS: level l holds indices [2^l, 2^(l+1))
  const int lds_levels = t.lds_n ? 31 - __clz(t.lds_n) : 0;
  for (;;) {
    // descend to a leaf (orchestrator/mod.rs:520-603)
    c.stems += uint32_t(lim - level);
    const O bnd = vis.bound();  // the bound only changes in a leaf visit
    // branch-free step on dimension d (= level % K): a sentinel stem (no pivot) sends the
    // query left and queues nothing
    auto step = [&](const A pivot, const int d) {
      const bool valid = pivot < Ax<A>::maxv();
      const bool right = valid && q[d] >= pivot;
      const O no = abs_diff(qw[d], O(pivot));
      const O rdf = met.rect(rd, off, d, no);
      const bool push = valid && (!check || !(rdf > bnd));
      pending |= (push ? 1u : 0u) << level;
      idx = (idx << 1) | (right ? 1u : 0u);
      ++level;
    };
    auto next_dim = [&]() { dim = (dim + 1 == K) ? 0 : dim + 1; };
    // the levels held in LDS, then the rest: each loop reads one kind of memory (a load
    // that picks LDS or global memory per step needs a generic address built every step).
    // The global levels go one at a time up to a multiple of K, then in groups of K whose
    // dimensions are constants (a run-time dimension costs a select per vector element).
    const int mid = lim < lds_levels ? lim : lds_levels;
    for (; level < mid; next_dim()) step(t.lds[idx], dim);
    for (; level < lim && dim != 0; next_dim()) step(t.stems[idx], dim);
    while (level + K <= lim) {
#pragma unroll
      for (int d = 0; d < K; ++d) step(t.stems[idx], d);
    }
#pragma unroll
    for (int d = 0; d < K - 1; ++d)
      if (level < lim) step(t.stems[idx], d);
    {
      const uint32_t nl = final ? 0u : 1u << (D - lim);
      idx <<= (D - lim);  // the block's first leaf: its ancestors above the block are unchanged
      level = D;
      lim = D;
      const uint32_t leaf = idx - (1u << D);
      c.leaves += nl;
      c.points += t.leaf_start[leaf + nl] - t.leaf_start[leaf];
      vis.leaf(t, qw, leaf, nl);  // the one call site: a second one doubles the top-k's registers
    }
    if (final) return;
    // pop frames from the deepest level upwards (orchestrator/mod.rs:450-515)
    bool entered = false;
    const O pbnd = vis.bound();  // constant until the next leaf visit
    while (pending | taken) {
      const int l = 31 - __clz(pending | taken);
      const uint32_t bit = 1u << l;
      if (taken & bit) {
        // every taken level below the deepest pending one is popped before it: clear them
        // together and rebuild once
        const uint32_t upto = pending ? (2u << (31 - __clz(pending))) - 1u : 0u;
        taken &= upto;
        // rebuild rd and off[] from the remaining taken levels, top-down
        rd = O(0);
        off.fill(O(0));
        auto replay = [&](const int tl, const A pv) {
          const int td = tl % K;
          c.stems += 1;
          const O no = abs_diff(qw[td], O(pv));
          rd = met.rect(rd, off, td, no);
          off.set(td, no);
        };
        // the LDS levels, then the rest (one kind of memory per loop, as in the descent)
        const uint32_t lds_mask = (1u << lds_levels) - 1u;
        for (uint32_t rest = taken & lds_mask; rest; rest &= rest - 1u) {
          const int tl = __ffs(rest) - 1;
          replay(tl, t.lds[idx >> (D - tl)]);
        }
        for (uint32_t rest = taken & ~lds_mask; rest; rest &= rest - 1u) {
          const int tl = __ffs(rest) - 1;
          replay(tl, t.stems[idx >> (D - tl)]);
        }
        continue;
      }
      pending ^= bit;
      const int ld = l % K;
      const A pivot = stem_at(t, idx >> (D - l));
      c.stems += 1;
      const O no = abs_diff(qw[ld], O(pivot));
      const O frd = met.rect(rd, off, ld, no);
      if (vis.prune_at(frd, pbnd)) continue;
      taken |= bit;
      off.set(ld, no);
      rd = frd;
      idx = (idx >> (D - l - 1)) ^ 1u;  // far child of the level-l ancestor
      level = l + 1;
      dim = level % K;
      check = true;
      entered = true;
      break;
    }
    if (!entered) final = true;
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
  __device__ bool prune_at(O rd, O b) const { return rd > b; }
  __device__ bool unbounded() const { return true; }
  // One entry at a time: the item is loaded only for an entry at or below the best
  // distance. (Loading 4 entries' coordinates and items together was 13% slower on a
  // 16M-point uniform f64 tree, 8.83 against 7.80 ms for 16M queries.)
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
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
// The middle of three values. For a <= c it is b clamped into [a, c]: slot j of a sorted
// list after b is inserted, from the old slots j-1 and j. Distances here are never NaN.
__device__ __forceinline__ float med3(float a, float b, float c) { return __builtin_amdgcn_fmed3f(a, b, c); }
__device__ __forceinline__ double med3(double a, double b, double c) { return fmin(c, fmax(a, b)); }

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
    return (ad < bd) | ((ad == bd) & (ai < bi));  // bitwise: no branch per slot
  }
  // inserts c, dropping the last slot
  __device__ __forceinline__ void insert(O cd, uint32_t ci) { insert_if(true, cd, ci); }
  // inserts c when ok (a predicate rather than a branch: no copy of the slots is made)
  // The distances shift by clamping (one med3, or a max and a min, per slot); the items
  // follow the (distance, item) comparisons, which also place equal distances.
  __device__ __forceinline__ void insert_if(bool ok, O cd, uint32_t ci) {
    bool lt[MAXK];
#pragma unroll
    for (int j = 0; j < MAXK; ++j) lt[j] = ok & less(cd, ci, d[j], it[j]);
    const O c = ok ? cd : oinf<O>();  // +inf leaves every distance slot as it is
#pragma unroll
    for (int j = MAXK - 1; j > 0; --j) {
      it[j] = lt[j - 1] ? it[j - 1] : (lt[j] ? ci : it[j]);
      d[j] = med3(d[j - 1], c, d[j]);
    }
    it[0] = lt[0] ? ci : it[0];
    d[0] = d[0] < c ? d[0] : c;
  }
  // Batcher's odd-even merge sort over the slots in (distance, item) order: the same list
  // as inserting the entries one at a time into an empty list (both keep the MAXK smallest
  // in order), for (MAXK/4) (log2 MAXK)^2 compare-exchanges instead of MAXK insertions of
  // MAXK slots each.
  __device__ __forceinline__ void sort() { sort_step<0>(); }
  template <int C> __device__ __forceinline__ void sort_step() {
    if constexpr (C < kNet.n) {
      constexpr int a = kNet.a[C], b = kNet.b[C];
      const bool sw = less(d[b], it[b], d[a], it[a]);
      const O lo = fmin(d[a], d[b]), hi = fmax(d[a], d[b]);  // equal distances: either
      const uint32_t il = sw ? it[b] : it[a], ih = sw ? it[a] : it[b];
      d[a] = lo;
      d[b] = hi;
      it[a] = il;
      it[b] = ih;
      sort_step<C + 1>();
    }
  }
  struct Net {
    int n = 0;
    unsigned char a[MAXK * 16] = {}, b[MAXK * 16] = {};
    constexpr Net() {
      for (int p = 1; p < MAXK; p <<= 1)
        for (int k = p; k >= 1; k >>= 1)
          for (int j = k % p; j + k < MAXK; j += 2 * k)
            for (int i = 0; i < k && i + j + k < MAXK; ++i)
              if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                a[n] = (unsigned char)(i + j);
                b[n] = (unsigned char)(i + j + k);
                ++n;
              }
    }
  };
  static constexpr Net kNet{};
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
// `cap` (default +inf) is an optional certified upper bound on the k-th distance, known
// before the traversal (k entries at most that far exist): entries and subtrees beyond it
// cannot reach the result, so it bounds the search from the start. Equal distances stay in.
template <class O, int MAXK> struct KnnVis {
  TopK<O, MAXK> top;
  int k, cnt;
  O radius;
  bool excl;
  O kd;
  uint32_t kit;
  O cap;
  __device__ void init(int k_, O r, bool ex) {
    top.init();
    k = k_;
    cnt = 0;
    radius = r;
    excl = ex;
    kd = oinf<O>();
    kit = 0xFFFFFFFFu;
    cap = oinf<O>();
  }
  __device__ O bound() const {
    const O b = (cnt >= k && kd < radius) ? kd : radius;
    return cap < b ? cap : b;
  }
  __device__ bool prune(O rd) const { return rd > bound() || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>() && cap == oinf<O>(); }
  __device__ __forceinline__ void offer(O d, uint32_t it) {
    top.insert(d, it);
    cnt = cnt < k ? cnt + 1 : k;
    if (cnt >= k) top.kth(k, kd, kit);
  }
  __device__ __forceinline__ void offer_if(bool ok, O d, uint32_t it) {
    top.insert_if(ok, d, it);
    cnt = ok ? (cnt < k ? cnt + 1 : k) : cnt;
    if (cnt >= k) top.kth(k, kd, kit);
  }
  __device__ __forceinline__ bool candidate(O d) const {
    return (excl ? d < radius : d <= radius) && d <= cap && (cnt < k || d <= kd);
  }
  // MAXK <= 4: one pass, the distances of 4 entries in one branch-free block (indices
  // clamped to the leaf, so their loads are in flight together), candidates offered in
  // entry order. The small top-k costs few registers even when the compiler keeps a
  // second copy of it for the four update sites.
  //
  // MAXK >= 8: two passes. The first measures the first 64 entries (4 at a time) and marks
  // in a bit mask those that pass the bound of the start of the leaf (all of them,
  // unmeasured, while the bound is +inf); the second measures the marked ones again and
  // offers them, in entry order, against the current bound. Entries past the first 64
  // (long leaves of equal coordinates) follow in the same loop. The top-k is updated at
  // one place inside one loop: a second update site, or a loop around this one, makes the
  // compiler hold a second copy of the top-k (+3 registers per slot).
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    constexpr uint32_t kChunk = 4;
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    if constexpr (MAXK <= 4) {
      for (uint32_t p0 = s; p0 < e; p0 += kChunk) {
        O dd[kChunk];
#pragma unroll
        for (uint32_t u = 0; u < kChunk; ++u) dd[u] = point_dist(t, p0 + u < e ? p0 + u : e - 1, qw);
#pragma unroll
        for (uint32_t u = 0; u < kChunk; ++u) {
          const uint32_t p = p0 + u;
          if (p < e && candidate(dd[u])) {
            const uint32_t it = t.items[p];
            offer_if(cnt < k || TopK<O, MAXK>::less(dd[u], it, kd, kit), dd[u], it);
          }
        }
      }
      return;
    }
    const uint32_t L = e - s, nm = L < 64u ? L : 64u;
    // unbounded (top-k not yet full, no radius): every entry is a candidate, unmeasured
    const bool bounded = bound() < oinf<O>();
    uint64_t mask = bounded ? 0ull : (nm == 64u ? ~0ull : (1ull << nm) - 1ull);
    for (uint32_t j0 = 0; bounded && j0 < nm; j0 += kChunk) {
      O dd[kChunk];
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) dd[u] = point_dist(t, s + (j0 + u < nm ? j0 + u : nm - 1), qw);
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u)
        mask |= uint64_t((j0 + u < nm) && candidate(dd[u]) ? 1u : 0u) << (j0 + u);
    }
    uint32_t next = nm;  // first entry not covered by the mask
    // one entry ahead: the loads of the next entry are issued before the current one is
    // offered (every step is branch-free, so they can overlap the insertion)
    bool have = mask != 0ull || next < L;
    uint32_t j = mask != 0ull ? uint32_t(__ffsll((unsigned long long)mask) - 1) : next;
    next += (mask == 0ull && have) ? 1u : 0u;
    mask &= mask - 1ull;
    O dn = have ? point_dist(t, s + j, qw) : O(0);
    uint32_t in = have ? t.items[s + j] : 0u;
    while (have) {
      const O d = dn;
      const uint32_t it = in;
      have = mask != 0ull || next < L;
      j = mask != 0ull ? uint32_t(__ffsll((unsigned long long)mask) - 1) : (next < L ? next : L - 1);
      next += (mask == 0ull && have) ? 1u : 0u;
      mask &= mask - 1ull;
      dn = point_dist(t, s + j, qw);  // in bounds: j < L (a repeat of the last entry when done)
      in = t.items[s + j];
      offer_if(candidate(d) && (cnt < k || TopK<O, MAXK>::less(d, it, kd, kit)), d, it);
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
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  // The distances of 4 entries are computed in one branch-free block (indices clamped to
  // the leaf), so their coordinate loads are in flight together; results follow in entry
  // order.
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    constexpr uint32_t kChunk = 4;
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    for (uint32_t p0 = s; p0 < e; p0 += kChunk) {
      O dd[kChunk];
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) dd[u] = point_dist(t, p0 + u < e ? p0 + u : e - 1, qw);
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) {
        const uint32_t p = p0 + u;
        const bool ok = p < e && (excl ? dd[u] < radius : dd[u] <= radius);
        if (!ok) continue;
        if (out_items && count < cap) {
          out_items[count] = t.items[p];
          out_d[count] = dd[u];
        }
        ++count;
      }
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
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  // 4 entries per step, their coordinates and items loaded together (indices clamped to
  // the leaf), then the ones in range offered in entry order at one insertion site (a
  // site per entry of the step would make the compiler keep copies of the list). One entry
  // at a time, a query waited for each entry's loads in turn: n = 5, r = 0.001 on a
  // clustered 12M-point f64 tree took 2.05 / 7.88 ms for 1k / 1M queries, 1.58 / 6.73 now.
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    constexpr uint32_t kChunk = 4;
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    for (uint32_t p0 = s; p0 < e; p0 += kChunk) {
      O dv[kChunk];
      uint32_t iv[kChunk];
      uint32_t okm = 0;
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) {
        const uint32_t p = p0 + u < e ? p0 + u : e - 1;
        dv[u] = point_dist(t, p, qw);
        iv[u] = t.items[p];
      }
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u)
        okm |= ((p0 + u < e) && (excl ? dv[u] < radius : dv[u] <= radius) ? 1u : 0u) << u;
      while (okm) {
        const uint32_t u = uint32_t(__ffs(okm) - 1);
        okm &= okm - 1u;
        O dd = dv[0];
        uint32_t ci = iv[0];
#pragma unroll
        for (uint32_t j = 1; j < kChunk; ++j)
          if (u == j) {
            dd = dv[j];
            ci = iv[j];
          }
        if (cnt >= n && !(ci < kit)) continue;
        insert(dd, ci);
      }
    }
  }
  // inserted after the kept entries of an equal item (the first one visited stays first)
  __device__ __forceinline__ void insert(O dd, uint32_t ci) {
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
};

// FP32-screened k-NN for f64 trees. Every leaf entry is first measured in FP32 from the
// shadow copy; only entries whose FP32 distance is within the certified error bound of the
// current FP64 bound are re-measured in FP64 and offered to the collector, so the result
// is the plain kernel's. Bound derivation, with u = 2^-24 and eta = 2^-150:
//   |fl32(p) - p| <= u|p| + eta, so each FP32 coordinate difference delta_i satisfies
//   |delta_i - (p_i - q_i)| <= u(2+u)(|p_i| + |q_i|) + 3 eta <= eps (per query, from the
//   largest |coordinate| of the tree and the query), hence ||delta|| <= ||t|| + sqrt(K) eps.
//   The FP32 sum of K <= 4 non-negative squares is at most ||delta||^2 (1+u)^K <=
//   ||delta||^2 (1 + 5u), plus K 2^-150 < 1e-44 when products fall in the subnormal range,
//   and the exact distance is at most the FP64 one times (1 + 1e-15). So FP64 dist <= T
//   implies FP32 dist <= (sqrt(T (1+1e-15)) + sqrt(K) eps)^2 (1 + 5u) + 1e-44, rounded up.
// The square is expanded as T' + 2 e s + e^2 with s >= sqrt(T'): an FP32 square root of T'
// rounded up to FP32, inflated by 2^-19 (far above any FP32 square root's error), so the
// threshold costs one FP32 square root instead of an FP64 one.
__device__ __forceinline__ float next_up(float f) { return __uint_as_float(__float_as_uint(f) + 1u); }

// FP32 squared distance of a shadow entry in Kiddo's order, every product and sum rounded.
// The shadow holds {x, y, z, item bits} for K <= 3 and {x, y, z, w} for K = 4.
template <int K>
__device__ __forceinline__ float shadow_sq(const float4 sh, const float (&qf)[4]) {
  const float dx = sub_rn(sh.x, qf[0]);
  float acc = mul_rn(dx, dx);
  const float dy = sub_rn(sh.y, qf[1]);
  acc = add_rn(acc, mul_rn(dy, dy));
  if (K > 2) {
    const float dz = sub_rn(sh.z, qf[2]);
    acc = add_rn(acc, mul_rn(dz, dz));
  }
  if (K > 3) {
    const float dw = sub_rn(sh.w, qf[3]);
    acc = add_rn(acc, mul_rn(dw, dw));
  }
  return acc;
}
__device__ __forceinline__ float screen_threshold(double T, double eps_k) {
  if (!(T < __builtin_huge_val())) return __builtin_huge_valf();
  const double tp = T * (1.0 + 1e-15);
  float tf = (float)tp;
  if ((double)tf < tp) tf = next_up(tf);  // tf >= tp (tp >= 0)
  const double s = double(sqrtf(tf)) * (1.0 + 0x1p-19);
  const double tt = (tp + 2.0 * eps_k * s + eps_k * eps_k) * (1.0 + 5.0 * 0x1p-24) + 1e-44;
  float f = (float)tt;
  if ((double)f < tt) f = next_up(f);
  return f;
}

// Per-query FP32 screen set-up: the query rounded to FP32, and sqrt(K) times the bound on
// the error of an FP32 coordinate difference (see screen_threshold).
template <int K>
__device__ __forceinline__ void screen_setup(const Small<double, K>& q, const double* __restrict__ max_abs,
                                             float (&qf)[4], double& eps_k) {
  double amax = 0.0;
#pragma unroll
  for (int d = 0; d < 4; ++d) qf[d] = 0.0f;
#pragma unroll
  for (int d = 0; d < K; ++d) {
    qf[d] = (float)q[d];
    amax = fmax(amax, max_abs[d] + fabs(q[d]));
  }
  const double u = 0x1p-24;
  const double eps = u * (2.0 + u) * amax + 3.0 * 0x1p-150;
  eps_k = sqrt((double)K) * eps * (1.0 + 1e-6);
}

// f32 trees: the query itself, and no error term (the shadow holds the coordinates exactly)
template <int K>
__device__ __forceinline__ void screen_setup(const Small<float, K>& q, const double* __restrict__,
                                             float (&qf)[4], double& eps_k) {
#pragma unroll
  for (int d = 0; d < 4; ++d) qf[d] = d < K ? q[d] : 0.0f;
  eps_k = 0.0;
}

// ------------------------------------------------ collectors for large k or n --------
// Above the register forms (k or n > 64) the collector is a binary max-heap held in the
// query's own output row (k entries of HBM, cached in L2), as Kiddo's BinaryHeap collector
// for large k (result_collection.rs:365-461, BinaryHeap of BestQueryResultItem for
// best_n_within, :463-561). The root is the largest kept entry, copied in registers: it
// is the bound once the heap is full. A full heap takes a candidate that is smaller than
// the root (replace the root, sift down); a heap not yet full appends and sifts up. The
// row is sorted in place when the traversal ends (sort_segment) and padded to k.
// Order: Key::less(a, b) over (distance, item) entries; the heap keeps the k smallest.
struct KnnKey {  // nearest_n: (distance, then lower item)
  template <class O> static __device__ __forceinline__ bool less(O ad, uint32_t ai, O bd, uint32_t bi) {
    return ad < bd || (ad == bd && ai < bi);
  }
};
struct ItemKey {  // best_n_within: item only (best_query_result_item.rs:27-31)
  template <class O> static __device__ __forceinline__ bool less(O, uint32_t ai, O, uint32_t bi) { return ai < bi; }
};

template <class O, class Key> struct RowHeap {
  uint32_t* __restrict__ it;
  O* __restrict__ d;
  uint32_t cap, cnt;
  O rd;          // root (largest kept entry) once cnt > 0
  uint32_t rit;
  __device__ void init(uint32_t* it_, O* d_, uint32_t cap_) {
    it = it_;
    d = d_;
    cap = cap_;
    cnt = 0;
    rd = oinf<O>();
    rit = 0xFFFFFFFFu;
  }
  __device__ bool full() const { return cnt >= cap; }
  // would the heap keep (cd, ci)?
  __device__ bool takes(O cd, uint32_t ci) const { return !full() || Key::less(cd, ci, rd, rit); }
  // places (cd, ci) at slot j of the heap [0, n) whose children below j are heaps
  __device__ void sift_down(uint32_t j, uint32_t n, O cd, uint32_t ci) {
    for (;;) {
      const uint32_t c = 2 * j + 1;
      if (c >= n) break;
      O bd = d[c];
      uint32_t bi = it[c], b = c;
      if (c + 1 < n) {
        const O d2 = d[c + 1];
        const uint32_t i2 = it[c + 1];
        if (Key::less(bd, bi, d2, i2)) {
          bd = d2;
          bi = i2;
          b = c + 1;
        }
      }
      if (!Key::less(cd, ci, bd, bi)) break;
      d[j] = bd;
      it[j] = bi;
      j = b;
    }
    d[j] = cd;
    it[j] = ci;
  }
  // requires takes(cd, ci)
  __device__ void add(O cd, uint32_t ci) {
    if (!full()) {
      uint32_t j = cnt++;
      while (j > 0) {  // sift up
        const uint32_t p = (j - 1) >> 1;
        const O pd = d[p];
        const uint32_t pi = it[p];
        if (!Key::less(pd, pi, cd, ci)) break;
        d[j] = pd;
        it[j] = pi;
        j = p;
      }
      d[j] = cd;
      it[j] = ci;
    } else {
      sift_down(0, cnt, cd, ci);  // replaces the root
    }
    rd = d[0];
    rit = it[0];
  }
  // Heapsort's extraction phase (the row is a max-heap already): the kept entries end up
  // ascending in Key order; the row is padded to cap with (item 0xFFFFFFFF, +inf).
  __device__ void finish() {
    for (uint32_t e = cnt; e > 1; --e) {
      const O td = d[e - 1];
      const uint32_t ti = it[e - 1];
      d[e - 1] = d[0];
      it[e - 1] = it[0];
      sift_down(0, e - 1, td, ti);
    }
    for (uint32_t j = cnt; j < cap; ++j) {
      it[j] = 0xFFFFFFFFu;
      d[j] = oinf<O>();
    }
  }
};

// nearest_n / nearest_n().within for k above the register forms: the rules of KnnVis with
// the top-k in a RowHeap. SCREEN (f64 trees with the FP32 shadow): an entry is measured in
// FP64 only when its FP32 distance passes the certified threshold of the current bound
// (screen_threshold), so the result is the plain one.
template <class O, bool SCREEN> struct HeapKnnVis {
  RowHeap<O, KnnKey> h;
  O radius;
  bool excl;
  float qf[4];   // SCREEN: the query in FP32
  double eps_k;  // SCREEN: sqrt(K) * per-query coordinate-difference error bound
  float thr;     // SCREEN: FP32 threshold of the current bound
  uint32_t rescored;
  __device__ O bound() const { return (h.full() && h.rd < radius) ? h.rd : radius; }
  __device__ bool prune(O rd) const { return rd > bound() || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    for (uint32_t p = s; p < e; ++p) {
      if constexpr (SCREEN) {
        const float4 sh = t.shadow[p];
        const float acc = shadow_sq<K>(sh, qf);
        if (acc > thr) continue;
        ++rescored;
      }
      const O d = point_dist(t, p, qw);
      if (!(excl ? d < radius : d <= radius)) continue;
      if (h.full() && !(d <= h.rd)) continue;
      const uint32_t it = t.items[p];
      if (!h.takes(d, it)) continue;
      h.add(d, it);
      if constexpr (SCREEN) {
        if (h.full()) thr = screen_threshold(double(bound()), eps_k);
      }
    }
  }
};

// best_n_within for n above the register form: the n smallest items within the radius
// (best_n_within.rs:163-196) in a RowHeap ordered by item.
template <class O> struct HeapBestNVis {
  RowHeap<O, ItemKey> h;
  O radius;
  bool excl;
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    for (uint32_t p = s; p < e; ++p) {
      const uint32_t it = t.items[p];
      if (!h.takes(O(0), it)) continue;
      const O d = point_dist(t, p, qw);
      if (!(excl ? d < radius : d <= radius)) continue;
      h.add(d, it);
    }
  }
};

// A bit set of W*64 entries of a visit, taken lowest first (constant word indices only, so
// it stays in registers).
template <int W> struct EntryMask {
  uint64_t m[W];
  __device__ __forceinline__ void clear() {
#pragma unroll
    for (int w = 0; w < W; ++w) m[w] = 0ull;
  }
  __device__ __forceinline__ void set_first(uint32_t n) {  // entries [0, n), n <= 64 W
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t lo = 64u * uint32_t(w);
      m[w] = n >= lo + 64u ? ~0ull : n > lo ? (1ull << (n - lo)) - 1ull : 0ull;
    }
  }
  __device__ __forceinline__ void set4(uint32_t j0, uint32_t bits) {  // j0 % 4 == 0
#pragma unroll
    for (int w = 0; w < W; ++w)
      if (j0 >= 64u * uint32_t(w) && j0 < 64u * uint32_t(w + 1)) m[w] |= uint64_t(bits) << (j0 - 64u * uint32_t(w));
  }
  __device__ __forceinline__ uint32_t count() const {
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) c += uint32_t(__popcll(m[w]));
    return c;
  }
  __device__ __forceinline__ bool any() const {
    uint64_t a = 0ull;
#pragma unroll
    for (int w = 0; w < W; ++w) a |= m[w];
    return a != 0ull;
  }
  __device__ __forceinline__ uint32_t pop() {  // requires any()
    uint32_t j = 0;
    bool found = false;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const bool here = !found && m[w] != 0ull;
      j = here ? 64u * uint32_t(w) + uint32_t(__ffsll((unsigned long long)m[w]) - 1) : j;
      m[w] = here ? m[w] & (m[w] - 1ull) : m[w];
      found = found || here;
    }
    return j;
  }
};

template <int MAXK> struct ScreenKnnVis {
  KnnVis<double, MAXK> v;
  float qf[4];
  double eps_k;
  float thr;
  uint32_t rescored;
  uint32_t visits, qhead, qn;  // visits so far; deferred survivors (a ring in LDS)
  uint32_t skip;                // entries of the first visit already offered (first_block)
  __device__ double bound() const { return v.bound(); }
  __device__ bool prune(double rd) const { return v.prune(rd); }
  __device__ bool prune_at(double rd, double b) const { return v.prune_at(rd, b); }
  __device__ bool unbounded() const { return v.unbounded(); }
  template <int K>
  __device__ __forceinline__ float screen_dist(const DevTree<double, K>& t, uint32_t p) const {
    const float4 sh = t.shadow[p];
    const float acc = shadow_sq<K>(sh, qf);
    return acc;
  }
  // The next survivor at or after entry `next` past the mask (entries screened one at a
  // time; long leaves of equal coordinates, or the tail of a large first visit): its index,
  // or L when there is none.
  template <int K>
  __device__ __forceinline__ uint32_t next_tail(const DevTree<double, K>& t, uint32_t s, uint32_t next,
                                                uint32_t L) const {
    while (next < L && screen_dist(t, s + next) > thr) ++next;
    return next;
  }
  // Survivors past kDeferBudget in a visit after the first are deferred: their entry
  // indices wait in an LDS ring of kDeferQ per lane and are offered in later visits within
  // the same budget, or in the final visit (n_leaves = 0), which drains them. A wave steps
  // through a visit's survivors at the pace of its busiest lane, so the budget bounds the
  // steps per visit. Deferring only delays tightening the bound (more is visited, nothing
  // is lost): a survivor is measured in FP64 and checked against the bound of its turn, and
  // an entry of the final result passes every bound.
  // Measured on 16.8M points, k = 16: budgets 2 / 4 / 6 / 9 took 24.3 / 20.1 / 20.9 / 21.7 ms;
  // small top-k lists gain less from deferring.
  static constexpr uint32_t kDeferBudget = MAXK >= 8 ? 4u : 8u, kDeferQ = 16;
  static constexpr int kWords = MAXK >= 16 ? 2 : 1;
  static constexpr uint32_t kMaskN = 64u * kWords;
  // marks in mk the entries among [s, s + nm) whose FP32 distance passes thr: shadow loads
  // issued 4 at a time from one address (up to 3 past the range: the shadow's padding),
  // bits gathered per 64-entry word
  template <int K>
  __device__ __forceinline__ void screen_into(const DevTree<double, K>& t, uint32_t s, uint32_t nm,
                                              EntryMask<kWords>& mk) const {
    constexpr uint32_t kBatch = 4;
#pragma unroll
    for (int w = 0; w < kWords; ++w) {
      const uint32_t lo = 64u * uint32_t(w), hi = nm < lo + 64u ? nm : lo + 64u;
      uint64_t word = 0;
      for (uint32_t j0 = lo; j0 < hi; j0 += kBatch) {
        const float4* __restrict__ sp = t.shadow + s + j0;
        float4 sh[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) sh[u] = sp[u];
        uint32_t bits = 0;
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) {
          const float acc = shadow_sq<K>(sh[u], qf);
          bits |= ((j0 + u < nm) && !(acc > thr) ? 1u : 0u) << u;
        }
        word |= uint64_t(bits) << (j0 - lo);
      }
      mk.m[w] |= word;
    }
  }
  // The first visit's block [b, e) (the pre-scan's, whose cap is the bound; b = e = 0 when
  // there is none; the tree must not be empty) offered before the traversal: the survivors of the screen among its first
  // kMaskN entries, up to MAXK of them, are measured in FP64 together and sorted into the
  // empty top-k, which is the list that offering them one at a time builds. `skip` records
  // how far into the block this reached (up to the last survivor taken when more remain);
  // the traversal's visit of the block starts there.
  template <int K, class QW>
  __device__ __forceinline__ void first_block(const DevTree<double, K>& t, const QW& qw, uint32_t b, uint32_t e) {
    const uint32_t L = e - b, nm = L < kMaskN ? L : kMaskN;
    EntryMask<kWords> mk;
    mk.clear();
    screen_into(t, b, nm, mk);
    // the survivors' indices first (an empty slot reads entry b, or entry 0 of a non-empty
    // tree when the block is empty), so that all their loads can be in flight together
    uint32_t pidx[MAXK];
    uint32_t last = 0, taken = 0;
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      const bool valid = mk.any();
      const uint32_t r = mk.pop();
      last = valid ? r : last;
      taken += valid ? 1u : 0u;
      pidx[j] = b + (valid ? r : 0u);
    }
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      const double dd = point_dist(t, pidx[j], qw);
      const uint32_t it = t.items[pidx[j]];
      // bitwise, not short-circuit: a branch here lets the loads sink under it, one at a time
      const bool ok = (uint32_t(j) < taken) & ((dd < v.radius) | (!v.excl & (dd == v.radius))) & (dd <= v.cap);
      // The item is masked rather than selected: the compiler turns a select of a loaded
      // value into a branch around its load, one candidate at a time. The empty asm keeps
      // the mask opaque so that it cannot fold the OR back into such a select.
      uint32_t drop = ok ? 0u : 0xFFFFFFFFu;
      asm("" : "+v"(drop));
      v.top.d[j] = ok ? dd : oinf<double>();
      v.top.it[j] = it | drop;
    }
    v.top.sort();
    int nok = 0;
#pragma unroll
    for (int j = 0; j < MAXK; ++j) nok += v.top.d[j] < oinf<double>() ? 1 : 0;
    v.cnt = nok < v.k ? nok : v.k;
    if (v.cnt >= v.k) v.top.kth(v.k, v.kd, v.kit);
    rescored += taken;
    skip = mk.any() ? last + 1u : nm;
    thr = screen_threshold(v.bound(), eps_k);
  }
  // Two passes per visit. The FP32 screen (screen_into) marks the survivors among the first
  // kMaskN entries (128 for a top-k of 16 or more, 64 below) in a bit mask against the
  // threshold at the start of the visit; then only survivors are measured in FP64 and
  // offered, in entry order, to the collector, one per step. Entries past the mask are
  // screened one at a time when the mask is empty, deferred ones come last. The first visit
  // (the pre-scan block) starts after the entries first_block offered. The top-k is updated at one
  // place inside one loop whose exit test is at the top (the next survivor is chosen at the
  // end of a step): a second update site, a loop around this one or an exit in the middle
  // makes the compiler hold a second copy of the top-k in registers.
  template <int K, class QW>
  __device__ void leaf(const DevTree<double, K>& t, const QW& qw, uint32_t leaf, uint32_t nl) {
    __shared__ uint32_t deferred[kDeferQ * kBlock];
    const uint32_t s0 = t.leaf_start[leaf], e = nl ? t.leaf_start[leaf + nl] : s0;
    const uint32_t s = visits == 0u ? s0 + skip : s0;
    const uint32_t L = e - s, nm = L < kMaskN ? L : kMaskN;
    // unbounded (top-k not yet full, no radius): every entry survives the screen
    const bool screen = thr < __builtin_huge_valf();
    EntryMask<kWords> mk;
    mk.clear();
    if (!screen) mk.set_first(nm);
    screen_into(t, s, screen ? nm : 0u, mk);
    const uint32_t lane = threadIdx.x;
    const uint32_t budget = (visits == 0u || nl == 0u) ? 0xFFFFFFFFu : kDeferBudget;
    visits += 1u;
    for (uint32_t c = mk.count(); c > budget && qn < kDeferQ; --c) {
      deferred[((qhead + qn) % kDeferQ) * kBlock + lane] = s + mk.pop();
      ++qn;
    }
    uint32_t next = nm;  // first entry not covered by the mask
    uint32_t used = 0, p = 0;
    bool have = true;
    // the next survivor: the mask, then the entries past it, then deferred ones
    auto select = [&]() {
      have = true;
      if (mk.any()) {
        p = s + mk.pop();
      } else {
        next = next_tail(t, s, next, L);
        if (next < L) {
          p = s + next;
          ++next;
        } else if (qn > 0u && used < budget) {
          p = deferred[qhead * kBlock + lane];
          qhead = (qhead + 1u) % kDeferQ;
          --qn;
        } else {
          have = false;
        }
      }
    };
    select();
    double dn = have ? point_dist(t, p, qw) : 0.0;
    uint32_t in = have ? t.items[p] : 0u;
    while (have) {
      const uint32_t pc = p;
      const double d = dn;
      const uint32_t it = in;
      ++used;
      select();
      // the next survivor's loads are issued before this one is offered
      const uint32_t pn = have ? p : pc;
      dn = point_dist(t, pn, qw);
      in = t.items[pn];
      ++rescored;
      bool ok = (v.excl ? d < v.radius : d <= v.radius) && d <= v.cap;
      ok = ok && (v.cnt < v.k || TopK<double, MAXK>::less(d, it, v.kd, v.kit));
      v.offer_if(ok, d, it);
      // entries past the mask are screened with the current bound (for a deferred entry the
      // difference wraps and the refresh is merely early)
      if (pc - s + 1u >= kMaskN) thr = screen_threshold(v.bound(), eps_k);
    }
    // The bound only tightens, so a threshold from the start of the visit stays valid for
    // the mask; refresh it once for the next visit.
    thr = screen_threshold(v.bound(), eps_k);
  }
};

// Certified upper bound on the FP64 distance of an entry whose FP32 screen distance is s:
// the FP32 sum of K <= 4 non-negative squares is at least ||delta||^2 (1-u)^K - K 2^-150
// (a product that underflows into the subnormal range loses up to 2^-150 absolute, a sum
// with a subnormal result is exact; 1e-44 > 4 * 2^-150, and 1 / (1-u)^4 < 1 + 5u), each FP32
// coordinate difference is within eps of the exact one, so the exact distance is at most
// (sqrt((s + 1e-44)(1 + 5u)) + sqrt(K) eps)^2, and the FP64 evaluation is within (1 + 1e-15)
// of it. The FP64 square root is within a few ulps, which the (1 + 1e-15) factors cover.
__device__ __forceinline__ double fp64_cap(float s, double eps_k) {
  if (!(s < __builtin_huge_valf())) return __builtin_huge_val();
  const double r = sqrt((double(s) + 1e-44) * (1.0 + 5.0 * 0x1p-24)) * (1.0 + 1e-15) + eps_k;
  return r * r * (1.0 + 2e-15);
}

// Largest FP32 screen distance that certifies an FP64 distance within the radius R:
// fp64_cap(s) <= R for every s <= the returned value (inverse of fp64_cap, rounded down and
// checked). +inf for an infinite radius, -1 when no screen distance certifies it.
__device__ __forceinline__ float certify_limit(double R, double eps_k) {
  if (!(R < __builtin_huge_val())) return __builtin_huge_valf();
  const double r = sqrt(R / (1.0 + 4e-15)) * (1.0 - 1e-15) - eps_k;
  if (!(r > 0.0)) return -1.0f;
  const double s = (r / (1.0 + 2e-15)) * (r / (1.0 + 2e-15)) / (1.0 + 6.0 * 0x1p-24) - 1e-44;
  if (!(s > 0.0)) return -1.0f;
  float f = (float)s;
  if ((double)f > s) f = __uint_as_float(__float_as_uint(f) - 1u);
  while (f > 0.0f && !(fp64_cap(f, eps_k) <= R)) f = __uint_as_float(__float_as_uint(f) - 1u);
  return f;
}

// Per-query constants of the FP32 bound refresh (Screen32Vis::refresh).
// With A = sqrt(L + 1e-44) for an FP32 screen distance L and c = 1 + 2^-16:
//   * fp64_cap(L) <= (A (1 + 2.6u) + eps_k)^2 (1 + 2e-15) <= fl32(fl32(c fl32(sqrt(L + 2^-146)) + e1)^2 c),
//     e1 = eps_k (1 + 2^-16) rounded up: every FP32 operation loses at most a factor (1 - u),
//     and c's 256u margin covers the square root's and the other operations' rounding;
//   * screen_threshold(fp64_cap(L)) <= (A (1 + 67u) + 2.0001 eps_k)^2 (1 + 6u) + 1e-44 (the
//     threshold takes a square root inflated by 2^-19 = 32u and rounds up),
//     <= fl32(fl32(c s' + e2)^2 c) + 2^-146 with e2 = 2 eps_k (1 + 2^-16) rounded up
//     (2^-146 > 1e-44; for results in the subnormal range the addition is exact, above it the
//     factor c covers it).
// So the FP32 forms bound the FP64 ones from above and prune and screen no more than they
// do: the refresh after each visit costs a handful of FP32 operations instead of two square
// roots in FP64.
__device__ __forceinline__ float round_up_f32(double x) {
  float f = (float)x;
  if ((double)f < x) f = next_up(f);
  return f;
}

// FP32 collector of the screened k-NN (k <= 16). The traversal runs on FP32 distances only;
// FP64 is used to rescore the final candidates and for the returned distances.
//   * top[] holds the MAXK smallest FP32 distances of the entries seen so far whose FP64
//     distance is certified within the radius (fp64_cap(s) <= radius), distances only.
//     With L = top[k-1], k entries have an FP64 distance of at most fp64_cap(L) <= U (the
//     FP32 form above), so the k-th FP64 distance of the answer is at most U: the traversal
//     prunes by bound = min(U, radius), like Kiddo's by the k-th distance (an entry of the
//     answer has a distance of at most the bound of every moment, so it is never pruned).
//   * The pre-scan's block (the first visit) is in top[] before the traversal starts; its
//     entries are screened again at the end.
//   * In later visits, an entry whose FP32 distance passes thr (every entry whose FP64
//     distance can be at most the bound passes) is appended to a list of candidates in LDS
//     (kCand = k + 16 entry indices per lane). When the list is full it is compacted against the
//     current threshold: an entry of the answer passes every threshold, so none is dropped.
//     A list that is still full after compaction cannot prove it holds the whole answer:
//     the query is marked for the exact redo pass (k_nearest_n_redo) and its traversal ends
//     (every subtree pruned).
//   * A listed entry's FP32 distance goes into top[] later: each visit inserts at most kIns
//     of the list's pending entries (the first ones not yet inserted), and everything
//     pending is inserted before a compaction and at the end. top[] holding a subset of the
//     certified entries seen keeps the bound an upper bound on the answer's k-th distance; an
//     entry that compaction drops fails a threshold at least the k-th distance of that
//     subset, so inserting it would have changed nothing.
//   * At the end the block's entries and the listed ones that pass the final threshold are
//     measured in FP64 and the k smallest in (distance, item) order are kept (Batcher's
//     network for the first MAXK, insertions for the rest): the plain kernel's answer.
// A survivor of the screen costs a med3 per slot of top[] and one LDS store, instead of an
// FP64 measurement and an insertion into an FP64 (distance, item) list. Inserting each
// survivor as it was found made a wave step through a visit's survivors at the pace of its
// busiest lane, with the insertion in every step (87.7 wave steps for 10.6 lane steps per
// query); appending is cheap, and the insertions are spread over the visits.
template <int MAXK, class A = double> struct Screen32Vis {
  static_assert(MAXK <= 16, "the FP32 collector is for k <= 16");
  using O = typename Ax<A>::O;
  // f32 trees: the shadow distance is the tree's own FP32 distance (same operations in the
  // same order), so nothing needs certifying: the bound and the threshold are the k-th
  // distance of top[] itself, and radius, bound and distances are FP32
  static constexpr bool kExact = std::is_same<A, float>::value;
  // candidate list entries per lane: k + 16 (fewer compactions than k + 8: 15.4 against
  // 16.0 ms on 16M queries at k = 16, with the LDS stems cut to 8 KiB to keep 4 blocks per CU)
  static constexpr uint32_t kCand = MAXK == 1 ? 8u : uint32_t(MAXK) + 16u;
  static constexpr int kWords = MAXK >= 16 ? 2 : 1;
  static constexpr uint32_t kMaskN = 64u * kWords;
  static constexpr float kC = 1.0f + 0x1p-16f;
  // pending list entries inserted into top[] per visit (measured on 16M queries at k = 16:
  // 1 / 2 / 3 / 4 / 5 / 6 / 8: 21.9 / 18.5 / 15.50 / 15.24 / 15.18 / 15.25 / 15.43 ms, 15.67 ms
  // with each survivor inserted as it was found)
  static constexpr uint32_t kIns = 5;
  float top[MAXK];  // ascending; see init
  int k;
  O radius;
  bool excl;
  bool pre;       // the first visit is the pre-scan block, whose distances are in top already
  bool ovf;       // the candidate list overflowed: the query goes to the redo pass
  float qf[4];    // the query in FP32
  double eps_k;   // sqrt(K) * per-query coordinate-difference error bound
  float e1, e2;   // eps_k and 2 eps_k, inflated and rounded up (refresh)
  float rlow;     // certify_limit(radius)
  float thr_r;    // screen_threshold(radius)
  float thr;      // FP32 threshold of the current bound
  O bnd;          // min(U, radius); -inf after an overflow
  uint32_t nc;    // candidates in the list
  uint32_t ins;   // list entries [0, ins) are in top[], [ins, nc) pending
  uint32_t visits, rescored;
  uint32_t* cand;  // LDS: kCand x kBlock entry indices, slot j of lane l at j * kBlock + l

  // after qf and eps_k are set
  __device__ void init(int k_, O r, bool ex, uint32_t* cand_lds) {
    // the first MAXK - k slots hold -inf and never leave the front, so the last slot is
    // always the k-th smallest distance (a constant index: a run-time one puts top[] in
    // scratch memory)
#pragma unroll
    for (int j = 0; j < MAXK; ++j) top[j] = j < MAXK - k_ ? -__builtin_huge_valf() : __builtin_huge_valf();
    k = k_;
    radius = r;
    excl = ex;
    pre = false;
    ovf = false;
    nc = 0;
    ins = 0;
    visits = 0;
    rescored = 0;
    cand = cand_lds;
    if constexpr (kExact) {
      e1 = e2 = 0.0f;
      rlow = r;
      thr_r = r;
    } else {
      e1 = round_up_f32(eps_k * (1.0 + 0x1p-16));
      e2 = round_up_f32(2.0 * eps_k * (1.0 + 0x1p-16));
      rlow = certify_limit(r, eps_k);
      thr_r = screen_threshold(r, eps_k);
    }
    bnd = r;
    thr = thr_r;
  }
  __device__ O bound() const { return bnd; }
  __device__ bool prune(O rd) const { return prune_at(rd, bnd); }
  __device__ bool prune_at(O rd, O b) const { return ovf || rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return bnd == oinf<O>(); }
  // bound and threshold from top[] (called when top[] may have changed)
  __device__ __forceinline__ void refresh() {
    const float L = top[MAXK - 1];
    if constexpr (kExact) {
      bnd = ovf ? -oinf<O>() : (L < radius ? L : radius);
      thr = L < thr_r ? L : thr_r;
      return;
    }
    float U32 = __builtin_huge_valf(), t32 = __builtin_huge_valf();
    if (L < __builtin_huge_valf()) {
      const float s = sqrtf(L + 0x1p-146f) * kC;
      const float u = s + e1, t = s + e2;
      U32 = u * u * kC;
      t32 = t * t * kC + 0x1p-146f;
    }
    const O U = O(U32);
    bnd = ovf ? -oinf<O>() : (U < radius ? U : radius);
    thr = t32 < thr_r ? t32 : thr_r;
  }
  // sorted insertion of a distance (+inf changes nothing)
  __device__ __forceinline__ void insert32(float c) {
#pragma unroll
    for (int j = MAXK - 1; j > 0; --j) top[j] = med3(top[j - 1], c, top[j]);
    top[0] = fminf(top[0], c);
  }
  template <int K>
  __device__ __forceinline__ float screen_dist(const DevTree<A, K>& t, uint32_t p) const {
    return shadow_sq<K>(t.shadow[p], qf);
  }
  template <int K>
  __device__ __forceinline__ uint32_t next_tail(const DevTree<A, K>& t, uint32_t s, uint32_t next,
                                                uint32_t L) const {
    while (next < L && screen_dist(t, s + next) > thr) ++next;
    return next;
  }
  // keeps the candidates that pass the current threshold, in order; 8 at a time, with their
  // shadow loads in flight together (one at a time, a compaction was a chain of 24 dependent
  // loads, which the wave of a lane whose list filled waited for)
  template <int K> __device__ __forceinline__ void compact(const DevTree<A, K>& t) {
    const uint32_t lane = threadIdx.x;
    constexpr uint32_t kChunk = 8;
    uint32_t w = 0, wi = 0;
    for (uint32_t j0 = 0; j0 < nc; j0 += kChunk) {
      uint32_t p[kChunk];
      float4 sh[kChunk];
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) p[u] = cand[(j0 + u < nc ? j0 + u : j0) * kBlock + lane];
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) sh[u] = t.shadow[p[u]];
#pragma unroll
      for (uint32_t u = 0; u < kChunk; ++u) {
        const bool keep = (j0 + u < nc) & !(shadow_sq<K>(sh[u], qf) > thr);
        // only kept entries are stored: w <= j0 + u < nc <= kCand, a slot already read. (An
        // unconditional store of the chunk's positions past nc reached slot kCand, past the
        // array, whenever the list was full, kept everything and kCand was not a multiple of 8.)
        if (keep) cand[w * kBlock + lane] = p[u];
        w += keep ? 1u : 0u;
        wi += (keep && j0 + u < ins) ? 1u : 0u;
      }
    }
    nc = w;
    ins = wi;
  }
  // inserts the next kIns pending list entries into top[] (their loads in flight together;
  // positions past the list read entry 0 of the tree and insert nothing)
  template <int K> __device__ __forceinline__ void insert_pending(const DevTree<A, K>& t) {
    const uint32_t lane = threadIdx.x;
    uint32_t p[kIns];
    float4 sh[kIns];
#pragma unroll
    for (uint32_t u = 0; u < kIns; ++u) p[u] = ins + u < nc ? cand[(ins + u) * kBlock + lane] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kIns; ++u) sh[u] = t.shadow[p[u]];
#pragma unroll
    for (uint32_t u = 0; u < kIns; ++u) {
      const float sc = shadow_sq<K>(sh[u], qf);
      insert32((ins + u < nc) & (sc <= rlow) ? sc : __builtin_huge_valf());
    }
    ins = ins + kIns < nc ? ins + kIns : nc;
  }
  template <int K> __device__ __forceinline__ void append(const DevTree<A, K>& t, uint32_t p) {
    if (nc == kCand) {
      while (ins < nc) insert_pending(t);
      refresh();
      compact(t);
      if (nc == kCand) {
        ovf = true;
        bnd = -oinf<O>();
        return;
      }
    }
    cand[nc * kBlock + threadIdx.x] = p;
    ++nc;
  }
  // appends p, and p2 when two (the list is compacted first when they do not both fit)
  template <int K> __device__ __forceinline__ void append2(const DevTree<A, K>& t, uint32_t p, uint32_t p2, bool two) {
    const uint32_t need = two ? 2u : 1u;
    if (nc + need > kCand) {
      while (ins < nc) insert_pending(t);
      refresh();
      compact(t);
      if (nc + need > kCand) {
        ovf = true;
        bnd = -oinf<O>();
        return;
      }
    }
    cand[nc * kBlock + threadIdx.x] = p;
    if (two) cand[(nc + 1u) * kBlock + threadIdx.x] = p2;
    nc += need;
  }
  // marks in mk the entries among [s, s + nm) whose FP32 distance passes thr (shadow loads
  // issued 4 at a time from one address, up to 3 past the range: the shadow's padding)
  template <int K>
  __device__ __forceinline__ void screen_into(const DevTree<A, K>& t, uint32_t s, uint32_t nm,
                                              EntryMask<kWords>& mk) const {
    constexpr uint32_t kBatch = 4;
#pragma unroll
    for (int w = 0; w < kWords; ++w) {
      const uint32_t lo = 64u * uint32_t(w), hi = nm < lo + 64u ? nm : lo + 64u;
      uint64_t word = 0;
      for (uint32_t j0 = lo; j0 < hi; j0 += kBatch) {
        const float4* __restrict__ sp = t.shadow + s + j0;
        float4 sh[kBatch];
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) sh[u] = sp[u];
        uint32_t bits = 0;
#pragma unroll
        for (uint32_t u = 0; u < kBatch; ++u) bits |= !(shadow_sq<K>(sh[u], qf) > thr) ? 1u << u : 0u;
        word |= uint64_t(bits) << (j0 - lo);
      }
      // the last batch's entries past nm are dropped here, once per word
      mk.m[w] |= hi > lo ? word & (hi - lo >= 64u ? ~0ull : (1ull << (hi - lo)) - 1ull) : 0ull;
    }
  }
  // A visit after the pre-scan block: the screen marks the survivors among the first kMaskN
  // entries against the threshold of the start of the visit and they are appended to the
  // candidate list, then up to kIns pending list entries go into top[]; entries past the mask
  // are screened one at a time. The bound and threshold are refreshed at the end of the visit
  // (and after each entry past the mask, for long leaves of equal coordinates).
  template <int K, class QW>
  __device__ __forceinline__ void leaf(const DevTree<A, K>& t, const QW&, uint32_t leaf, uint32_t nl) {
    const bool block = visits == 0u && pre;
    visits += 1u;
    if (nl == 0u || ovf || block) return;
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    const uint32_t L = e - s, nm = L < kMaskN ? L : kMaskN;
    const bool screen = thr < __builtin_huge_valf();
    EntryMask<kWords> mk;
    mk.clear();
    if (!screen) mk.set_first(nm);
    screen_into(t, s, screen ? nm : 0u, mk);
    uint32_t next = nm;  // first entry not covered by the mask
    // two survivors per step (16M queries at k = 16: 15.11 ms; one per step 15.16, four 15.15)
    while (mk.any() && !ovf) {
      const uint32_t j = mk.pop();
      const bool two = mk.any();
      const uint32_t j2 = two ? mk.pop() : j;
      append2(t, s + j, s + j2, two);
    }
    if (!ovf) insert_pending(t);
    while (!ovf) {
      next = next_tail(t, s, next, L);
      if (next >= L) break;
      append(t, s + next);
      ++next;
      if (!ovf) insert_pending(t);
      refresh();
    }
    refresh();
  }
  // an entry's distance in the tree's own arithmetic and its item: on f32 trees both from the
  // shadow entry (the same operations in the same order; the item bits in w for K <= 3), one
  // 16-byte load instead of one per coordinate plane and one for the item; on f64 trees the
  // item from the shadow entry too, whose line the screen read (16M queries at k = 16:
  // 15.11 against 15.13 ms with the items array)
  template <int K, class QW>
  __device__ __forceinline__ void final_entry(const DevTree<A, K>& t, uint32_t p, const QW& qw, O& dd,
                                              uint32_t& it) const {
    if constexpr (kExact) {
      const float4 sh = t.shadow[p];
      dd = shadow_sq<K>(sh, qf);
      it = K <= 3 ? __float_as_uint(sh.w) : t.items[p];  // K = 4: w is a coordinate
    } else {
      dd = point_dist(t, p, qw);
      it = K <= 3 ? __float_as_uint(t.shadow[p].w) : t.items[p];
    }
  }
  // The answer: the pre-scan block's entries [bb, be) and the listed candidates that pass
  // the final threshold, measured in FP64, in (distance, item) order; ok entries are those
  // within the radius.
  template <int K, class QW>
  __device__ __forceinline__ void finish(const DevTree<A, K>& t, const QW& qw, uint32_t bb, uint32_t be,
                                         TopK<O, MAXK>& res, int& cnt) {
    while (ins < nc) insert_pending(t);
    refresh();
    compact(t);
    const uint32_t lane = threadIdx.x;
    const uint32_t L = be - bb, nm = L < kMaskN ? L : kMaskN;
    EntryMask<kWords> mk;
    mk.clear();
    const bool screen = thr < __builtin_huge_valf();
    if (!screen) mk.set_first(nm);
    screen_into(t, bb, screen ? nm : 0u, mk);
    uint32_t next = nm, li = 0, taken = 0;
    // the next candidate: the block's mask, the block past the mask, then the list
    auto next_cand = [&](uint32_t& p) -> bool {
      if (mk.any()) {
        p = bb + mk.pop();
        return true;
      }
      next = next_tail(t, bb, next, L);
      if (next < L) {
        p = bb + next;
        ++next;
        return true;
      }
      if (li < nc) {
        p = cand[li * kBlock + lane];
        ++li;
        return true;
      }
      return false;
    };
    // the first MAXK together (an empty slot reads entry bb, in the tree), so that all their
    // loads are in flight at once, sorted by Batcher's network
    uint32_t pidx[MAXK];
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      uint32_t p = bb;
      taken += next_cand(p) ? 1u : 0u;
      pidx[j] = p;
    }
    int nok = 0;
#pragma unroll
    for (int j = 0; j < MAXK; ++j) {
      O dd;
      uint32_t it;
      final_entry(t, pidx[j], qw, dd, it);
      const bool ok = (uint32_t(j) < taken) & ((dd < radius) | (!excl & (dd == radius)));
      // the item is masked rather than selected (a select of a loaded value becomes a
      // branch around its load); the empty asm keeps the mask opaque
      uint32_t drop = ok ? 0u : 0xFFFFFFFFu;
      asm("" : "+v"(drop));
      res.d[j] = ok ? dd : oinf<O>();
      res.it[j] = it | drop;
      nok += ok ? 1 : 0;
    }
    res.sort();
    uint32_t p = bb;
    while (next_cand(p)) {
      ++taken;
      O dd;
      uint32_t it;
      final_entry(t, p, qw, dd, it);
      const bool ok = (dd < radius) | (!excl & (dd == radius));
      nok += ok ? 1 : 0;
      res.insert_if(ok && TopK<O, MAXK>::less(dd, it, res.d[MAXK - 1], res.it[MAXK - 1]), dd, it);
    }
    rescored = taken;
    cnt = nok < k ? nok : k;
  }
};

// Sums the wave's counters and adds them, one atomic each, to the block's slot of
// kStatSlots groups of 4 (read back summed). One group for a whole launch made a million
// atomics on one cache line, which serialise: about 6 ms per 16M-query launch.
// Every lane of the wave must reach this call.
__device__ __forceinline__ void flush_counters(const Counters& c, unsigned long long* __restrict__ out) {
  if (!out) return;
  unsigned long long v[4] = {c.leaves, c.points, c.stems, c.rescored};
#pragma unroll
  for (int j = 0; j < 4; ++j)
    for (int o = 32; o > 0; o >>= 1) v[j] += __shfl_xor(v[j], o);
  if ((threadIdx.x & 63) == 0) {
    unsigned long long* slot = out + 4u * (blockIdx.x % uint32_t(kStatSlots));
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (v[j]) atomicAdd(&slot[j], v[j]);
  }
}

template <class A, int K>
__device__ __forceinline__ void load_query(const A* __restrict__ queries, uint64_t qi, Small<A, K>& q,
                                           Small<typename Ax<A>::O, K>& qw) {
  using O = typename Ax<A>::O;
  q.fill(A(0));
  qw.fill(O(0));
#pragma unroll
  for (int d = 0; d < K; ++d) {
    const A v = queries[qi * K + d];
    q.set(d, v);
    qw.set(d, O(v));
  }
}


// Kernel prologue shared by every family: one thread per query, queries optionally
// visited in a permuted (Morton) order; results always land at the query's own index.
// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i to XCD i mod 8), and
// each XCD has its own L2. When queries run in Morton order, neighbouring blocks share
// stems and leaves, so block i takes the logical block that gives XCD (i mod 8) one
// contiguous run of the batch (a bijection on [0, gridDim.x)).
__device__ __forceinline__ uint32_t xcd_block(uint32_t i, uint32_t g) {
  constexpr uint32_t kXcds = 8;
  const uint32_t q = g / kXcds, r = g % kXcds, x = i % kXcds;
  return x * q + (x < r ? x : r) + i / kXcds;
}

#define KD_QUERY_PROLOGUE                                             \
  using O = typename Ax<A>::O;                                        \
  pin_stems(t);                                                       \
  const uint64_t i = uint64_t(order ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x) * kBlock + threadIdx.x; \
  const bool active = i < m;                                          \
  const uint64_t qi = active ? (order ? order[i] : i) : 0;            \
  Small<A, K> q;                                                      \
  Small<O, K> qw;                                                     \
  if (active) load_query<A, K>(queries, qi, q, qw);                   \
  Counters cnt;

// ------------------------------------------------------------ group-per-query --------
// G lanes (G divides 64) answer one query together. Control flow is uniform inside a
// group: every lane walks the same traversal, the lanes split each leaf's entries between
// them, and lane gl keeps
//   * slot gl of the sorted top-k (k <= G), so inserting a candidate is one shift across
//     the group's lanes;
//   * the DFS stack entries of levels gl, gl + G, ... in registers.

template <int G> struct GroupStack {
  static constexpr int R = (kMaxLevels + G - 1) / G;
};

template <class O, int G>
__device__ __forceinline__ void gs_write(O (&reg)[GroupStack<G>::R], int gl, int l, O v) {
  const int r = l / G, owner = l % G;
#pragma unroll
  for (int j = 0; j < GroupStack<G>::R; ++j)
    if (j == r && gl == owner) reg[j] = v;
}

template <class O, int G>
__device__ __forceinline__ O gs_read(const O (&reg)[GroupStack<G>::R], int l) {
  const int r = l / G;
  O v = reg[0];
#pragma unroll
  for (int j = 1; j < GroupStack<G>::R; ++j)
    if (j == r) v = reg[j];
  return __shfl(v, l % G, G);
}

// The traversal of `traverse` with the stack spread over the group's lanes. V::leaf is
// called by every lane of the group.
template <class A, int K, int G, class V>
__device__ __forceinline__ void group_traverse(const DevTree<A, K>& t, const Small<A, K>& q,
                                               const Small<typename Ax<A>::O, K>& qw, V& vis,
                                               Counters& c, int gl) {
  using O = typename Ax<A>::O;
  if (t.n == 0) return;
  const int D = t.nlev;
  Small<O, K> off;
  off.fill(O(0));
  O reg[GroupStack<G>::R];
#pragma unroll
  for (int j = 0; j < GroupStack<G>::R; ++j) reg[j] = O(0);
  uint32_t pending = 0, taken = 0;
  O rd = O(0);
  uint32_t idx = 1;
  int level = 0, dim = 0;
  bool check = !vis.unbounded();
  for (;;) {
    c.stems += uint32_t(D - level);
    while (level < D) {
      const A pivot = stem_at(t, idx);
      if (pivot < Ax<A>::maxv()) {
        const bool right = q[dim] >= pivot;
        const O no = abs_diff(qw[dim], O(pivot));
        const O oo = off[dim];
        const O rdf = add_rn(sub_rn(rd, mul_rn(oo, oo)), mul_rn(no, no));
        if (!check || !(rdf > vis.bound())) {
          gs_write<O, G>(reg, gl, level, rdf);
          pending |= 1u << level;
        }
        idx = (idx << 1) | (right ? 1u : 0u);
      } else {
        idx <<= 1;
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
    bool entered = false;
    while (pending | taken) {
      const int l = 31 - __clz(pending | taken);
      const uint32_t bit = 1u << l;
      const int ld = l % K;
      if (taken & bit) {
        off.set(ld, gs_read<O, G>(reg, l));
        taken ^= bit;
        continue;
      }
      pending ^= bit;
      const O frd = gs_read<O, G>(reg, l);
      if (vis.prune(frd)) continue;
      const A pivot = stem_at(t, idx >> (D - l));
      c.stems += 1;
      const O no = abs_diff(qw[ld], O(pivot));
      gs_write<O, G>(reg, gl, l, off[ld]);
      taken |= bit;
      off.set(ld, no);
      rd = frd;
      idx = (idx >> (D - l - 1)) ^ 1u;
      level = l + 1;
      dim = level % K;
      check = true;
      entered = true;
      break;
    }
    if (!entered) return;
  }
}

// Group within collector (fill pass for dense outputs): the G lanes of a query scan a
// leaf G entries at a time, and the entries in range are written in entry order, lane j
// at the query's next free slot plus the number of hits in lanes below j, so consecutive
// lanes store consecutive results (coalesced) instead of one lane writing them one by one.
template <class O, int G> struct GroupWithinVis {
  O radius;
  bool excl;
  int gl;
  uint64_t count;  // results so far (group-uniform)
  uint32_t* out_items;
  O* out_d;
  uint64_t cap;
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl = 1u) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    const int base = int(threadIdx.x & 63u) - gl;  // first lane of this group in the wave
    for (uint32_t p0 = s; p0 < e; p0 += G) {
      const uint32_t p = p0 + uint32_t(gl);
      O d = O(0);
      bool hit = false;
      if (p < e) {
        d = point_dist(t, p, qw);
        hit = excl ? d < radius : d <= radius;
      }
      const uint64_t mine = (uint64_t(__ballot(hit)) >> base) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
      if (hit) {
        const uint64_t o = count + uint64_t(__popcll(mine & ((1ull << gl) - 1ull)));
        if (o < cap) {
          out_items[o] = t.items[p];
          out_d[o] = d;
        }
      }
      count += uint64_t(__popcll(mine));
    }
  }
};

// Group within fill with staged writes: the G lanes of a query write the hits of each step
// into a ring of kRing results in LDS (position = result index mod kRing), and the group
// stores the ring to global memory in whole segments that end on a multiple of 32 results
// (128 bytes of items), so the L2 receives complete lines from one wave at a time. Writing
// each step's few hits straight to global memory left lines partly written when the L2
// evicted them: 2.4 bytes of HBM writes per result byte on a dense workload. Only the
// query's first and last lines are partial. Every lane of the group must call leaf() and
// finish(); lanes of one group are in one wave, so a wave barrier orders the ring's stores
// before its loads.
template <class O, int G> struct GroupWithinStageVis {
  // (the flush leaves fewer than kAlign pending and comes when more than kRing - G are, so
  // at most kRing are ever pending). A ring of 32 in segments of 16 at 4 lanes (16 KiB of LDS
  // per block) filled the 10M-point mixture as fast as 8 lanes but read 3.4x the HBM bytes
  // (65.7 against 19.3 GB), and the autotuner, which prefers fewer lanes, then took it.
  static constexpr uint32_t kRing = 64;
  static constexpr uint32_t kAlign = 32;
  static_assert(kRing - G >= kAlign && kRing % kAlign == 0, "ring and segment sizes");
  O radius;
  bool excl;
  int gl;
  uint64_t count;    // results so far (group-uniform)
  uint64_t flushed;  // results already stored to global memory
  uint64_t base;     // the query's first result index in the output arrays (alignment)
  uint32_t* out_items;  // the query's segment
  O* out_d;
  uint64_t cap;
  uint32_t* ring_it;  // this group's ring (LDS)
  O* ring_d;
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  // stores results [flushed, end) (end = the last 32-result boundary up to count, or count
  // when final)
  __device__ __forceinline__ void flush(bool final) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint64_t end = count;
    if (!final) {
      const uint64_t a = ((base + count) / kAlign) * kAlign;
      end = a > base + flushed ? a - base : flushed;
    }
    for (uint64_t j = flushed + uint64_t(gl); j < end; j += G) {
      if (j < cap) {
        out_items[j] = ring_it[j % kRing];
        out_d[j] = ring_d[j % kRing];
      }
    }
    flushed = end;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl = 1u) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    const int first = int(threadIdx.x & 63u) - gl;  // first lane of this group in the wave
    for (uint32_t p0 = s; p0 < e; p0 += G) {
      const uint32_t p = p0 + uint32_t(gl);
      O d = O(0);
      bool hit = false;
      if (p < e) {
        d = point_dist(t, p, qw);
        hit = excl ? d < radius : d <= radius;
      }
      const uint64_t mine = (uint64_t(__ballot(hit)) >> first) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
      if (hit) {
        const uint64_t o = count + uint64_t(__popcll(mine & ((1ull << gl) - 1ull)));
        ring_it[o % kRing] = t.items[p];
        ring_d[o % kRing] = d;
      }
      count += uint64_t(__popcll(mine));
      if (count - flushed > kRing - G) flush(false);  // room for the next step's G hits
    }
  }
};

// Group nearest_one collector (nearest_one.rs:330-352, canonical rule as NearestOneVis):
// the G lanes of a query walk group_traverse together and measure G entries of a leaf per
// step; the step's smallest (distance, item) is found by a butterfly over the group and
// replaces the best when it is smaller in that order. The thread collector's result is the
// smallest (distance, item) of (+inf, 0) and every entry it measures, which does not depend
// on the order the entries are seen in, so both return the same.
template <class O, int G> struct GroupNearestOneVis {
  O best;
  uint32_t item;
  int gl;
  __device__ O bound() const { return best; }
  __device__ bool prune(O rd) const { return rd > best; }
  __device__ bool prune_at(O rd, O b) const { return rd > b; }
  __device__ bool unbounded() const { return true; }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    for (uint32_t p0 = s; p0 < e; p0 += G) {
      const uint32_t p = p0 + uint32_t(gl);
      O d = oinf<O>();
      uint32_t it = 0xFFFFFFFFu;
      if (p < e) {
        d = point_dist(t, p, qw);
        it = t.items[p];
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) {
        const O od = __shfl_xor(d, o, G);
        const uint32_t oi = __shfl_xor(it, o, G);
        if (od < d || (od == d && oi < it)) {
          d = od;
          it = oi;
        }
      }
      if (d < best || (d == best && it < item)) {
        best = d;
        item = it;
      }
    }
  }
};

// Group best_n_within collector (n <= G; best_n_within.rs:163-196): the G lanes of a query
// run its traversal together (the radius is the only bound, so they take one path) and
// split each leaf G entries at a time. Lane gl keeps slot gl of the n smallest items in
// range. Entries in range are offered in entry order, and an entry goes after kept entries
// of an equal item, as in the thread collector (BestNVis), so both keep the same entries.
template <class O, int G> struct GroupBestNVis {
  O sd;  // this lane's slot: distance and item
  uint32_t si;
  int n, cnt, gl;
  O radius;
  bool excl;
  uint32_t kit;  // the n-th item once n are kept (group-uniform)
  __device__ void init(int n_, O r, bool ex, int gl_) {
    sd = oinf<O>();
    si = 0xFFFFFFFFu;
    n = n_;
    cnt = 0;
    gl = gl_;
    radius = r;
    excl = ex;
    kit = 0xFFFFFFFFu;
  }
  __device__ O bound() const { return radius; }
  __device__ bool prune(O rd) const { return rd > radius || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  // every lane of the group calls it with the same candidate
  __device__ __forceinline__ void insert(O cd, uint32_t ci) {
    const bool lt = gl < n && ci < si;
    const O pd = __shfl_up(sd, 1, G);
    const uint32_t pi = __shfl_up(si, 1, G);
    int plt = __shfl_up(int(lt), 1, G);
    if (gl == 0) plt = 0;
    if (plt) {
      sd = pd;
      si = pi;
    } else if (lt) {
      sd = cd;
      si = ci;
    }
    cnt = cnt < n ? cnt + 1 : n;
    if (cnt >= n) kit = __shfl(si, n - 1, G);
  }
  template <class A, int K, class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf, uint32_t nl = 1u) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + nl];
    const int first = int(threadIdx.x & 63u) - gl;  // first lane of this group in the wave
    for (uint32_t p0 = s; p0 < e; p0 += G) {
      const uint32_t p = p0 + uint32_t(gl);
      O d = oinf<O>();
      uint32_t it = 0xFFFFFFFFu;
      bool ok = false;
      if (p < e) {
        d = point_dist(t, p, qw);
        it = t.items[p];
        ok = (excl ? d < radius : d <= radius) && (cnt < n || it < kit);
      }
      uint64_t mask = (uint64_t(__ballot(ok)) >> first) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
      while (mask) {
        const int src = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1ull;
        const O cd = __shfl(d, src, G);
        const uint32_t ci = __shfl(it, src, G);
        if (cnt >= n && !(ci < kit)) continue;
        insert(cd, ci);
      }
    }
  }
};

// Group k-NN collector (k <= G). SCREEN: f64 trees measured first in FP32 from the shadow
// against the certified bound of screen_threshold, exact FP64 only for survivors.
template <class A, int K, int G, bool SCREEN> struct GroupKnnVis {
  using O = typename Ax<A>::O;
  O sd;          // this lane's slot of the sorted top-k
  uint32_t si;
  int k, cnt, gl;
  O radius;
  bool excl;
  O kd;          // k-th (last kept) entry once full, group-uniform
  uint32_t kit;
  float qf[4];   // SCREEN: the query in FP32
  double eps_k;  // SCREEN: sqrt(K) * per-query coordinate-difference error bound
  float thr;     // SCREEN: FP32 threshold certified to admit every FP64 candidate
  uint32_t rescored;

  __device__ void init(int k_, O r, bool ex, int gl_) {
    sd = oinf<O>();
    si = 0xFFFFFFFFu;
    k = k_;
    cnt = 0;
    gl = gl_;
    radius = r;
    excl = ex;
    kd = oinf<O>();
    kit = 0xFFFFFFFFu;
    thr = __builtin_huge_valf();
    rescored = 0;
  }
  __device__ O bound() const { return (cnt >= k && kd < radius) ? kd : radius; }
  __device__ bool prune(O rd) const { return rd > bound() || (excl && !(rd < radius)); }
  __device__ bool prune_at(O rd, O b) const { return rd > b || (excl && !(rd < radius)); }
  __device__ bool unbounded() const { return radius == oinf<O>(); }
  static __device__ __forceinline__ bool less(O ad, uint32_t ai, O bd, uint32_t bi) {
    return ad < bd || (ad == bd && ai < bi);
  }
  __device__ __forceinline__ void insert(O cd, uint32_t ci) {
    const bool lt = gl < k && less(cd, ci, sd, si);
    const O pd = __shfl_up(sd, 1, G);
    const uint32_t pi = __shfl_up(si, 1, G);
    // every lane of the group must execute the shuffle; lane 0 then drops what it read
    int plt = __shfl_up(int(lt), 1, G);
    if (gl == 0) plt = 0;
    if (gl < k) {
      if (plt) {
        sd = pd;
        si = pi;
      } else if (lt) {
        sd = cd;
        si = ci;
      }
    }
    cnt = cnt < k ? cnt + 1 : k;
    if (cnt >= k) {
      kd = __shfl(sd, k - 1, G);
      kit = __shfl(si, k - 1, G);
      if (SCREEN) thr = screen_threshold(double(bound()), eps_k);
    }
  }
  template <class QW>
  __device__ void leaf(const DevTree<A, K>& t, const QW& qw, uint32_t leaf) {
    const uint32_t s = t.leaf_start[leaf], e = t.leaf_start[leaf + 1];
    const int base = (threadIdx.x & 63) & ~(G - 1);
    for (uint32_t c0 = s; c0 < e; c0 += G) {
      const uint32_t p = c0 + uint32_t(gl);
      O d = oinf<O>();
      uint32_t it = 0xFFFFFFFFu;
      bool ok = false;
      if (p < e) {
        if constexpr (SCREEN) {
          const float4 sh = t.shadow[p];
          const float acc = shadow_sq<K>(sh, qf);
          if (!(acc > thr)) {
            ++rescored;
            d = point_dist(t, p, qw);
            it = K <= 3 ? __float_as_uint(sh.w) : t.items[p];  // K = 4: w is a coordinate
            ok = true;
          }
        } else {
          d = point_dist(t, p, qw);
          ok = true;
        }
        ok = ok && (excl ? d < radius : d <= radius) && (cnt < k || d <= kd);
        if (ok) {
          if constexpr (!SCREEN) it = t.items[p];
          ok = cnt < k || less(d, it, kd, kit);
        }
      }
      // the group's bits of the 64-lane ballot (all 64 when G = 64)
      uint64_t mask = (uint64_t(__ballot(ok)) >> base) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
      while (mask) {
        const int src = __ffsll((unsigned long long)mask) - 1;
        mask &= mask - 1ull;
        const O cd = __shfl(d, src, G);
        const uint32_t ci = __shfl(it, src, G);
        if (cnt >= k && !less(cd, ci, kd, kit)) continue;
        insert(cd, ci);
      }
    }
  }
};

// Sorts a segment of at most N entries in registers (Batcher's network over N slots, the
// slots past len padded with (+inf, ~0u), which sort last): one read and one write of each
// entry.
template <class O, int N>
__device__ __forceinline__ void sort_segment_regs(uint32_t* __restrict__ it, O* __restrict__ d, uint32_t len) {
  TopK<O, N> r;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    const uint32_t src = uint32_t(j) < len ? uint32_t(j) : 0u;  // in the segment, len >= 1
    const O dv = d[src];
    const uint32_t iv = it[src];
    r.d[j] = uint32_t(j) < len ? dv : oinf<O>();
    r.it[j] = uint32_t(j) < len ? iv : 0xFFFFFFFFu;
  }
  r.sort();
#pragma unroll
  for (int j = 0; j < N; ++j)
    if (uint32_t(j) < len) {
      d[j] = r.d[j];
      it[j] = r.it[j];
    }
}

// Sorts one query's segment by (distance, item), in place in global memory: insertion sort
// up to 48 entries, heapsort above. Few registers, for use inside a query kernel.
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

// sort_segment with 5 to 32 entries sorted in registers (one read and one write of each
// entry; insertion sort in global memory moves entries many times: on 250M results in
// segments of 15.6 on average the sort wrote 17.2 GB of HBM for 3 GB of results, 3.8 GB with
// 5 to 32 entries in registers). Takes 141 VGPRs for f64, so it runs in a kernel of its own
// (k_sort_segments).
template <class O>
__device__ void sort_segment_any(uint32_t* __restrict__ it, O* __restrict__ d, uint64_t len) {
  if (len > 4 && len <= 8) {
    sort_segment_regs<O, 8>(it, d, uint32_t(len));
  } else if (len > 8 && len <= 16) {
    sort_segment_regs<O, 16>(it, d, uint32_t(len));
  } else if (len > 16 && len <= 32) {
    sort_segment_regs<O, 32>(it, d, uint32_t(len));
  } else {
    sort_segment(it, d, len);
  }
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


}  // namespace dev
}  // namespace kd
