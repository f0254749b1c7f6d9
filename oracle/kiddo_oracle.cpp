// kiddo_oracle.cpp — CPU restatement of Kiddo's ImmutableKdTree builder and scalar query
// path (crate kiddo 6.0.0-alpha.3, default features: no `simd`, no FMA).
//
// TEST INFRASTRUCTURE ONLY. This file is the checker for the GPU engine in kiddo_amd/:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. The
// product never links, imports or calls it.
//
// What it restates (file:line relative to the kiddo crate root):
//   * builder          src/kd_tree/construction.rs:413-539 (new_from_source_with),
//                      :641-691 (write_leaf_from_sort_index, soft budgets), :868-968
//                      (populate_recursive_soft), :1089-1175 (update_pivot)
//   * stem layout      src/stem_strategy/eytzinger/mod.rs:82-174 (root 1, child 2i+r,
//                      dim cycles per level, leaf_idx = idx & !(1<<level))
//   * leaf arena       src/leaf_strategy/vec_of_arenas.rs:292-323 (8-byte aligned leaf
//                      start, tiles 32/8/4/2/1), src/leaf_view/mod.rs:18-41,431-505
//   * traversal        src/kd_tree/orchestrator/mod.rs:388-603 (arithmetic DFS with a
//                      restore stack and a far stack), :133-151 + traits/stem_strategy.rs
//                      :182-200 (straight descent for approx)
//   * metric           src/dist/squared_euclidean/mod.rs:36-39 (dist1 = d*d),
//                      src/dist/distance_metric_core.rs:38-40,89-98 (acc += c from 0;
//                      rd' = (rd - old^2) + new^2), src/traits/axis.rs:83-137,283-344
//   * families         query/nearest_one.rs:111-143,321-352; nearest_n.rs:21-38;
//                      nearest_n_within.rs:127-179,280-321,441-479; within.rs:18-41;
//                      within_unsorted.rs:133-156,267-286; best_n_within.rs:163-196,
//                      290-306; approx_nearest_one.rs:49-71
//   * collectors       src/results/result_collection.rs:147-215 (ThresholdVec),
//                      :463-561 (BinaryHeap of BestQueryResultItem), :1269-1317 (Vec)
//   * leaf kernels     src/leaf_view_chunked/{nearest_one,nearest_n_within,
//                      best_n_within}/fallback.rs
//
// Two collector modes:
//   canonical (mode 0): results ordered by (distance, then lower item); equal-distance
//                       subtrees are never pruned, so the whole boundary tie group is seen.
//   faithful  (mode 1): Kiddo's own rules (distance-only ordering, strict '<' for
//                       nearest_one and for a full k-NN collector, prune-on-equal for
//                       nearest_one). Tie choices then depend on within-leaf order, which
//                       Kiddo takes from Rust std's select_nth_unstable_by and which no
//                       restatement can reproduce; this mode exists for divergence counts.
//
// Output type O: A for f32/f64 trees; f64 for u16/u32 trees (every value is an exact
// integer below 2^53 there, so f64 results equal exact integer arithmetic).
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off: every product and sum is rounded
// separately, matching Kiddo's scalar path).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

enum { DT_F32 = 0, DT_F64 = 1, DT_U16 = 2, DT_U32 = 3 };
constexpr int kMaxLeafLen = 1024;  // src/leaf_view/mod.rs:18, asserted construction.rs:526

template <class A> struct Ax;
template <> struct Ax<float> {
  using O = float;
  static float maxv() { return std::numeric_limits<float>::infinity(); }
};
template <> struct Ax<double> {
  using O = double;
  static double maxv() { return std::numeric_limits<double>::infinity(); }
};
template <> struct Ax<uint16_t> {
  using O = double;
  static uint16_t maxv() { return 0xFFFFu; }
};
template <> struct Ax<uint32_t> {
  using O = double;
  static uint32_t maxv() { return 0xFFFFFFFFu; }
};

template <class O> O out_max() { return std::numeric_limits<O>::infinity(); }

// Kiddo's pop-time rule: prune if rd > bound, or rd == bound when prune-on-equal is set
// (Axis::cmp treats NaN as Equal, axis.rs:108-120).
template <class O> inline bool kiddo_prune(O rd, O bound, bool prune_on_equal) {
  bool greater = rd > bound;
  bool equal = !(rd < bound) && !greater;
  return greater || (prune_on_equal && equal);
}

struct Layout {
  int dtype = 0, K = 0, B = 0;
  uint64_t n = 0;
  int max_stem_level = -1;
  std::vector<uint8_t> stems;      // raw A values, Eytzinger order
  uint64_t stems_len = 0;          // number of A values
  std::vector<uint64_t> extents;   // (byte offset, len) per leaf
  std::vector<uint8_t> arena;      // VecOfArenas leaf_bytes
  uint64_t max_leaf_len = 0;
  int pad_levels = 0;              // sentinel levels above the real root (block layouts)
};

// ---------------------------------------------------------------- builder ---------

int tile_len(uint64_t remaining) {
  // src/leaf_view/mod.rs:22-30
  static const int widths[5] = {32, 8, 4, 2, 1};
  for (int w : widths)
    if (remaining >= (uint64_t)w) return w;
  return 1;
}

template <class A, int K>
void append_leaf(Layout& L, const A* pts, const uint32_t* items, const uint64_t* idx,
                 uint64_t len) {
  // src/leaf_strategy/vec_of_arenas.rs:292-323: start aligned to max(align A, align T, 8)
  uint64_t start = (L.arena.size() + 7) & ~uint64_t(7);
  L.arena.resize(start, 0);
  L.extents.push_back(start);
  L.extents.push_back(len);
  uint64_t base = 0;
  while (base < len) {
    int tl = tile_len(len - base);
    for (int d = 0; d < K; ++d)
      for (int i = 0; i < tl; ++i) {
        A v = pts[idx[base + i] * K + d];
        const uint8_t* b = reinterpret_cast<const uint8_t*>(&v);
        L.arena.insert(L.arena.end(), b, b + sizeof(A));
      }
    for (int i = 0; i < tl; ++i) {
      uint32_t it = items ? items[idx[base + i]] : (uint32_t)idx[base + i];
      const uint8_t* b = reinterpret_cast<const uint8_t*>(&it);
      L.arena.insert(L.arena.end(), b, b + 4);
    }
    base += tl;
  }
  L.max_leaf_len = std::max<uint64_t>(L.max_leaf_len, len);
}

template <class A, int K> struct Builder {
  const A* pts;
  const uint32_t* items;
  Layout& L;
  A* stems;
  int max_level;
  std::vector<A> scratch;

  A val(uint64_t i, int d) const { return pts[i * K + d]; }

  // construction.rs:1089-1175. Kiddo selects the init_pivot-th value v, then moves the
  // pivot left to the start of v's run of equal values; if that reaches 0 it moves right
  // past the end of the run instead, and an all-equal slice stays unsplit (pivot = len).
  // The restatement computes those positions from counts and partitions stably.
  uint64_t update_pivot(uint64_t* idx, uint64_t len, int dim, uint64_t init) {
    scratch.resize(len);
    for (uint64_t i = 0; i < len; ++i) scratch[i] = val(idx[i], dim);
    std::nth_element(scratch.begin(), scratch.begin() + init, scratch.end());
    A v = scratch[init];
    uint64_t lt = 0, le = 0;
    for (uint64_t i = 0; i < len; ++i) {
      A x = val(idx[i], dim);
      lt += (x < v);
      le += (x <= v);
    }
    uint64_t pivot;
    if (lt > 0) {
      pivot = lt;
      std::stable_partition(idx, idx + len, [&](uint64_t i) { return val(i, dim) < v; });
      stem_value = v;  // the run of v starts the right side
    } else if (le < len) {
      pivot = le;
      std::stable_partition(idx, idx + len, [&](uint64_t i) { return val(i, dim) <= v; });
      // Kiddo sorted the slice here, so its stem is the smallest value above v
      stem_value = val(idx[pivot], dim);
      for (uint64_t i = pivot; i < len; ++i) stem_value = std::min(stem_value, val(idx[i], dim));
    } else {
      pivot = len;
    }
    return pivot;
  }
  A stem_value{};

  // construction.rs:868-968
  void populate(uint64_t stem_idx, int level, uint64_t budget, uint64_t* idx, uint64_t len) {
    if (budget == 0) return;
    if (level > max_level) {
      append_leaf<A, K>(L, pts, items, idx, len);
      return;
    }
    int dim = level % K;
    uint64_t left_budget, right_budget, pivot;
    if (budget == 1) {
      left_budget = 1;
      right_budget = 0;
      pivot = len;
    } else {
      // construction.rs:673-691
      int lz = __builtin_clzll(budget - 1);
      left_budget = uint64_t(1) << (63 - lz);
      right_budget = budget - left_budget;
      if (len == 0) {
        pivot = 0;
      } else {
        pivot = (len * left_budget + budget - 1) / budget;
        pivot = std::min<uint64_t>(std::max<uint64_t>(pivot, 1), len);
      }
      if (pivot < len) pivot = update_pivot(idx, len, dim, pivot);
    }
    // the stem is the smallest value of the right side (Kiddo reads it at sort_index[pivot]
    // after select_nth / sort, construction.rs:927-933)
    if (pivot < len) stems[stem_idx] = stem_value;
    uint64_t split = std::min(pivot, len);
    populate(stem_idx * 2, level + 1, left_budget, idx, split);
    populate(stem_idx * 2 + 1, level + 1, right_budget, idx + split, len - split);
  }
};

// block > 1: the stem depth is padded up to a multiple of `block` as Kiddo does for
// block-based stem strategies (construction.rs:431-444): the padding levels sit above the
// real root, hold no pivot (sentinel), and are pre-traversed to the left, so the real root
// is Eytzinger node 2^pad at level pad and splits dimension pad % K; the soft leaf budget
// is 2^(padded depth) (construction.rs:446-450). The leaves under the padding levels'
// right children are empty and follow the real ones.
template <class A, int K>
int build_layout(Layout& L, const A* pts, const uint32_t* items, uint64_t n, int B, int block = 1) {
  L.K = K;
  L.B = B;
  L.n = n;
  L.extents.clear();
  L.arena.clear();
  L.max_leaf_len = 0;
  uint64_t leaf_node_count = (n + B - 1) / B;
  std::vector<uint64_t> idx(n);
  for (uint64_t i = 0; i < n; ++i) idx[i] = i;
  if (leaf_node_count < 2) {
    // construction.rs:540-588: one leaf, no stems
    L.max_stem_level = -1;
    L.stems_len = 0;
    L.stems.clear();
    if (n > 0) append_leaf<A, K>(L, pts, items, idx.data(), n);
  } else {
    int d = 0;
    while ((uint64_t(1) << d) < leaf_node_count) ++d;  // log2(next_pow2)
    const int pad = (block > 1 && d % block) ? block - d % block : 0;
    d += pad;
    if (d > 40) return -3;
    uint64_t budget = uint64_t(1) << d;
    uint64_t stem_count = uint64_t(1) << (d + 1);  // rightmost leaf's stem index + 1
    uint64_t padded = (stem_count + 7) / 8 * 8;     // construction.rs:465-469
    L.stems_len = padded;
    L.stems.assign(padded * sizeof(A), 0);
    A* stems = reinterpret_cast<A*>(L.stems.data());
    for (uint64_t i = 0; i < padded; ++i) stems[i] = Ax<A>::maxv();
    L.max_stem_level = d - 1;
    L.pad_levels = pad;
    Builder<A, K> b{pts, items, L, stems, d - 1, {}};
    b.populate(uint64_t(1) << pad, pad, budget, idx.data(), n);
    // the empty leaves under the padding levels' right children
    const uint64_t real = uint64_t(1) << (d - pad);
    for (uint64_t l = real; l < budget; ++l) append_leaf<A, K>(L, pts, items, idx.data(), 0);
  }
  if (L.max_leaf_len > (uint64_t)kMaxLeafLen) return -2;  // construction.rs:526
  return 0;
}

// ------------------------------------------------------------- query engine ---------

struct Stats {
  uint64_t leaves = 0, points = 0, stems = 0;
};

// Distance metrics (src/dist/*/mod.rs): dist1 of a difference or offset, the combine of
// per-axis terms from 0 in axis order, and the rect distance after one axis offset changes.
//   0 SquaredEuclidean  d*d, sum, (rd - old^2) + new^2     (squared_euclidean/mod.rs:36-39)
//   1 Manhattan         |d|, sum, (rd - old) + new         (manhattan/mod.rs:36-49)
//   2 Chebyshev         |d|, max, rescan of the offsets    (chebyshev/mod.rs:34-61)
//   3 Minkowski<P>      |d|^P, sum, (rd - old^P) + new^P   (minkowski/mod.rs:55-59; powi
//                       restated as binary exponentiation, the usual expansion of llvm.powi)
template <class O> struct Metric {
  int kind = 0, p = 0;
  O d1(O x) const {
    switch (kind) {
      case 1:
      case 2:
        return std::fabs(x);
      case 3: {
        O a = std::fabs(x), r = O(1);
        for (int b = p;;) {
          if (b & 1) r = r * a;
          b >>= 1;
          if (!b) break;
          a = a * a;
        }
        return r;
      }
      default:
        return x * x;
    }
  }
  O combine(O acc, O c) const {
    if (kind == 2) return c > acc ? c : acc;
    return acc + c;
  }
  template <int K> O rect(O rd, const O* off, int dim, O no) const {
    switch (kind) {
      case 1:
        return (rd - off[dim]) + no;
      case 2: {
        O acc = O(0);
        for (int a = 0; a < K; ++a) acc = combine(acc, a == dim ? no : off[a]);
        return acc;
      }
      case 3:
        return (rd - d1(off[dim])) + d1(no);
      default:
        return (rd - off[dim] * off[dim]) + no * no;
    }
  }
};

template <class A, int K> struct View {
  using O = typename Ax<A>::O;
  const A* stems;
  const uint64_t* ext;
  uint64_t leaf_count;
  const uint8_t* arena;
  uint64_t n;
  int max_stem_level;
  const O* pbox = nullptr;  // periodic box (widened); null: Cartesian
  Metric<O> met{};

  // src/leaf_view/mod.rs:431-446 tile walk; calls f(O dist, uint32_t item) per entry
  // with dist computed as ((0 + dist1(x)) + dist1(y)) + ... in O, no FMA.
  template <class F> void scan_leaf(uint64_t leaf, const O* qw, F&& f) const {
    uint64_t off = ext[2 * leaf], len = ext[2 * leaf + 1];
    const uint8_t* p = arena + off;
    uint64_t base = 0;
    while (base < len) {
      int tl = tile_len(len - base);
      for (int i = 0; i < tl; ++i) {
        O acc = O(0);
        for (int d = 0; d < K; ++d) {
          A c;
          std::memcpy(&c, p + (size_t(d) * tl + i) * sizeof(A), sizeof(A));
          O diff;
          if (pbox) {
            // minimum-image axis delta (query/periodic.rs:282-292, :322-336):
            // delta = |q - c|, wrapped = L - delta, the smaller of the two
            O delta = std::fabs(qw[d] - O(c));
            O wrapped = pbox[d] - delta;
            diff = wrapped < delta ? wrapped : delta;
          } else {
            diff = O(c) - qw[d];
          }
          acc = met.combine(acc, met.d1(diff));
        }
        uint32_t it;
        std::memcpy(&it, p + size_t(K) * tl * sizeof(A) + size_t(i) * 4, 4);
        f(acc, it);
      }
      p += size_t(K) * tl * sizeof(A) + size_t(tl) * 4;
      base += tl;
    }
  }

  // orchestrator/mod.rs:133-151 + traits/stem_strategy.rs:182-200
  uint64_t straight_leaf(const A* q, Stats& st) const {
    uint64_t idx = 1;
    int level = 0, dim = 0;
    while (level <= max_stem_level) {
      A pivot = stems[idx];
      ++st.stems;
      bool right = q[dim] >= pivot;
      idx = idx * 2 + (right ? 1 : 0);
      ++level;
      dim = (dim + 1 == K) ? 0 : dim + 1;
    }
    return idx & ~(uint64_t(1) << level);
  }

  // orchestrator/mod.rs:388-603. Ctx provides max_dist() (descent check), prune(rd)
  // (re-check of a far frame when it is popped), unbounded() and leaf(leaf_idx, qw, st).
  template <class Ctx> void dfs(const A* q, Ctx& ctx, Stats& st) const {
    if (n == 0) return;
    O qw[K];
    for (int d = 0; d < K; ++d) qw[d] = O(q[d]);
    O off[K];
    for (int d = 0; d < K; ++d) off[d] = O(0);
    struct Restore { O old_off; bool has_far; };
    struct Far { uint64_t idx; int level; int dim; O far_off; O rd; };
    // fixed-capacity stacks, like Kiddo's inline continuation stacks (query_stack.rs:5-7)
    struct RStack {
      Restore v[72]; int n = 0;
      void push_back(Restore r) { v[n++] = r; }
      Restore back() const { return v[n - 1]; }
      void pop_back() { --n; }
      bool empty() const { return n == 0; }
    } rs;
    struct FStack {
      Far v[72]; int n = 0;
      void push_back(Far f) { v[n++] = f; }
      Far back() const { return v[n - 1]; }
      void pop_back() { --n; }
    } fs;
    uint64_t idx = 1;
    int level = 0, sdim = 0;
    O rd = O(0);

    auto descend = [&](bool check_bound) -> uint64_t {
      while (level <= max_stem_level) {
        int dim = sdim;
        A pivot = stems[idx];
        ++st.stems;
        int ndim = (dim + 1 == K) ? 0 : dim + 1;
        if (pivot < Ax<A>::maxv()) {
          bool right = q[dim] >= pivot;
          uint64_t near_idx = idx * 2 + (right ? 1 : 0);
          uint64_t far_idx = idx * 2 + (right ? 0 : 1);
          O pw = O(pivot);
          O new_off = qw[dim] >= pw ? qw[dim] - pw : pw - qw[dim];  // |a-b| (axis.rs:124)
          O old_off = off[dim];
          O rd_far = met.template rect<K>(rd, off, dim, new_off);
          bool cand = check_bound ? !(rd_far > ctx.max_dist()) : true;
          if (cand) {
            rs.push_back({old_off, true});
            fs.push_back({far_idx, level + 1, ndim, new_off, rd_far});
          } else {
            rs.push_back({old_off, false});
          }
          idx = near_idx;
        } else {
          rs.push_back({off[dim], false});
          idx = idx * 2;
        }
        ++level;
        sdim = ndim;
      }
      return idx & ~(uint64_t(1) << level);
    };

    uint64_t leaf = descend(!ctx.unbounded());
    ++st.leaves;
    ctx.leaf(leaf, qw, st);
    int dim = sdim;
    for (;;) {
      bool entered = false;
      while (!rs.empty()) {
        Restore fr = rs.back();
        rs.pop_back();
        int restore_dim = dim == 0 ? K - 1 : dim - 1;
        off[restore_dim] = fr.old_off;
        dim = restore_dim;
        if (!fr.has_far) continue;
        Far far = fs.back();
        fs.pop_back();
        if (ctx.prune(far.rd)) continue;  // orchestrator/mod.rs:472-483
        rs.push_back({fr.old_off, false});
        off[restore_dim] = far.far_off;
        idx = far.idx;
        level = far.level;
        sdim = far.dim;
        rd = far.rd;
        leaf = descend(true);
        ++st.leaves;
        ctx.leaf(leaf, qw, st);
        dim = sdim;
        entered = true;
        break;
      }
      if (!entered) break;
    }
  }

  // Periodic backtracking (query/periodic.rs:439-618). Every frame carries the node's box
  // (lower/upper per axis, starting at [0, L]), its per-axis periodic interval offsets and
  // rect distance. At a stem both children get rd = (rd - off[dim]^2) + child_off^2 with
  // child_off the distance from the query to the child's interval or its +-L images
  // (:294-317); the far child is pushed unless pruned, a pruned near child pops the stack,
  // and a popped frame is not re-checked. Requires pbox.
  static O iv1(O q, O lo, O hi) {  // interval_distance_1d, stem_strategy/donnelly/simd_full/mod.rs:209-216
    O below = std::max(O(0), lo - q);
    O above = std::max(O(0), q - hi);
    return below + above;
  }
  static O piv(O q, O lo, O hi, O L) {  // periodic_interval_distance_1d
    O best = iv1(q, lo, hi);
    O minus = iv1(q, lo - L, hi - L);
    if (minus < best) best = minus;
    O plus = iv1(q, lo + L, hi + L);
    if (plus < best) best = plus;
    return best;
  }
  template <class Ctx> void periodic_dfs(const A* q, Ctx& ctx, Stats& st) const {
    if (n == 0) return;
    struct Frame {
      uint64_t idx;
      int level;
      O lower[K], upper[K], off[K];
      O rd;
    };
    O qw[K];
    Frame cur;
    for (int d = 0; d < K; ++d) {
      qw[d] = O(q[d]);
      cur.lower[d] = O(0);
      cur.upper[d] = pbox[d];
      cur.off[d] = O(0);
    }
    cur.rd = O(0);
    cur.idx = 1;
    cur.level = 0;
    std::vector<Frame> stack;
    stack.reserve(size_t(max_stem_level) + 2);
    for (;;) {
      for (;;) {  // descend until a leaf is scanned (periodic.rs:495-606)
        if (cur.level > max_stem_level) {
          uint64_t leaf = cur.idx & ~(uint64_t(1) << cur.level);
          ++st.leaves;
          ctx.leaf(leaf, qw, st);
          break;
        }
        const int dim = cur.level % K;
        A pivot = stems[cur.idx];
        ++st.stems;
        if (!(pivot < Ax<A>::maxv())) {  // sentinel: left child, nothing changes
          cur.idx *= 2;
          ++cur.level;
          continue;
        }
        const O pw = O(pivot), ol = cur.lower[dim], ou = cur.upper[dim];
        const bool right = q[dim] >= pivot;
        const O r_lo = std::max(ol, pw), r_hi = ou;
        const O l_lo = ol, l_hi = ou < pw ? ou : pw;
        const O n_lo = right ? r_lo : l_lo, n_hi = right ? r_hi : l_hi;
        const O f_lo = right ? l_lo : r_lo, f_hi = right ? l_hi : r_hi;
        const O n_off = piv(qw[dim], n_lo, n_hi, pbox[dim]);
        const O f_off = piv(qw[dim], f_lo, f_hi, pbox[dim]);
        const O n_rd = met.template rect<K>(cur.rd, cur.off, dim, n_off);
        const O f_rd = met.template rect<K>(cur.rd, cur.off, dim, f_off);
        const bool n_pruned = ctx.prune(n_rd), f_pruned = ctx.prune(f_rd);
        if (!f_pruned) {
          Frame f = cur;
          f.idx = cur.idx * 2 + (right ? 0 : 1);
          f.level = cur.level + 1;
          f.lower[dim] = f_lo;
          f.upper[dim] = f_hi;
          f.off[dim] = f_off;
          f.rd = f_rd;
          stack.push_back(f);
        }
        if (n_pruned) {
          if (stack.empty()) return;
          cur = stack.back();
          stack.pop_back();
          continue;
        }
        cur.idx = cur.idx * 2 + (right ? 1 : 0);
        ++cur.level;
        cur.lower[dim] = n_lo;
        cur.upper[dim] = n_hi;
        cur.off[dim] = n_off;
        cur.rd = n_rd;
      }
      if (stack.empty()) break;
      cur = stack.back();
      stack.pop_back();
    }
  }
};

// ---- collectors / contexts

template <class O> struct Res {
  O d;
  uint32_t it;
};
template <class O> inline bool canon_less(const Res<O>& a, const Res<O>& b) {
  return a.d < b.d || (a.d == b.d && a.it < b.it);
}

template <class A, int K> struct NearestOneCtx {
  using O = typename Ax<A>::O;
  const View<A, K>& t;
  bool canonical;
  O best = out_max<O>();
  uint32_t item = 0;  // T::default()
  O max_dist() const { return best; }
  bool unbounded() const { return true; }
  // nearest_one.rs:347-351 prunes on equal; canonical keeps equal-distance subtrees
  bool prune(O rd) const { return kiddo_prune(rd, best, !canonical); }
  void leaf(uint64_t l, const O* qw, Stats& st) {
    t.scan_leaf(l, qw, [&](O d, uint32_t it) {
      ++st.points;
      if (canonical) {
        if (d < best || (d == best && it < item)) { best = d; item = it; }
      } else if (d < best) {  // nearest_one/fallback.rs:98
        best = d;
        item = it;
      }
    });
  }
};

template <class A, int K> struct NearestNCtx {
  using O = typename Ax<A>::O;
  const View<A, K>& t;
  bool canonical;
  uint64_t k;
  O radius;
  bool exclusive;
  std::vector<Res<O>> v;  // faithful: ThresholdVec; canonical: sorted by (d, item)
  bool full() const { return v.size() >= k; }
  O max_dist() const {  // nearest_n_within.rs:461-468
    if (full() && v.back().d < radius) return v.back().d;
    return radius;
  }
  bool unbounded() const { return radius == out_max<O>(); }
  bool prune(O rd) const {
    if (!canonical) return kiddo_prune(rd, max_dist(), exclusive);
    // canonical: an equal-distance subtree may hold a lower item, so only the exclusive
    // radius prunes on equality
    return rd > max_dist() || (exclusive && !(rd < radius));
  }
  void add_faithful(Res<O> e) {  // result_collection.rs:171-189
    if (v.size() < k) {
      v.push_back(e);
      if (v.size() == k)
        std::stable_sort(v.begin(), v.end(), [](const Res<O>& a, const Res<O>& b) { return a.d < b.d; });
    } else if (e.d < v[k - 1].d) {
      size_t i = k - 1;
      while (i > 0 && v[i - 1].d > e.d) { v[i] = v[i - 1]; --i; }
      v[i] = e;
    }
  }
  void add_canonical(Res<O> e) {
    if (v.size() < k) {
      v.insert(std::upper_bound(v.begin(), v.end(), e, canon_less<O>), e);
    } else if (canon_less(e, v[k - 1])) {
      v.pop_back();
      v.insert(std::upper_bound(v.begin(), v.end(), e, canon_less<O>), e);
    }
  }
  void leaf(uint64_t l, const O* qw, Stats& st) {
    O leaf_max = max_dist();  // nearest_n_within.rs:307-315: bound taken at leaf start
    t.scan_leaf(l, qw, [&](O d, uint32_t it) {
      ++st.points;
      if (canonical) {
        bool ok = exclusive ? d < radius : d <= radius;
        if (ok) add_canonical({d, it});
      } else {
        bool ok = exclusive ? d < leaf_max : d <= leaf_max;
        if (ok) add_faithful({d, it});
      }
    });
  }
};

template <class A, int K> struct WithinCtx {
  using O = typename Ax<A>::O;
  const View<A, K>& t;
  O radius;
  bool exclusive;
  std::vector<Res<O>>* out;  // nullptr: count only
  uint64_t count = 0;
  O max_dist() const { return radius; }
  bool unbounded() const { return radius == out_max<O>(); }
  bool prune(O rd) const { return kiddo_prune(rd, radius, exclusive); }
  void leaf(uint64_t l, const O* qw, Stats& st) {
    t.scan_leaf(l, qw, [&](O d, uint32_t it) {
      ++st.points;
      bool ok = exclusive ? d < radius : d <= radius;
      if (ok) {
        ++count;
        if (out) out->push_back({d, it});
      }
    });
  }
};

template <class A, int K> struct BestNCtx {
  using O = typename Ax<A>::O;
  const View<A, K>& t;
  uint64_t nmax;
  O radius;
  bool exclusive;
  std::vector<Res<O>> heap;  // max-heap by item (best_query_result_item.rs:27-31)
  O max_dist() const { return radius; }
  bool unbounded() const { return radius == out_max<O>(); }
  bool prune(O rd) const { return kiddo_prune(rd, radius, exclusive); }
  static bool item_less(const Res<O>& a, const Res<O>& b) { return a.it < b.it; }
  void leaf(uint64_t l, const O* qw, Stats& st) {
    bool have_thr = heap.size() >= nmax;  // best_n_within.rs: threshold_item snapshot
    uint32_t thr = have_thr ? heap.front().it : 0;
    t.scan_leaf(l, qw, [&](O d, uint32_t it) {
      ++st.points;
      bool ok = exclusive ? d < radius : d <= radius;
      if (!ok) return;
      if (have_thr && it >= thr) return;
      if (heap.size() < nmax) {  // result_collection.rs:481-500
        heap.push_back({d, it});
        std::push_heap(heap.begin(), heap.end(), item_less);
      } else if (it < heap.front().it) {
        std::pop_heap(heap.begin(), heap.end(), item_less);
        heap.back() = {d, it};
        std::push_heap(heap.begin(), heap.end(), item_less);
      }
    });
  }
};

// ------------------------------------------------------------------ helpers ---------

template <class F> void parallel_for(uint64_t m, int threads, F&& f) {
  if (threads <= 1 || m < 256) {
    for (uint64_t i = 0; i < m; ++i) f(i, 0);
    return;
  }
  std::atomic<uint64_t> next{0};
  std::vector<std::thread> th;
  for (int t = 0; t < threads; ++t)
    th.emplace_back([&, t] {
      for (;;) {
        uint64_t s = next.fetch_add(1024);
        if (s >= m) break;
        uint64_t e = std::min<uint64_t>(m, s + 1024);
        for (uint64_t i = s; i < e; ++i) f(i, t);
      }
    });
  for (auto& x : th) x.join();
}

struct StatsAcc {
  std::vector<Stats> per;
  explicit StatsAcc(int threads) : per(std::max(threads, 1)) {}
  void total(uint64_t* out) const {
    if (!out) return;
    Stats s;
    for (auto& p : per) { s.leaves += p.leaves; s.points += p.points; s.stems += p.stems; }
    out[0] = s.leaves;
    out[1] = s.points;
    out[2] = s.stems;
  }
};

}  // namespace

// ==================================================================== C ABI =========

struct or_tree {
  Layout L;
  int metric_kind = 0, metric_p = 0;  // metric of the following queries (or_set_metric)
};

extern "C" {

#define DISPATCH(dtype, K, ...)                                              \
  switch ((dtype) * 8 + (K)) {                                                \
    case DT_F32 * 8 + 2: { using A = float; constexpr int KK = 2; __VA_ARGS__; }     \
    case DT_F32 * 8 + 3: { using A = float; constexpr int KK = 3; __VA_ARGS__; }     \
    case DT_F32 * 8 + 4: { using A = float; constexpr int KK = 4; __VA_ARGS__; }     \
    case DT_F64 * 8 + 2: { using A = double; constexpr int KK = 2; __VA_ARGS__; }    \
    case DT_F64 * 8 + 3: { using A = double; constexpr int KK = 3; __VA_ARGS__; }    \
    case DT_F64 * 8 + 4: { using A = double; constexpr int KK = 4; __VA_ARGS__; }    \
    case DT_U16 * 8 + 2: { using A = uint16_t; constexpr int KK = 2; __VA_ARGS__; }  \
    case DT_U16 * 8 + 3: { using A = uint16_t; constexpr int KK = 3; __VA_ARGS__; }  \
    case DT_U16 * 8 + 4: { using A = uint16_t; constexpr int KK = 4; __VA_ARGS__; }  \
    case DT_U32 * 8 + 2: { using A = uint32_t; constexpr int KK = 2; __VA_ARGS__; }  \
    case DT_U32 * 8 + 3: { using A = uint32_t; constexpr int KK = 3; __VA_ARGS__; }  \
    case DT_U32 * 8 + 4: { using A = uint32_t; constexpr int KK = 4; __VA_ARGS__; }  \
    default: return -1;                                                       \
  }

// Builds a tree from n points (row-major n x K array of the dtype) exactly as
// ImmutableKdTree::new_from_slice (items = source index) or new_from_entries (items given).
int or_build(int dtype, int K, int B, uint64_t n, const void* pts, const uint32_t* items,
             or_tree** out) {
  if (B < 1 || !out) return -1;
  auto* t = new or_tree;
  t->L.dtype = dtype;
  int rc = 0;
  DISPATCH(dtype, K, rc = (build_layout<A, KK>(t->L, static_cast<const A*>(pts), items, n, B)); break);
  if (rc != 0) {
    delete t;
    return rc;
  }
  *out = t;
  return 0;
}

// As or_build, with the stem depth padded to a multiple of `block` levels as for a
// block-based stem strategy (Donnelly<block>); the tree is held in Eytzinger order.
int or_build_blocked(int dtype, int K, int B, uint64_t n, const void* pts, const uint32_t* items,
                     int block, or_tree** out) {
  if (B < 1 || block < 1 || block > 16 || !out) return -1;
  auto* t = new or_tree;
  t->L.dtype = dtype;
  int rc = 0;
  DISPATCH(dtype, K, rc = (build_layout<A, KK>(t->L, static_cast<const A*>(pts), items, n, B, block)); break);
  if (rc != 0) {
    delete t;
    return rc;
  }
  *out = t;
  return 0;
}

int or_pad_levels(const or_tree* t) { return t->L.pad_levels; }

// Donnelly<BH> stem order (src/stem_strategy/donnelly/core.rs): stems are laid out in
// minor triangles of BH levels, one per line of 2^BH slots. One step down
// (step_pure, core.rs:254-289): within a triangle the children of slot i are 2i + 1 and
// 2i + 2 of the line (offset by the line base); from the bottom row of a triangle (minor
// level BH - 1) the child is the root of the child triangle, whose line starts at
// (line base + 1 + 2 * column + is_right) << BH, column = slot - minor level - 1.
// All arithmetic wraps in u32, as in the reference.
struct DonPos {
  uint32_t idx, minor;
};
static DonPos donnelly_step(uint32_t BH, DonPos p, bool right) {
  const uint32_t r = right ? 1u : 0u;
  const uint32_t mask = (1u << BH) - 1u;
  const uint32_t min_idx = p.idx & mask;
  const uint32_t col = min_idx - p.minor - 1u;
  const uint32_t base = (p.idx & ~mask) + 1u;
  const uint32_t same = base + r + (min_idx << 1);
  const uint32_t next = ((base + (col << 1)) << BH) + (r << BH);
  const bool major = p.minor + 1u == BH;
  return DonPos{major ? next : same, major ? 0u : p.minor + 1u};
}
// traverse_block (core.rs:161-200): from a triangle root (minor level 0), `block_size`
// levels at once to child `child` (0 .. 2^block_size - 1) of the block.
static uint32_t donnelly_block_step(uint32_t BH, uint32_t idx, uint32_t child, uint32_t block_size) {
  const uint32_t mask = (1u << BH) - 1u;
  const uint32_t off = (1u << BH) - (block_size << 1) - 1u;
  return ((idx & ~mask) + off + child) << BH;
}

// Stem index (and minor level) reached by following `n` choices from the root.
int or_donnelly_path(int BH, const uint8_t* bits, uint64_t n, uint32_t* out2) {
  if (BH < 1 || BH > 16) return -1;
  DonPos p{0, 0};
  for (uint64_t i = 0; i < n; ++i) p = donnelly_step(uint32_t(BH), p, bits[i] != 0);
  out2[0] = p.idx;
  out2[1] = p.minor;
  return 0;
}
// Stem index reached by `n` block steps of `block_size` levels (children in [0, 2^bs)).
int or_donnelly_block_path(int BH, const uint8_t* children, uint64_t n, int block_size, uint32_t* out) {
  if (BH < 1 || BH > 16 || block_size < 1) return -1;
  uint32_t idx = 0;
  for (uint64_t i = 0; i < n; ++i) idx = donnelly_block_step(uint32_t(BH), idx, children[i], uint32_t(block_size));
  *out = idx;
  return 0;
}

// The stems of a padded tree (or_build_blocked with block = BH) in Donnelly<BH> order,
// sized as Kiddo sizes them (construction.rs:441-469: from the real root's position,
// follow the right-most path of stem-depth levels; that index + 1, rounded up to a
// multiple of 8), unused slots holding the sentinel. out == null: only *out_len is set.
int or_donnelly_stems(const or_tree* t, int BH, void* out, uint64_t* out_len) {
  const Layout& L = t->L;
  if (L.max_stem_level < 0) {
    *out_len = 0;
    return 0;
  }
  const int D = L.max_stem_level + 1, pad = L.pad_levels;
  if (BH < 3 || BH > 16 || D % BH != 0 || pad >= BH) return -2;  // BH < 3: the child rule wraps
  DonPos root{0, 0};
  for (int i = 0; i < pad; ++i) root = donnelly_step(uint32_t(BH), root, false);
  DonPos far = root;
  for (int i = 0; i < D; ++i) far = donnelly_step(uint32_t(BH), far, true);
  const uint64_t len = (uint64_t(far.idx) + 1 + 7) / 8 * 8;
  *out_len = len;
  if (!out) return 0;
  DISPATCH(L.dtype, L.K, {
    (void)KK;
    A* dst = static_cast<A*>(out);
    const A* src = reinterpret_cast<const A*>(L.stems.data());
    for (uint64_t i = 0; i < len; ++i) dst[i] = Ax<A>::maxv();
    // depth-first over the real subtree: Eytzinger node e at level l <-> Donnelly slot
    struct Frame { uint64_t e; int level; DonPos p; };
    std::vector<Frame> stack{{uint64_t(1) << pad, pad, root}};
    while (!stack.empty()) {
      const Frame f = stack.back();
      stack.pop_back();
      if (f.p.idx >= len) return -3;
      dst[f.p.idx] = src[f.e];
      if (f.level + 1 < D) {
        stack.push_back({2 * f.e, f.level + 1, donnelly_step(uint32_t(BH), f.p, false)});
        stack.push_back({2 * f.e + 1, f.level + 1, donnelly_step(uint32_t(BH), f.p, true)});
      }
    }
    return 0;
  });
}

// Wraps an existing Kiddo-layout tree (for example one made by another builder).
int or_from_layout(int dtype, int K, int B, uint64_t n, int max_stem_level, const void* stems,
                   uint64_t stems_len, const uint64_t* extents, uint64_t leaf_count,
                   const uint8_t* arena, uint64_t arena_len, or_tree** out) {
  size_t asz = dtype == DT_F32 ? 4 : dtype == DT_F64 ? 8 : dtype == DT_U16 ? 2 : dtype == DT_U32 ? 4 : 0;
  if (!asz || K < 2 || K > 4 || !out) return -1;
  auto* t = new or_tree;
  Layout& L = t->L;
  L.dtype = dtype; L.K = K; L.B = B; L.n = n; L.max_stem_level = max_stem_level;
  L.stems_len = stems_len;
  L.stems.assign((const uint8_t*)stems, (const uint8_t*)stems + stems_len * asz);
  L.extents.assign(extents, extents + 2 * leaf_count);
  L.arena.assign(arena, arena + arena_len);
  for (uint64_t i = 0; i < leaf_count; ++i) L.max_leaf_len = std::max(L.max_leaf_len, extents[2 * i + 1]);
  *out = t;
  return 0;
}

void or_free(or_tree* t) { delete t; }

// The metric functions themselves, for the reference's metric tests (src/dist/mod.rs):
// the distance of two K-vectors (f64) and the rect distance after one offset changes.
int or_metric_dist(uint32_t flags, const double* a, const double* b, int K, double* out) {
  Metric<double> m;
  m.kind = int((flags >> 8) & 0xF);
  m.p = int((flags >> 12) & 0xFF);
  double acc = 0.0;
  for (int d = 0; d < K; ++d) acc = m.combine(acc, m.d1(a[d] - b[d]));
  *out = acc;
  return 0;
}
int or_metric_rect(uint32_t flags, double rd, const double* off, int K, int dim, double no, double* out) {
  Metric<double> m;
  m.kind = int((flags >> 8) & 0xF);
  m.p = int((flags >> 12) & 0xFF);
  switch (K) {
    case 2: *out = m.rect<2>(rd, off, dim, no); return 0;
    case 3: *out = m.rect<3>(rd, off, dim, no); return 0;
    case 4: *out = m.rect<4>(rd, off, dim, no); return 0;
    default: return -1;
  }
}

// Metric of the following queries on this tree: flags bits 8-11 kind (0 SquaredEuclidean,
// 1 Manhattan, 2 Chebyshev, 3 Minkowski), bits 12-19 the Minkowski exponent.
int or_set_metric(or_tree* t, uint32_t flags) {
  const int kind = int((flags >> 8) & 0xF), p = int((flags >> 12) & 0xFF);
  if (kind > 3 || (kind == 3 && (p == 1 || p == 2))) return -2;
  t->metric_kind = kind;
  t->metric_p = p;
  return 0;
}

// update_pivot on its own (construction.rs:1089-1175), for its unit test
// (construction.rs:1187-1213). Returns the pivot; sort_index is reordered in place.
int64_t or_update_pivot(int dtype, int K, const void* pts, uint64_t n, uint64_t* sort_index,
                        int dim, uint64_t init) {
  Layout L;
  DISPATCH(dtype, K, {
    Builder<A, KK> b{static_cast<const A*>(pts), nullptr, L, nullptr, 0, {}};
    return (int64_t)b.update_pivot(sort_index, n, dim, init);
  });
}

// Layout accessors (pointers stay valid until or_free).
int or_info(const or_tree* t, int64_t* out6) {
  // n, max_stem_level, stems_len, leaf_count, arena_len, max_leaf_len
  out6[0] = (int64_t)t->L.n;
  out6[1] = t->L.max_stem_level;
  out6[2] = (int64_t)t->L.stems_len;
  out6[3] = (int64_t)(t->L.extents.size() / 2);
  out6[4] = (int64_t)t->L.arena.size();
  out6[5] = (int64_t)t->L.max_leaf_len;
  return 0;
}
const void* or_stems(const or_tree* t) { return t->L.stems.data(); }
const uint64_t* or_extents(const or_tree* t) { return t->L.extents.data(); }
const uint8_t* or_arena(const or_tree* t) { return t->L.arena.data(); }

}  // extern "C"

namespace {
template <class A, int K> View<A, K> view_of(const or_tree* t) {
  const Layout& L = t->L;
  View<A, K> v{reinterpret_cast<const A*>(L.stems.data()), L.extents.data(),
               L.extents.size() / 2, L.arena.data(), L.n, L.max_stem_level};
  v.met.kind = t->metric_kind;
  v.met.p = t->metric_p;
  return v;
}
}  // namespace

extern "C" {

// mode: 0 canonical, 1 faithful. dists: O per query. stats3: leaves, points, stems.
int or_nearest_one(const or_tree* t, const void* q, uint64_t m, int mode, uint32_t* items,
                   void* dists, int threads, uint64_t* stats3) {
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      NearestOneCtx<A, KK> c{v, mode == 0};
      v.dfs(qs + i * KK, c, acc.per[th]);
      items[i] = c.item;
      od[i] = c.best;
    });
    acc.total(stats3);
    return 0;
  });
}

// approx_nearest_one.rs:49-71 (requires Output == A, so only float trees).
int or_approx_nearest_one(const or_tree* t, const void* q, uint64_t m, int mode, uint32_t* items,
                          void* dists, int threads) {
  if (t->L.dtype != DT_F32 && t->L.dtype != DT_F64) return -3;
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      O qw[KK];
      for (int d = 0; d < KK; ++d) qw[d] = O(qs[i * KK + d]);
      O best = out_max<O>();
      uint32_t item = 0;
      if (v.n > 0) {
        uint64_t leaf = v.straight_leaf(qs + i * KK, acc.per[th]);
        v.scan_leaf(leaf, qw, [&](O d, uint32_t it) {
          if (d < best || (mode == 0 && d == best && it < item)) { best = d; item = it; }
        });
      }
      items[i] = item;
      od[i] = best;
    });
    return 0;
  });
}

// nearest_n (radius = +inf) and nearest_n().within(radius). Results row-major m x k;
// rows shorter than k are padded with item 0xFFFFFFFF and distance +inf.
int or_nearest_n(const or_tree* t, const void* q, uint64_t m, uint64_t k, double radius,
                 int exclusive, int sorted, int mode, uint32_t* items, void* dists,
                 uint32_t* counts, int threads, uint64_t* stats3) {
  if (k == 0) return -1;
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      NearestNCtx<A, KK> c{v, mode == 0, k, O(radius), exclusive != 0, {}};
      c.v.reserve(k + 1);
      v.dfs(qs + i * KK, c, acc.per[th]);
      auto& r = c.v;
      if (mode != 0 && sorted && r.size() < k)  // into_sorted_vec: sort when not full
        std::stable_sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.d < b.d; });
      counts[i] = (uint32_t)r.size();
      for (uint64_t j = 0; j < k; ++j) {
        items[i * k + j] = j < r.size() ? r[j].it : 0xFFFFFFFFu;
        od[i * k + j] = j < r.size() ? r[j].d : out_max<O>();
      }
    });
    acc.total(stats3);
    return 0;
  });
}

int or_within_count(const or_tree* t, const void* q, uint64_t m, double radius, int exclusive,
                    uint64_t* counts, int threads, uint64_t* stats3) {
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      WithinCtx<A, KK> c{v, O(radius), exclusive != 0, nullptr};
      v.dfs(qs + i * KK, c, acc.per[th]);
      counts[i] = c.count;
    });
    acc.total(stats3);
    return 0;
  });
}

// offsets: m+1 exclusive prefix sums of the counts. sorted: 0 visit order (unsorted),
// 1 sorted (canonical: by (dist, item); faithful: stable by distance, within.rs:18-41).
int or_within_fill(const or_tree* t, const void* q, uint64_t m, double radius, int exclusive,
                   int sorted, int mode, const uint64_t* offsets, uint32_t* items, void* dists,
                   int threads) {
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      std::vector<Res<O>> r;
      WithinCtx<A, KK> c{v, O(radius), exclusive != 0, &r};
      v.dfs(qs + i * KK, c, acc.per[th]);
      if (sorted) {
        if (mode == 0)
          std::sort(r.begin(), r.end(), canon_less<O>);
        else
          std::stable_sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.d < b.d; });
      }
      uint64_t o = offsets[i];
      for (size_t j = 0; j < r.size() && o + j < offsets[i + 1]; ++j) {
        items[o + j] = r[j].it;
        od[o + j] = r[j].d;
      }
    });
    return 0;
  });
}

// best_n_within: rows sorted by item ascending (Kiddo returns the BinaryHeap; its
// into_sorted_vec order is ascending item, best_n_within.rs tests).
int or_best_n_within(const or_tree* t, const void* q, uint64_t m, double radius, uint64_t nmax,
                     int exclusive, uint32_t* items, void* dists, uint32_t* counts, int threads,
                     uint64_t* stats3) {
  if (nmax == 0) return -1;
  StatsAcc acc(threads);
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      BestNCtx<A, KK> c{v, nmax, O(radius), exclusive != 0, {}};
      v.dfs(qs + i * KK, c, acc.per[th]);
      auto& r = c.heap;
      std::sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.it < b.it; });
      counts[i] = (uint32_t)r.size();
      for (uint64_t j = 0; j < nmax; ++j) {
        items[i * nmax + j] = j < r.size() ? r[j].it : 0xFFFFFFFFu;
        od[i * nmax + j] = j < r.size() ? r[j].d : out_max<O>();
      }
    });
    acc.total(stats3);
    return 0;
  });
}

// ------------------------------------------------------------ periodic ------------
// Periodic boundary conditions (query/periodic.rs, builder.rs:1280-1291): float trees only
// (PeriodicAxis is implemented for f32/f64), every box length strictly positive
// (builder.rs:29-35). box: K values of the tree's coordinate type.

}  // extern "C"

#define DISPATCH_FLOAT(dtype, K, ...)                                        \
  switch ((dtype) * 8 + (K)) {                                                \
    case DT_F32 * 8 + 2: { using A = float; constexpr int KK = 2; __VA_ARGS__; }     \
    case DT_F32 * 8 + 3: { using A = float; constexpr int KK = 3; __VA_ARGS__; }     \
    case DT_F32 * 8 + 4: { using A = float; constexpr int KK = 4; __VA_ARGS__; }     \
    case DT_F64 * 8 + 2: { using A = double; constexpr int KK = 2; __VA_ARGS__; }    \
    case DT_F64 * 8 + 3: { using A = double; constexpr int KK = 3; __VA_ARGS__; }    \
    case DT_F64 * 8 + 4: { using A = double; constexpr int KK = 4; __VA_ARGS__; }    \
    default: return -3;                                                       \
  }

namespace {
template <class A, int K> bool periodic_view(View<A, K>& v, const void* box, typename Ax<A>::O* bw) {
  const A* b = static_cast<const A*>(box);
  for (int d = 0; d < K; ++d) {
    if (!(b[d] > A(0))) return false;
    bw[d] = typename Ax<A>::O(b[d]);
  }
  v.pbox = bw;
  return true;
}
}  // namespace

extern "C" {

// nearest_one under periodic boundaries (periodic.rs:196-244): Cartesian nearest_one of the
// home query, then of each shifted image (every axis shifted by -L, 0 or +L, not all zero,
// enumerated axis 0 outermost, shift -1, 0, +1, :47-142) whose lower bound
// ((0 + o0^2) + o1^2) + ... (o = distance of the shifted coordinate outside [0, L], :27-37)
// passes the best distance so far, visited in ascending lower bound (stable: ties in
// enumeration order). Faithful: exclusive lower-bound checks and strict '<' updates, as
// Kiddo. Canonical: inclusive checks and (distance, item) order, so equal-distance images
// are seen.
int or_periodic_nearest_one(const or_tree* t, const void* q, uint64_t m, const void* box, int mode,
                            uint32_t* items, void* dists, int threads, uint64_t* stats3) {
  if (t->L.dtype != DT_F32 && t->L.dtype != DT_F64) return -3;
  StatsAcc acc(threads);
  DISPATCH_FLOAT(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    O bw[KK];
    if (!periodic_view(v, box, bw)) return -2;
    auto cv = view_of<A, KK>(t);  // Cartesian view for the image searches
    const A* qs = static_cast<const A*>(q);
    const A* bx = static_cast<const A*>(box);
    O* od = static_cast<O*>(dists);
    const bool canonical = mode == 0;
    parallel_for(m, threads, [&](uint64_t i, int th) {
      const A* qi = qs + i * KK;
      NearestOneCtx<A, KK> home{cv, canonical};
      cv.dfs(qi, home, acc.per[th]);
      O best = home.best;
      uint32_t item = home.item;
      struct Img { A wq[KK]; O lb; };
      std::vector<Img> cands;
      A wq[KK];
      for (int d = 0; d < KK; ++d) wq[d] = qi[d];
      auto accepts = [&](O lb) { return canonical ? !(lb > best) : lb < best; };
      // recursion of periodic_image_candidates (:55-124), threshold = home best
      std::function<void(int, bool, O)> rec = [&](int axis, bool shifted, O lb) {
        if (axis == KK) {
          if (shifted) {
            Img c;
            std::memcpy(c.wq, wq, sizeof(wq));
            c.lb = lb;
            cands.push_back(c);
          }
          return;
        }
        const A orig = qi[axis], L = bx[axis];
        for (int sh = -1; sh <= 1; ++sh) {
          A wc = sh < 0 ? A(orig - L) : sh == 0 ? orig : A(orig + L);
          wq[axis] = wc;
          A off = wc < A(0) ? A(std::fabs(wc - A(0))) : wc > L ? A(std::fabs(wc - L)) : A(0);
          O ow = O(off);
          O nlb = v.met.combine(lb, v.met.d1(ow));
          if (!accepts(nlb)) continue;
          rec(axis + 1, shifted || sh != 0, nlb);
        }
        wq[axis] = orig;
      };
      rec(0, false, O(0));
      std::stable_sort(cands.begin(), cands.end(), [](const Img& a, const Img& b) { return a.lb < b.lb; });
      for (const Img& c : cands) {
        if (!accepts(c.lb)) continue;
        NearestOneCtx<A, KK> img{cv, canonical};
        cv.dfs(c.wq, img, acc.per[th]);
        const bool better = canonical ? (img.best < best || (img.best == best && img.item < item))
                                      : img.best < best;
        if (better) {
          best = img.best;
          item = img.item;
        }
      }
      items[i] = item;
      od[i] = best;
    });
    acc.total(stats3);
    return 0;
  });
}

// periodic nearest_n (radius +inf) and nearest_n().within(radius) (periodic.rs:620-791).
int or_periodic_nearest_n(const or_tree* t, const void* q, uint64_t m, const void* box, uint64_t k,
                          double radius, int exclusive, int sorted, int mode, uint32_t* items,
                          void* dists, uint32_t* counts, int threads, uint64_t* stats3) {
  if (k == 0) return -1;
  if (t->L.dtype != DT_F32 && t->L.dtype != DT_F64) return -3;
  StatsAcc acc(threads);
  DISPATCH_FLOAT(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    O bw[KK];
    if (!periodic_view(v, box, bw)) return -2;
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      NearestNCtx<A, KK> c{v, mode == 0, k, O(radius), exclusive != 0, {}};
      c.v.reserve(k + 1);
      v.periodic_dfs(qs + i * KK, c, acc.per[th]);
      auto& r = c.v;
      if (mode != 0 && sorted && r.size() < k)
        std::stable_sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.d < b.d; });
      counts[i] = (uint32_t)r.size();
      for (uint64_t j = 0; j < k; ++j) {
        items[i * k + j] = j < r.size() ? r[j].it : 0xFFFFFFFFu;
        od[i * k + j] = j < r.size() ? r[j].d : out_max<O>();
      }
    });
    acc.total(stats3);
    return 0;
  });
}

int or_periodic_within_count(const or_tree* t, const void* q, uint64_t m, const void* box,
                             double radius, int exclusive, uint64_t* counts, int threads,
                             uint64_t* stats3) {
  if (t->L.dtype != DT_F32 && t->L.dtype != DT_F64) return -3;
  StatsAcc acc(threads);
  DISPATCH_FLOAT(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    O bw[KK];
    if (!periodic_view(v, box, bw)) return -2;
    const A* qs = static_cast<const A*>(q);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      WithinCtx<A, KK> c{v, O(radius), exclusive != 0, nullptr};
      v.periodic_dfs(qs + i * KK, c, acc.per[th]);
      counts[i] = c.count;
    });
    acc.total(stats3);
    return 0;
  });
}

int or_periodic_within_fill(const or_tree* t, const void* q, uint64_t m, const void* box,
                            double radius, int exclusive, int sorted, int mode,
                            const uint64_t* offsets, uint32_t* items, void* dists, int threads) {
  if (t->L.dtype != DT_F32 && t->L.dtype != DT_F64) return -3;
  StatsAcc acc(threads);
  DISPATCH_FLOAT(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    O bw[KK];
    if (!periodic_view(v, box, bw)) return -2;
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int th) {
      std::vector<Res<O>> r;
      WithinCtx<A, KK> c{v, O(radius), exclusive != 0, &r};
      v.periodic_dfs(qs + i * KK, c, acc.per[th]);
      if (sorted) {
        if (mode == 0)
          std::sort(r.begin(), r.end(), canon_less<O>);
        else
          std::stable_sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.d < b.d; });
      }
      uint64_t o = offsets[i];
      for (size_t j = 0; j < r.size() && o + j < offsets[i + 1]; ++j) {
        items[o + j] = r[j].it;
        od[o + j] = r[j].d;
      }
    });
    return 0;
  });
}

// ------------------------------------------------------------ brute force ---------
// Canonical answers by scanning every entry: the independent check of the traversal.

int or_brute_nearest_n(const or_tree* t, const void* q, uint64_t m, uint64_t k, double radius,
                       int exclusive, uint32_t* items, void* dists, uint32_t* counts, int threads) {
  if (k == 0) return -1;
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int) {
      O qw[KK];
      for (int d = 0; d < KK; ++d) qw[d] = O(qs[i * KK + d]);
      std::vector<Res<O>> r;
      O rad = O(radius);
      for (uint64_t l = 0; l < v.leaf_count; ++l)
        v.scan_leaf(l, qw, [&](O d, uint32_t it) {
          bool ok = exclusive ? d < rad : d <= rad;
          if (!ok) return;
          Res<O> e{d, it};
          if (r.size() < k) {
            r.insert(std::upper_bound(r.begin(), r.end(), e, canon_less<O>), e);
          } else if (canon_less(e, r.back())) {
            r.pop_back();
            r.insert(std::upper_bound(r.begin(), r.end(), e, canon_less<O>), e);
          }
        });
      counts[i] = (uint32_t)r.size();
      for (uint64_t j = 0; j < k; ++j) {
        items[i * k + j] = j < r.size() ? r[j].it : 0xFFFFFFFFu;
        od[i * k + j] = j < r.size() ? r[j].d : out_max<O>();
      }
    });
    return 0;
  });
}

int or_brute_within(const or_tree* t, const void* q, uint64_t m, double radius, int exclusive,
                    const uint64_t* offsets, uint64_t* counts, uint32_t* items, void* dists,
                    int threads) {
  // offsets == nullptr: count only. Otherwise fill sorted by (dist, item).
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int) {
      O qw[KK];
      for (int d = 0; d < KK; ++d) qw[d] = O(qs[i * KK + d]);
      std::vector<Res<O>> r;
      O rad = O(radius);
      for (uint64_t l = 0; l < v.leaf_count; ++l)
        v.scan_leaf(l, qw, [&](O d, uint32_t it) {
          bool ok = exclusive ? d < rad : d <= rad;
          if (ok) r.push_back({d, it});
        });
      if (counts) counts[i] = r.size();
      if (offsets) {
        std::sort(r.begin(), r.end(), canon_less<O>);
        for (size_t j = 0; j < r.size(); ++j) {
          items[offsets[i] + j] = r[j].it;
          od[offsets[i] + j] = r[j].d;
        }
      }
    });
    return 0;
  });
}

int or_brute_best_n_within(const or_tree* t, const void* q, uint64_t m, double radius,
                           uint64_t nmax, int exclusive, uint32_t* items, void* dists,
                           uint32_t* counts, int threads) {
  if (nmax == 0) return -1;
  DISPATCH(t->L.dtype, t->L.K, {
    using O = typename Ax<A>::O;
    auto v = view_of<A, KK>(t);
    const A* qs = static_cast<const A*>(q);
    O* od = static_cast<O*>(dists);
    parallel_for(m, threads, [&](uint64_t i, int) {
      O qw[KK];
      for (int d = 0; d < KK; ++d) qw[d] = O(qs[i * KK + d]);
      std::vector<Res<O>> r;
      O rad = O(radius);
      for (uint64_t l = 0; l < v.leaf_count; ++l)
        v.scan_leaf(l, qw, [&](O d, uint32_t it) {
          bool ok = exclusive ? d < rad : d <= rad;
          if (ok) r.push_back({d, it});
        });
      std::sort(r.begin(), r.end(), [](const Res<O>& a, const Res<O>& b) { return a.it < b.it; });
      if (r.size() > nmax) r.resize(nmax);
      counts[i] = (uint32_t)r.size();
      for (uint64_t j = 0; j < nmax; ++j) {
        items[i * nmax + j] = j < r.size() ? r[j].it : 0xFFFFFFFFu;
        od[i * nmax + j] = j < r.size() ? r[j].d : out_max<O>();
      }
    });
    return 0;
  });
}

}  // extern "C"
