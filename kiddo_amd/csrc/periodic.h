// periodic.h — device side of queries under periodic boundary conditions
// (kiddo src/kd_tree/query/periodic.rs; float trees only, every box length L > 0).
//
//   * leaf distance: per axis delta = |q - c|, the smaller of delta and L - delta, squared
//     and summed from 0 in axis order (:282-292, :322-336);
//   * nearest_n / within: Kiddo's periodic backtracking (:439-618). Each node carries its
//     box [lower, upper] per axis (the root's is [0, L]), the periodic interval offset of
//     the query per axis (the distance to the interval or to its images shifted by -L and
//     +L, :294-317) and rd = (rd - off[dim]^2) + off'^2. At a stem both children are
//     measured; the far one is queued unless pruned, a pruned near child pops the queue,
//     and a popped child is entered without a second check;
//   * nearest_one: Cartesian nearest_one of the home query, then of the shifted query
//     images in ascending lower bound (:47-244).
//
// As in the Cartesian traversal there is no stack in memory: a queued far child is a bit
// in `pending`, and a node's box, offsets and rd are a function of its root path alone,
// so after a pop they are rebuilt by replaying the path from the root with the same
// operations Kiddo applied when it pushed the frame (bit-identical values).
#pragma once

#include "device.h"

namespace kd {
namespace dev {

template <class O, int K, class M = SqEuclidean> struct PeriodicQW {
  Small<O, K> q;    // widened query
  Small<O, K> box;  // widened box lengths
  M met;
  __device__ O operator[](int d) const { return q[d]; }
};

// Minimum-image distance of entry p, in Kiddo's periodic leaf order:
// acc = combine(acc, dist1(m, 0)) per axis.
template <class A, int K, class O, class M>
__device__ __forceinline__ O point_dist(const DevTree<A, K>& t, uint32_t p, const PeriodicQW<O, K, M>& pq) {
  O acc = O(0);
#pragma unroll
  for (int d = 0; d < K; ++d) {
    const O c = O(t.coords[d * t.plane + p]);
    const O delta = fabs(sub_rn(pq.q[d], c));
    const O wrapped = sub_rn(pq.box[d], delta);
    const O m = wrapped < delta ? wrapped : delta;
    acc = pq.met.combine(acc, pq.met.d1(m));
  }
  return acc;
}

// interval_distance_1d (stem_strategy/donnelly/simd_full/mod.rs:209-216):
// max(0, lo - q) + max(0, q - hi)
template <class O> __device__ __forceinline__ O interval_dist(O q, O lo, O hi) {
  const O a = sub_rn(lo, q), b = sub_rn(q, hi);
  return add_rn(O(0) < a ? a : O(0), O(0) < b ? b : O(0));
}

// periodic_interval_distance_1d (periodic.rs:294-317): the interval, then its -L and +L
// images, keeping a later one only when strictly smaller
template <class O> __device__ __forceinline__ O periodic_interval_dist(O q, O lo, O hi, O L) {
  O best = interval_dist(q, lo, hi);
  const O minus = interval_dist(q, sub_rn(lo, L), sub_rn(hi, L));
  best = minus < best ? minus : best;
  const O plus = interval_dist(q, add_rn(lo, L), add_rn(hi, L));
  best = plus < best ? plus : best;
  return best;
}

// Periodic backtracking. V supplies prune(rd) (the family's rule, canonical: equal
// distances are only pruned by an exclusive radius) and leaf(tree, pq, leaf_index).
// Counters: one stem per descent step (sentinel stems included), leaves and entries scanned.
template <class A, int K, class V, class M>
__device__ __forceinline__ void periodic_traverse(const DevTree<A, K>& t, const Small<A, K>& q,
                                                  const PeriodicQW<typename Ax<A>::O, K, M>& pq, V& vis,
                                                  Counters& c) {
  using O = typename Ax<A>::O;
  if (t.n == 0) return;
  const int D = t.nlev;
  Small<O, K> lo, hi, off;
  lo.fill(O(0));
  hi = pq.box;
  off.fill(O(0));
  O rd = O(0);
  uint32_t idx = 1, pending = 0;
  int level = 0;
  // One level on dimension d (= level % K); false when the near child is pruned (then pop).
  auto descend = [&](const int d) -> bool {
    const A pivot = stem_at(t, idx);
    c.stems += 1;
    if (!(pivot < Ax<A>::maxv())) {  // sentinel stem: left child, state unchanged
      idx <<= 1;
      ++level;
      return true;
    }
    const O pw = O(pivot), ol = lo[d], ou = hi[d];
    const bool right = q[d] >= pivot;
    const O r_lo = ol < pw ? pw : ol;  // max(lower, pivot)
    const O l_hi = ou < pw ? ou : pw;  // min(upper, pivot)
    const O n_lo = right ? r_lo : ol, n_hi = right ? ou : l_hi;
    const O f_lo = right ? ol : r_lo, f_hi = right ? l_hi : ou;
    const O qd = pq.q[d], L = pq.box[d];
    const O n_off = periodic_interval_dist(qd, n_lo, n_hi, L);
    const O f_off = periodic_interval_dist(qd, f_lo, f_hi, L);
    const O n_rd = pq.met.rect(rd, off, d, n_off);
    const O f_rd = pq.met.rect(rd, off, d, f_off);
    if (!vis.prune(f_rd)) pending |= 1u << level;
    idx = (idx << 1) | (right ? 1u : 0u);
    ++level;
    if (vis.prune(n_rd)) return false;
    lo.set(d, n_lo);
    hi.set(d, n_hi);
    off.set(d, n_off);
    rd = n_rd;
    return true;
  };
  for (;;) {
    // levels one at a time up to a multiple of K (a run-time dimension), then groups of K
    // levels whose dimensions are constants, then the rest
    bool reached = true;
    while (reached && level < D && level % K != 0) reached = descend(level % K);
    while (reached && level + K <= D) {
#pragma unroll
      for (int d = 0; d < K; ++d)
        if (reached) reached = descend(d);
    }
#pragma unroll
    for (int d = 0; d < K - 1; ++d)
      if (reached && level < D) reached = descend(d);
    if (reached) {
      const uint32_t leaf = idx - (1u << D);
      c.leaves += 1;
      c.points += t.leaf_start[leaf + 1] - t.leaf_start[leaf];
      vis.leaf(t, pq, leaf, 1u);
    }
    if (!pending) return;
    // the deepest queued far child: the sibling of the current path's level-(l+1) node
    const int l = 31 - __clz(pending);
    pending &= ~(1u << l);
    idx = (idx >> (level - l - 1)) ^ 1u;
    level = l + 1;
    // rebuild the node's box, offsets and rd along its root path. Level j splits dimension
    // j % K and the path starts at level 0, so the levels go in groups of K with each
    // level's dimension a constant (a run-time dimension costs a select per element of
    // every vector it reads or writes).
    lo.fill(O(0));
    hi = pq.box;
    off.fill(O(0));
    rd = O(0);
    auto replay = [&](const int j, const int d) {
      const uint32_t node = idx >> (level - j);
      const bool side = ((idx >> (level - j - 1)) & 1u) != 0;
      const A pivot = stem_at(t, node);
      if (pivot < Ax<A>::maxv()) {
        const O pw = O(pivot), ol = lo[d], ou = hi[d];
        const O n_lo = side ? (ol < pw ? pw : ol) : ol;
        const O n_hi = side ? ou : (ou < pw ? ou : pw);
        const O n_off = periodic_interval_dist(pq.q[d], n_lo, n_hi, pq.box[d]);
        rd = pq.met.rect(rd, off, d, n_off);
        lo.set(d, n_lo);
        hi.set(d, n_hi);
        off.set(d, n_off);
      }
    };
    int j = 0;
    for (; j + K <= level; j += K) {
#pragma unroll
      for (int d = 0; d < K; ++d) replay(j + d, d);
    }
#pragma unroll
    for (int d = 0; d < K - 1; ++d)
      if (j + d < level) replay(j + d, d);
  }
}

}  // namespace dev
}  // namespace kd
