// host_build.cpp — parallel host builder producing Kiddo's ImmutableKdTree layout.
//
// Stands in for the Rust side's `ImmutableKdTree::new_from_slice` /
// `new_from_entries` (kiddo src/kd_tree/construction.rs:278-303, :403-411), which cannot be
// compiled in this toolchain. The output is the exact byte layout a Rust FFI shim would
// read out of a real tree through `KdTreeAccessor` (src/traits/kd_tree.rs:16-46):
//   * stems: Eytzinger order, index 0 unused, root at 1, children 2i / 2i+1, splitting
//     dimension = level mod K, unused slots hold A::max_value(); length is 2^(d+1)
//     rounded up to a multiple of 8 (construction.rs:454-469), d = log2(next_pow2(ceil(n/B)))
//   * leaves: 2^d leaves in left-to-right order (Soft bucket limit, so some may be empty
//     and some may exceed B), each an 8-byte aligned arena of tiles of 32/8/4/2/1 entries,
//     every tile laid out as K coordinate runs followed by the u32 items
//     (src/leaf_strategy/vec_of_arenas.rs:292-323, src/leaf_view/mod.rs:18-41).
//
// Split rule (construction.rs:673-691, :1089-1175): a stem with leaf budget b sends b/2
// leaves left; the ideal pivot is ceil(len * (b/2) / b) clamped to [1, len]; the pivot
// value v is the value of that rank; the split moves to the start of v's run of equal
// values, or, when that run starts the slice, past its end; an all-equal slice is not
// split (the stem keeps the sentinel and everything goes left). So every value left of a
// stem is < stem <= every value right of it, and leaf membership depends on the values
// only. The order of entries inside a leaf follows this builder's partitioning (Kiddo's
// follows Rust std's select_nth_unstable_by); the query engine never depends on it.
//
// Implementation: one index array, std::nth_element + partitions per stem, subtrees built
// on separate threads below the top levels, leaf arenas encoded in parallel afterwards.

#include "kiddo_amd_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace kd {

namespace {

template <class A, int K> struct SoftBuilder {
  const A* pts;
  A* stems;
  int max_level;            // d - 1
  uint64_t* leaf_begin;     // per leaf: first position in idx
  uint64_t* leaf_len;
  std::atomic<int>* threads_left;

  A val(uint64_t i, int d) const { return pts[i * K + d]; }

  uint64_t split(uint64_t* idx, uint64_t len, int dim, uint64_t init) {
    auto cmp = [&](uint64_t a, uint64_t b) { return val(a, dim) < val(b, dim); };
    std::nth_element(idx, idx + init, idx + len, cmp);
    const A v = val(idx[init], dim);
    // [0, init) holds values <= v and (init, len) values >= v.
    uint64_t lt = std::partition(idx, idx + init, [&](uint64_t i) { return val(i, dim) < v; }) - idx;
    if (lt > 0) return lt;
    // every value in [0, init] equals v: move past the end of v's run
    uint64_t eq_right =
        std::partition(idx + init + 1, idx + len, [&](uint64_t i) { return !(v < val(i, dim)); }) -
        (idx + init + 1);
    uint64_t le = init + 1 + eq_right;
    if (le < len) {
      // the stem is read at the pivot position: bring the smallest value above v there
      uint64_t* mn = std::min_element(idx + le, idx + len, cmp);
      std::swap(*mn, idx[le]);
    }
    return le;  // == len when the slice is unsplittable on this dimension
  }

  void run(uint64_t stem_idx, int level, uint64_t budget, uint64_t* idx, uint64_t begin,
           uint64_t len) {
    if (budget == 0) return;
    if (level > max_level) {
      uint64_t leaf = stem_idx - (uint64_t(1) << level);
      leaf_begin[leaf] = begin;
      leaf_len[leaf] = len;
      return;
    }
    const int dim = level % K;
    uint64_t left_budget, right_budget, pivot;
    if (budget == 1) {
      left_budget = 1;
      right_budget = 0;
      pivot = len;
    } else {
      left_budget = uint64_t(1) << (63 - __builtin_clzll(budget - 1));
      right_budget = budget - left_budget;
      if (len == 0) {
        pivot = 0;
      } else {
        pivot = (len * left_budget + budget - 1) / budget;
        pivot = std::min<uint64_t>(std::max<uint64_t>(pivot, 1), len);
      }
      if (pivot < len) pivot = split(idx + begin, len, dim, pivot);
    }
    if (pivot < len) stems[stem_idx] = val(idx[begin + pivot], dim);
    const uint64_t sp = std::min(pivot, len);
    // Fork the left subtree onto a new thread while big slices remain.
    bool fork = len > (uint64_t(1) << 16) && threads_left->fetch_sub(1) > 0;
    if (fork) {
      std::thread th([&, sp] { run(stem_idx * 2, level + 1, left_budget, idx, begin, sp); });
      run(stem_idx * 2 + 1, level + 1, right_budget, idx, begin + sp, len - sp);
      th.join();
    } else {
      if (len > (uint64_t(1) << 16)) threads_left->fetch_add(1);
      run(stem_idx * 2, level + 1, left_budget, idx, begin, sp);
      run(stem_idx * 2 + 1, level + 1, right_budget, idx, begin + sp, len - sp);
    }
  }
};

int tile_width(uint64_t remaining) {
  if (remaining >= 32) return 32;
  if (remaining >= 8) return 8;
  if (remaining >= 4) return 4;
  if (remaining >= 2) return 2;
  return 1;
}

template <class A, int K>
void encode_leaf(uint8_t* dst, const A* pts, const uint32_t* items, const uint64_t* idx,
                 uint64_t len) {
  uint64_t base = 0;
  while (base < len) {
    const int tl = tile_width(len - base);
    for (int d = 0; d < K; ++d)
      for (int i = 0; i < tl; ++i) {
        A v = pts[idx[base + i] * K + d];
        std::memcpy(dst, &v, sizeof(A));
        dst += sizeof(A);
      }
    for (int i = 0; i < tl; ++i) {
      uint32_t it = items ? items[idx[base + i]] : uint32_t(idx[base + i]);
      std::memcpy(dst, &it, 4);
      dst += 4;
    }
    base += tl;
  }
}

template <class F> void parallel_chunks(uint64_t count, int threads, F&& f) {
  if (threads <= 1 || count < 4096) {
    f(0, count);
    return;
  }
  std::vector<std::thread> th;
  uint64_t chunk = (count + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    uint64_t s = t * chunk, e = std::min(count, s + chunk);
    if (s >= e) break;
    th.emplace_back([&f, s, e] { f(s, e); });
  }
  for (auto& x : th) x.join();
}

template <class A, int K>
int build_typed(HostLayout& L, const A* pts, const uint32_t* items, uint64_t n, int B,
                int threads) {
  L.dims = K;
  L.bucket = B;
  L.size = n;
  const uint64_t leaf_node_count = (n + B - 1) / B;
  std::vector<uint64_t> idx(n);
  parallel_chunks(n, threads, [&](uint64_t s, uint64_t e) {
    for (uint64_t i = s; i < e; ++i) idx[i] = i;
  });
  std::vector<uint64_t> leaf_begin, leaf_len;
  if (leaf_node_count < 2) {
    // construction.rs:540-588: a single leaf and no stems
    L.max_stem_level = -1;
    L.stems.clear();
    L.stems_len = 0;
    if (n > 0) {
      leaf_begin.push_back(0);
      leaf_len.push_back(n);
    }
  } else {
    int d = 0;
    while ((uint64_t(1) << d) < leaf_node_count) ++d;
    if (d > kMaxStemLevels) return set_error(KD_ERR_TOO_DEEP, "tree depth %d exceeds %d", d, kMaxStemLevels);
    const uint64_t budget = uint64_t(1) << d;
    const uint64_t stem_count = uint64_t(1) << (d + 1);
    L.stems_len = (stem_count + 7) / 8 * 8;
    L.stems.assign(L.stems_len * sizeof(A), 0);
    A* stems = reinterpret_cast<A*>(L.stems.data());
    const A sentinel = AxisHost<A>::max_value();
    for (uint64_t i = 0; i < L.stems_len; ++i) stems[i] = sentinel;
    L.max_stem_level = d - 1;
    leaf_begin.assign(budget, 0);
    leaf_len.assign(budget, 0);
    std::atomic<int> threads_left{std::max(threads - 1, 0)};
    SoftBuilder<A, K> b{pts, stems, d - 1, leaf_begin.data(), leaf_len.data(), &threads_left};
    b.run(1, 0, budget, idx.data(), 0, n);
  }
  const uint64_t leaves = leaf_len.size();
  L.extents.assign(2 * leaves, 0);
  uint64_t off = 0, max_len = 0;
  for (uint64_t l = 0; l < leaves; ++l) {
    off = (off + 7) & ~uint64_t(7);
    L.extents[2 * l] = off;
    L.extents[2 * l + 1] = leaf_len[l];
    off += leaf_len[l] * (K * sizeof(A) + 4);
    max_len = std::max(max_len, leaf_len[l]);
  }
  L.max_leaf_len = max_len;
  if (max_len > kMaxLeafLen)
    return set_error(KD_ERR_LEAF_OVERFLOW,
                     "a leaf holds %llu entries; Kiddo's limit is %d (too many equal points)",
                     (unsigned long long)max_len, kMaxLeafLen);
  L.arena.assign(off, 0);
  parallel_chunks(leaves, threads, [&](uint64_t s, uint64_t e) {
    for (uint64_t l = s; l < e; ++l)
      encode_leaf<A, K>(L.arena.data() + L.extents[2 * l], pts, items, idx.data() + leaf_begin[l],
                        leaf_len[l]);
  });
  return KD_OK;
}

template <class A, int K> int check_points(const A* pts, uint64_t n) {
  for (uint64_t i = 0; i < n * K; ++i) {
    A v = pts[i];
    if (AxisHost<A>::is_nan(v))
      return set_error(KD_ERR_INVALID, "coordinate %llu is NaN", (unsigned long long)i);
    if (v == AxisHost<A>::max_value())
      return set_error(KD_ERR_INVALID,
                       "coordinate %llu equals the type's maximum, which Kiddo reserves as the "
                       "no-pivot sentinel", (unsigned long long)i);
  }
  return KD_OK;
}

}  // namespace

int build_layout(HostLayout& L, int dtype, int dims, int bucket, uint64_t n, const void* points,
                 const uint32_t* items, int threads) {
  if (bucket < 1) return set_error(KD_ERR_INVALID, "bucket size must be >= 1");
  if (n > 0 && !points) return set_error(KD_ERR_INVALID, "points is null");
  if (!items && n > 0xFFFFFFFFull)
    return set_error(KD_ERR_INVALID, "auto-generated u32 item index would overflow");
  if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  L.dtype = dtype;
  int rc = KD_ERR_INVALID;
  KD_DISPATCH_AK(dtype, dims, {
    rc = check_points<A, KK>(static_cast<const A*>(points), n);
    if (rc == KD_OK) rc = build_typed<A, KK>(L, static_cast<const A*>(points), items, n, bucket, threads);
  });
  return rc;
}

}  // namespace kd
