// k_build.hip — ImmutableKdTree construction on the GPU, from points already in HBM
// (kiddo src/kd_tree/construction.rs:413-539, :641-691, :868-968, :1089-1175).
//
// At every stem the leaf budget is a power of two and half of it goes left, so Kiddo's
// Soft split of a node with len entries sorted on the level's dimension (s[0..len)) is:
//   init = ceil(len / 2); len < 2: no split (the stem keeps the sentinel, all go left);
//   v = s[init]; lt = #{s < v}; lt > 0: split at lt, stem v;
//   else le = #{s <= v}; le < len: split at le, stem s[le]; else no split.
// Every value left of a stem is < stem <= every value right of it, so the stems and the
// leaf membership depend on the values only; this builder reproduces the host builder's
// (and the reference's) stems bit for bit and the same entries in every leaf. The order of
// entries inside a leaf differs (this one is sorted on the last split dimension), which no
// query observes. (A +0.0 / -0.0 tie at a split may leave a stem with the other zero sign,
// as the choice between equal values in the reference's select_nth_unstable_by is not
// specified either.)
//
// Level-synchronous: the positions of a level are grouped by node (contiguous, node order);
// a stable radix sort by the level's coordinate followed by a stable radix sort by node id
// leaves every node's positions sorted on that coordinate; one thread per node then finds
// its split by binary search and writes the stem, and each position learns its child.
// Cost per level: two device radix sorts of n pairs and three streaming passes.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "tree.h"

namespace kd {
namespace dev {

template <class A> struct BuildAx;
template <> struct BuildAx<float> {
  static __device__ bool bad(float v) { return v != v || v == __builtin_huge_valf(); }
};
template <> struct BuildAx<double> {
  static __device__ bool bad(double v) { return v != v || v == __builtin_huge_val(); }
};
template <> struct BuildAx<uint16_t> {
  static __device__ bool bad(uint16_t v) { return v == 0xFFFFu; }
};
template <> struct BuildAx<uint32_t> {
  static __device__ bool bad(uint32_t v) { return v == 0xFFFFFFFFu; }
};

// Per block: min, max and max |x| per dimension and a bad-value count (NaN or the
// type's maximum, which Kiddo reserves as the no-pivot sentinel). 13 doubles per block.
template <class A, int K>
__global__ __launch_bounds__(kBlock) void k_build_scan(const A* __restrict__ pts, uint64_t n,
                                                       double* __restrict__ partial) {
  __shared__ double red[kBlock];
  double lo[4], hi[4], ab[4], bad = 0.0;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    lo[d] = __builtin_huge_val();
    hi[d] = -__builtin_huge_val();
    ab[d] = 0.0;
  }
  for (uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
#pragma unroll
    for (int d = 0; d < K; ++d) {
      const A v = pts[i * K + d];
      if (BuildAx<A>::bad(v)) {
        bad += 1.0;
        continue;
      }
      const double x = double(v);
      lo[d] = fmin(lo[d], x);
      hi[d] = fmax(hi[d], x);
      ab[d] = fmax(ab[d], fabs(x));
    }
  }
  auto reduce = [&](double v, int op) -> double {  // op 0 min, 1 max, 2 sum
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if (int(threadIdx.x) < s) {
        const double a = red[threadIdx.x], b = red[threadIdx.x + s];
        red[threadIdx.x] = op == 0 ? fmin(a, b) : op == 1 ? fmax(a, b) : a + b;
      }
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  double out[13];
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    out[d] = reduce(lo[d], 0);
    out[4 + d] = reduce(hi[d], 1);
    out[8 + d] = reduce(ab[d], 1);
  }
  out[12] = reduce(bad, 2);
  if (threadIdx.x == 0)
    for (int j = 0; j < 13; ++j) partial[blockIdx.x * 13 + j] = out[j];
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = uint32_t(i);
}

template <class T> __global__ void k_fill(T* __restrict__ p, uint64_t n, T v) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

// keys = the level's coordinate of each position's point; payload = (node << 32) | point
template <class A, int K>
__global__ void k_level_keys(const A* __restrict__ pts, const uint32_t* __restrict__ P,
                             const uint32_t* __restrict__ node, uint64_t n, int dim, A* __restrict__ keys,
                             uint64_t* __restrict__ payload) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = P[i];
  keys[i] = pts[uint64_t(p) * K + dim];
  payload[i] = (uint64_t(node[i]) << 32) | p;
}

__global__ void k_unpack(const uint64_t* __restrict__ payload, uint64_t n, uint32_t* __restrict__ node,
                         uint32_t* __restrict__ P) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint64_t v = payload[i];
  node[i] = uint32_t(v >> 32);
  P[i] = uint32_t(v);
}

// begin/end of every node's run of positions (nodes with no position keep 0 / 0)
__global__ void k_node_bounds(const uint32_t* __restrict__ node, uint64_t n, uint32_t* __restrict__ begin,
                              uint32_t* __restrict__ end) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = node[i];
  if (i == 0 || node[i - 1] != s) begin[s] = uint32_t(i);
  if (i == n - 1 || node[i + 1] != s) end[s] = uint32_t(i + 1);
}

// One thread per node of the level: the split (construction.rs:673-691, :1089-1175).
template <class A, int K>
__global__ void k_node_split(const A* __restrict__ pts, const uint32_t* __restrict__ P,
                             const uint32_t* __restrict__ begin, const uint32_t* __restrict__ end, uint32_t nodes,
                             int level, int dim, A* __restrict__ stems, uint32_t* __restrict__ pivot) {
  const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= nodes) return;
  const uint32_t b = begin[s], len = end[s] - b;
  auto val = [&](uint32_t j) { return pts[uint64_t(P[b + j]) * K + dim]; };
  uint32_t pv = len;  // no split: everything goes left, the stem keeps the sentinel
  if (len >= 2) {
    const uint32_t init = (len + 1) / 2;
    const A v = val(init);
    uint32_t lo = 0, hi = init;  // first position in [0, init) whose value is not < v
    while (lo < hi) {
      const uint32_t mid = (lo + hi) / 2;
      if (val(mid) < v) lo = mid + 1;
      else hi = mid;
    }
    if (lo > 0) {
      pv = lo;
      stems[(1u << level) + s] = v;
    } else {
      lo = init + 1;  // past the end of v's run
      hi = len;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (!(v < val(mid))) lo = mid + 1;
        else hi = mid;
      }
      if (lo < len) {
        pv = lo;
        stems[(1u << level) + s] = val(lo);
      }
    }
  }
  pivot[s] = pv;
}

__global__ void k_child_node(const uint32_t* __restrict__ node, const uint32_t* __restrict__ begin,
                             const uint32_t* __restrict__ pivot, uint64_t n, uint32_t* __restrict__ next) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = node[i];
  next[i] = 2u * s + ((uint32_t(i) - begin[s]) >= pivot[s] ? 1u : 0u);
}

// Leaf-ordered SoA coordinates, items and (optionally) the FP32 shadow.
template <class A, int K>
__global__ void k_gather_tree(const A* __restrict__ pts, const uint32_t* __restrict__ items_in,
                              const uint32_t* __restrict__ P, uint64_t n, uint64_t plane, A* __restrict__ coords,
                              uint32_t* __restrict__ items, float4* __restrict__ shadow) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = P[i];
  A c[4] = {A(0), A(0), A(0), A(0)};
#pragma unroll
  for (int d = 0; d < K; ++d) {
    c[d] = pts[uint64_t(p) * K + d];
    coords[d * plane + i] = c[d];
  }
  const uint32_t it = items_in ? items_in[p] : p;
  items[i] = it;
  // {x, y, z, item bits} for K <= 3, {x, y, z, w} for K = 4
  if (shadow)
    shadow[i] = make_float4(float(c[0]), float(c[1]), K > 2 ? float(c[2]) : 0.0f,
                            K > 3 ? float(c[3]) : __uint_as_float(it));
}

}  // namespace dev

namespace {

template <class A, int K>
int build_typed(kd_tree* t, const A* pts, const uint32_t* items, uint64_t n, int B, uint32_t flags,
                hipStream_t s) {
  std::vector<void*> scratch;
  auto tmp = [&](size_t bytes) -> void* {
    void* p = nullptr;
    if (hipMallocAsync(&p, std::max<size_t>(bytes, 256), s) != hipSuccess) return nullptr;
    scratch.push_back(p);
    return p;
  };
  struct Release {
    std::vector<void*>& v;
    hipStream_t s;
    ~Release() {
      for (void* p : v) (void)hipFreeAsync(p, s);
      (void)hipStreamSynchronize(s);
    }
  } release{scratch, s};
  auto dev_alloc = [&](void** p, size_t bytes) -> int {
    bytes = std::max<size_t>(bytes, 8);
    if (hipMalloc(p, bytes) != hipSuccess) {
      *p = nullptr;
      return set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed", bytes);
    }
    t->device_bytes += bytes;
    return KD_OK;
  };
#define KD_B(call)                          \
  do {                                      \
    int rc_ = (call);                       \
    if (rc_ != KD_OK) return rc_;           \
  } while (0)

  // 1. bounds, |max| and the NaN / sentinel check
  double lo[4] = {0, 0, 0, 0}, hi[4] = {0, 0, 0, 0}, mabs[4] = {0, 0, 0, 0};
  if (n > 0) {
    const uint32_t blocks = uint32_t(std::min<uint64_t>(1024, (n + dev::kBlock - 1) / dev::kBlock));
    double* partial = static_cast<double*>(tmp(size_t(blocks) * 13 * sizeof(double)));
    if (!partial) return set_error(KD_ERR_NO_MEMORY, "device allocation failed");
    hipLaunchKernelGGL((dev::k_build_scan<A, K>), dim3(blocks), dim3(dev::kBlock), 0, s, pts, n, partial);
    KD_HIP(hipGetLastError());
    std::vector<double> h(size_t(blocks) * 13);
    KD_HIP(hipMemcpyAsync(h.data(), partial, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    double bad = 0;
    for (int d = 0; d < 4; ++d) {
      lo[d] = INFINITY;
      hi[d] = -INFINITY;
    }
    for (uint32_t b = 0; b < blocks; ++b)
      for (int d = 0; d < K; ++d) {
        lo[d] = std::min(lo[d], h[b * 13 + d]);
        hi[d] = std::max(hi[d], h[b * 13 + 4 + d]);
        mabs[d] = std::max(mabs[d], h[b * 13 + 8 + d]);
        if (d == 0) bad += h[b * 13 + 12];
      }
    if (bad > 0)
      return set_error(KD_ERR_INVALID, "a coordinate is NaN or equals the type's maximum, which Kiddo reserves "
                                       "as the no-pivot sentinel");
  }
  if ((flags & KD_TREE_SHADOW_F32))
    for (int d = 0; d < K; ++d)
      if (!(mabs[d] <= 0x1p60)) return set_error(KD_ERR_UNSUPPORTED, "FP32 shadow needs |coordinates| <= 2^60");

  // 2. depth (construction.rs:454-469): d = log2(next_pow2(ceil(n / B)))
  const uint64_t leaf_nodes = (n + uint64_t(B) - 1) / uint64_t(B);
  int depth = 0;
  if (leaf_nodes >= 2)
    while ((uint64_t(1) << depth) < leaf_nodes) ++depth;
  if (depth > kMaxStemLevels) return set_error(KD_ERR_TOO_DEEP, "tree depth %d exceeds %d", depth, kMaxStemLevels);
  const uint64_t leaves = depth > 0 ? (uint64_t(1) << depth) : (n > 0 ? 1 : 0);
  t->max_stem_level = depth - 1;
  t->leaf_count = leaves;
  t->stems_len = depth > 0 ? ((uint64_t(1) << (depth + 1)) + 7) / 8 * 8 : 0;
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->stems), std::max<uint64_t>(t->stems_len, 1) * sizeof(A)));
  if (t->stems_len)
    hipLaunchKernelGGL((dev::k_fill<A>), dim3(grid_for(t->stems_len)), dim3(dev::kBlock), 0, s,
                       static_cast<A*>(t->stems), t->stems_len, dev::Ax<A>::maxv());

  // 3. level-synchronous split
  uint32_t* P = static_cast<uint32_t*>(tmp(n * 4));
  uint32_t* node = static_cast<uint32_t*>(tmp(n * 4));
  uint32_t* node2 = static_cast<uint32_t*>(tmp(n * 4));
  uint32_t* P2 = static_cast<uint32_t*>(tmp(n * 4));
  A* keys = static_cast<A*>(tmp(n * sizeof(A)));
  A* keys2 = static_cast<A*>(tmp(n * sizeof(A)));
  uint64_t* pay = static_cast<uint64_t*>(tmp(n * 8));
  uint64_t* pay2 = static_cast<uint64_t*>(tmp(n * 8));
  const uint64_t max_nodes = depth > 0 ? (uint64_t(1) << depth) : 1;
  uint32_t* nb = static_cast<uint32_t*>(tmp(max_nodes * 4));
  uint32_t* ne = static_cast<uint32_t*>(tmp(max_nodes * 4));
  uint32_t* piv = static_cast<uint32_t*>(tmp(max_nodes * 4));
  if (!P || !node || !node2 || !P2 || !keys || !keys2 || !pay || !pay2 || !nb || !ne || !piv)
    return set_error(KD_ERR_NO_MEMORY, "device allocation failed");
  size_t t1 = 0, t2 = 0;
  KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, keys, keys2, pay, pay2, int(std::max<uint64_t>(n, 1)), 0,
                                            int(sizeof(A) * 8), s));
  KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, node, node2, P, P2, int(std::max<uint64_t>(n, 1)), 0, 32, s));
  void* sort_tmp = tmp(std::max(t1, t2));
  if (!sort_tmp) return set_error(KD_ERR_NO_MEMORY, "device allocation failed");
  const size_t sort_bytes = std::max(t1, t2);
  if (n > 0) {
    hipLaunchKernelGGL(dev::k_iota_u32, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, P, n);
    hipLaunchKernelGGL(dev::k_fill<uint32_t>, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, node, n, 0u);
  }
  for (int level = 0; level < depth && n > 0; ++level) {
    const int dim = level % K;
    const uint64_t nodes = uint64_t(1) << level;
    hipLaunchKernelGGL((dev::k_level_keys<A, K>), dim3(grid_for(n)), dim3(dev::kBlock), 0, s, pts, P, node, n, dim,
                       keys, pay);
    size_t tb = sort_bytes;
    KD_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, tb, keys, keys2, pay, pay2, int(n), 0, int(sizeof(A) * 8),
                                              s));
    hipLaunchKernelGGL(dev::k_unpack, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, pay2, n, node2, P2);
    if (level > 0) {  // stable sort by node id brings every node's positions back together
      tb = sort_bytes;
      KD_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, tb, node2, node, P2, P, int(n), 0, level, s));
    } else {
      std::swap(node, node2);
      std::swap(P, P2);
    }
    hipLaunchKernelGGL(dev::k_fill<uint32_t>, dim3(grid_for(nodes)), dim3(dev::kBlock), 0, s, nb, nodes, 0u);
    hipLaunchKernelGGL(dev::k_fill<uint32_t>, dim3(grid_for(nodes)), dim3(dev::kBlock), 0, s, ne, nodes, 0u);
    hipLaunchKernelGGL(dev::k_node_bounds, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, node, n, nb, ne);
    hipLaunchKernelGGL((dev::k_node_split<A, K>), dim3(grid_for(nodes)), dim3(dev::kBlock), 0, s, pts, P, nb, ne,
                       uint32_t(nodes), level, dim, static_cast<A*>(t->stems), piv);
    hipLaunchKernelGGL(dev::k_child_node, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, node, nb, piv, n, node2);
    std::swap(node, node2);
    KD_HIP(hipGetLastError());
  }

  // 4. leaves: runs of the final node ids (leaf indices, non-decreasing along positions)
  std::vector<uint32_t> leaf_start(leaves + 1, 0);
  uint64_t max_len = 0;
  if (depth > 0 && n > 0) {
    hipLaunchKernelGGL(dev::k_fill<uint32_t>, dim3(grid_for(leaves)), dim3(dev::kBlock), 0, s, nb, leaves, 0u);
    hipLaunchKernelGGL(dev::k_fill<uint32_t>, dim3(grid_for(leaves)), dim3(dev::kBlock), 0, s, ne, leaves, 0u);
    hipLaunchKernelGGL(dev::k_node_bounds, dim3(grid_for(n)), dim3(dev::kBlock), 0, s, node, n, nb, ne);
    std::vector<uint32_t> hb(leaves), he(leaves);
    KD_HIP(hipMemcpyAsync(hb.data(), nb, leaves * 4, hipMemcpyDeviceToHost, s));
    KD_HIP(hipMemcpyAsync(he.data(), ne, leaves * 4, hipMemcpyDeviceToHost, s));
    KD_HIP(hipStreamSynchronize(s));
    uint64_t total = 0;
    for (uint64_t l = 0; l < leaves; ++l) {
      const uint64_t len = he[l] - hb[l];
      leaf_start[l] = uint32_t(total);
      total += len;
      max_len = std::max(max_len, len);
    }
    leaf_start[leaves] = uint32_t(total);
    if (total != n) return set_error(KD_ERR_DEVICE, "device build lost entries (%llu of %llu)",
                                     (unsigned long long)total, (unsigned long long)n);
  } else if (n > 0) {
    leaf_start[1] = uint32_t(n);
    max_len = n;
  }
  if (max_len > uint64_t(kMaxLeafLen))
    return set_error(KD_ERR_LEAF_OVERFLOW, "a leaf holds %llu entries; Kiddo's limit is %d (too many equal points)",
                     (unsigned long long)max_len, kMaxLeafLen);
  t->max_leaf_len = max_len;

  // 5. the device tree
  const uint64_t plane = (n + 63) / 64 * 64;
  t->plane = plane;
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->leaf_start), (leaves + 1) * 4));
  KD_B(dev_alloc(&t->coords, std::max<uint64_t>(plane * K, 1) * sizeof(A)));
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->items), std::max<uint64_t>(n, 1) * 4));
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->d_max_abs), 4 * sizeof(double)));
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->d_bbox), 8 * sizeof(double)));
  KD_B(dev_alloc(reinterpret_cast<void**>(&t->counters), dev::kCounterWords * sizeof(unsigned long long)));
  if (flags & KD_TREE_SHADOW_F32) {
    KD_B(dev_alloc(reinterpret_cast<void**>(&t->shadow), (n + dev::kShadowPad) * 16));
    KD_HIP(hipMemsetAsync(t->shadow + n, 0, dev::kShadowPad * 16, s));  // the padding
  }
  KD_HIP(hipMemcpyAsync(t->leaf_start, leaf_start.data(), (leaves + 1) * 4, hipMemcpyHostToDevice, s));
  if (n > 0)
    hipLaunchKernelGGL((dev::k_gather_tree<A, K>), dim3(grid_for(n)), dim3(dev::kBlock), 0, s, pts, items, P, n,
                       plane, static_cast<A*>(t->coords), t->items, t->shadow);
  KD_HIP(hipGetLastError());
  double bb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int d = 0; d < K; ++d) {
    t->bbox_min[d] = n ? lo[d] : 0.0;
    t->bbox_max[d] = n ? hi[d] : 0.0;
    bb[d] = t->bbox_min[d];
    const double ext = t->bbox_max[d] - t->bbox_min[d];
    bb[4 + d] = ext > 0 ? double((1u << (32 / K)) - 1u) / ext : 0.0;
  }
  KD_HIP(hipMemcpyAsync(t->d_max_abs, mabs, 4 * sizeof(double), hipMemcpyHostToDevice, s));
  KD_HIP(hipMemcpyAsync(t->d_bbox, bb, 8 * sizeof(double), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(dev::k_fill<unsigned long long>, dim3(grid_for(dev::kCounterWords)), dim3(dev::kBlock), 0, s,
                     t->counters, dev::kCounterWords, 0ull);
  KD_HIP(hipGetLastError());
  KD_HIP(hipStreamSynchronize(s));
#undef KD_B
  return KD_OK;
}

}  // namespace

int build_device(kd_tree* t, int dtype, int dims, int bucket, uint64_t n, const void* points,
                 const uint32_t* items, uint32_t flags, hipStream_t s) {
  if (bucket < 1) return set_error(KD_ERR_INVALID, "bucket size must be >= 1");
  if (n > 0 && !points) return set_error(KD_ERR_INVALID, "points is null");
  if (n >= 0x7FFFFFFFull) return set_error(KD_ERR_UNSUPPORTED, "device builds hold < 2^31 entries");
  if ((flags & KD_TREE_SHADOW_F32) && dtype != KD_F64 && dtype != KD_F32)
    return set_error(KD_ERR_UNSUPPORTED, "the FP32 shadow is provided for f64 and f32 trees");
  t->dtype = dtype;
  t->dims = dims;
  t->bucket = bucket;
  t->size = n;
  (void)hipGetDevice(&t->device);
  KD_DISPATCH_AK(dtype, dims,
                 return (build_typed<A, KK>(t, static_cast<const A*>(points), items, n, bucket, flags, s)));
  return KD_ERR_UNSUPPORTED;
}

}  // namespace kd
