// api.hip — C ABI of the engine (include/kiddo_amd.h): tree upload and re-layout,
// Morton ordering of query batches, launch tuning, timing, and the per-family entry points
// (whose kernels live in k_knn.hip and k_families.hip).

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

#include "tree.h"

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


namespace kd {
namespace dev {

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// Autotuning sample: `runs` contiguous runs of the batch's Morton order, evenly spaced, so
// the sample keeps the spatial coherence of the Morton-ordered launch it stands for.
template <class A, int K>
__global__ void k_gather_sample(const A* __restrict__ queries, const uint32_t* __restrict__ perm, uint64_t m,
                                uint64_t ms, uint64_t runs, A* __restrict__ out) {
  const uint64_t i = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  if (i >= ms) return;
  const uint64_t run_len = ms / runs, c = i / run_len, j = i % run_len;
  const uint64_t src = perm[c * (m / runs) + j];
#pragma unroll
  for (int d = 0; d < K; ++d) out[i * K + d] = queries[src * K + d];
}

// Clears n counters (a kernel rather than hipMemsetAsync so that callers may capture it).
__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int i = int(threadIdx.x); i < n; i += int(blockDim.x)) p[i] = 0ull;
}

}  // namespace dev

uint32_t grid_for(uint64_t threads) { return uint32_t((threads + dev::kBlock - 1) / dev::kBlock); }

void reset_counters(kd_tree* t, hipStream_t s) {
  hipLaunchKernelGGL(dev::k_zero_u64, dim3(1), dim3(dev::kBlock), 0, s, t->counters, dev::kCounterWords);
}

uint32_t lds_stems_for(const kd_tree* t, size_t a_size) {
  // Default: the top levels that fit in 16 KiB, so that eight 256-thread blocks can
  // share a CU's 160 KiB of LDS.
  int levels = t->lds_levels;
  if (levels < 0) {
    levels = 0;
    while (levels < 14 && (size_t(1) << (levels + 1)) * a_size <= 16384) ++levels;
  }
  const uint64_t want = levels > 0 ? (uint64_t(1) << levels) : 0;
  return uint32_t(std::min<uint64_t>(want, t->stems_len));
}

namespace {

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
  if ((flags & KD_TREE_SHADOW_F32) && desc->dtype != KD_F64 && desc->dtype != KD_F32)
    return set_error(KD_ERR_UNSUPPORTED, "the FP32 shadow is provided for f64 and f32 trees");
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
  // zeroed padding past the last point (dev::kShadowPad)
  std::vector<float4> shadow((flags & KD_TREE_SHADOW_F32) ? n + dev::kShadowPad : 0, float4{0.0f, 0.0f, 0.0f, 0.0f});
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
        // {x, y, z, item bits} for K <= 3, {x, y, z, w} for K = 4
        float4 f;
        f.x = float(coords[0 * plane + i]);
        f.y = float(coords[1 * plane + i]);
        f.z = K > 2 ? float(coords[2 * plane + i]) : 0.0f;
        if (K > 3)
          f.w = float(coords[3 * plane + i]);
        else
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
  KD_TRY(dmalloc(&t->counters, dev::kCounterWords * sizeof(unsigned long long), tot));
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
  if (hipMemset(t->counters, 0, dev::kCounterWords * sizeof(unsigned long long)) != hipSuccess) {
    free_tree(t);
    return set_error(KD_ERR_DEVICE, "hipMemset failed");
  }
  *out = t;
  return KD_OK;
}

// Kiddo's host layout of a device tree (stems, per-leaf extents, arena of tiles), the
// format kd_tree_create takes and a Rust shim hands to KdTree (vec_of_arenas.rs:292-323).
template <class A, int K> int export_typed(const kd_tree* t, HostLayout& L) {
  const uint64_t n = t->size, leaves = t->leaf_count;
  std::vector<uint32_t> ls(leaves + 1, 0);
  std::vector<A> coords(t->plane * K);
  std::vector<uint32_t> items(n);
  auto down = [&](void* dst, const void* src, size_t bytes) -> int {
    if (bytes == 0) return KD_OK;
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return set_error(KD_ERR_DEVICE, "download failed: %s", hipGetErrorString(e));
    return KD_OK;
  };
  L.dtype = t->dtype;
  L.dims = K;
  L.bucket = t->bucket;
  L.size = n;
  L.max_stem_level = t->max_stem_level;
  L.stems_len = t->stems_len;
  L.stems.assign(t->stems_len * sizeof(A), 0);
  if (int rc = down(L.stems.data(), t->stems, L.stems.size())) return rc;
  if (int rc = down(ls.data(), t->leaf_start, ls.size() * 4)) return rc;
  if (int rc = down(coords.data(), t->coords, coords.size() * sizeof(A))) return rc;
  if (int rc = down(items.data(), t->items, items.size() * 4)) return rc;
  L.extents.assign(2 * leaves, 0);
  uint64_t off = 0, max_len = 0;
  for (uint64_t l = 0; l < leaves; ++l) {
    off = (off + 7) & ~uint64_t(7);
    const uint64_t len = ls[l + 1] - ls[l];
    L.extents[2 * l] = off;
    L.extents[2 * l + 1] = len;
    off += len * (K * sizeof(A) + 4);
    max_len = std::max(max_len, len);
  }
  L.max_leaf_len = max_len;
  L.arena.assign(off, 0);
  host_parallel(leaves, [&](uint64_t s, uint64_t e) {
    for (uint64_t l = s; l < e; ++l) {
      uint8_t* dst = L.arena.data() + L.extents[2 * l];
      const uint64_t b = ls[l], len = ls[l + 1] - b;
      for (uint64_t base = 0; base < len;) {
        const uint64_t rem = len - base;
        const int tl = rem >= 32 ? 32 : rem >= 8 ? 8 : rem >= 4 ? 4 : rem >= 2 ? 2 : 1;
        for (int d = 0; d < K; ++d)
          for (int i = 0; i < tl; ++i) {
            std::memcpy(dst, &coords[d * t->plane + b + base + i], sizeof(A));
            dst += sizeof(A);
          }
        for (int i = 0; i < tl; ++i) {
          std::memcpy(dst, &items[b + base + i], 4);
          dst += 4;
        }
        base += tl;
      }
    }
  });
  return KD_OK;
}

// Optional Morton ordering of a batch: returns a device permutation (or null).
struct Order {
  uint32_t* perm = nullptr;
  void* mem = nullptr;
  hipStream_t s = nullptr;
  ~Order() {
    if (mem) (void)hipFreeAsync(mem, s);
  }
};

// ext (optional): caller-provided device scratch of ext_bytes used instead of a per-call
// allocation (the keys and ids are written in full before the sort reads them, so it need
// not be cleared).
template <class A, int K>
int make_order(const kd_tree* t, const A* queries, uint64_t m, uint32_t flags, hipStream_t s, Order& o,
               void* ext = nullptr, size_t ext_bytes = 0) {
  if (!(flags & KD_MORTON) || m < 2) return KD_OK;
  if (m > 0xFFFFFFFFull) return set_error(KD_ERR_INVALID, "Morton ordering supports < 2^32 queries");
  o.s = s;
  size_t temp = 0;
  uint32_t *k0 = nullptr, *k1 = nullptr, *v0 = nullptr, *v1 = nullptr;
  KD_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp, k0, k1, v0, v1, int(m), 0, 32, s));
  const size_t arr = ((m * 4 + 255) / 256) * 256;
  const size_t total = 4 * arr + temp;
  uint8_t* base = nullptr;
  if (ext) {
    if (total > ext_bytes)
      return set_error(KD_ERR_CAPACITY, "Morton ordering of %llu queries needs %zu bytes of scratch, %zu held",
                       (unsigned long long)m, total, ext_bytes);
    base = static_cast<uint8_t*>(ext);
  } else {
    KD_HIP(hipMallocAsync(&o.mem, total, s));
    base = static_cast<uint8_t*>(o.mem);
  }
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

int check_radius(double radius) {
  if (std::isnan(radius)) return set_error(KD_ERR_INVALID, "radius is NaN");
  return KD_OK;
}

}  // namespace
}  // namespace kd

// ------------------------------------------------ allocation-free entry point --------
// kd_nearest_n_noalloc takes its Morton-ordering scratch (keys and ids, two buffers each,
// and the radix sort's temporary storage) from this array instead of a per-call
// allocation: up to KD_NOALLOC_MAX_QUERIES queries, 32 MiB per GPU, held until the process
// ends. Host code never passes `scratch` itself to a kernel or copy (on the host it only
// names the symbol); it looks up each device's address once with hipGetSymbolAddress.
namespace kd {
namespace {
constexpr size_t kNoAllocArr = size_t(KD_NOALLOC_MAX_QUERIES) * 4;  // one u32 array
constexpr size_t kNoAllocSortTemp = size_t(16) << 20;  // checked per call against the sort's request
constexpr size_t kScratchBytes = 4 * kNoAllocArr + kNoAllocSortTemp;
}  // namespace
}  // namespace kd

static __device__ __align__(256) unsigned char scratch[kd::kScratchBytes];

namespace kd {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
int device_scratch(int device, void** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(KD_ERR_UNSUPPORTED, "device %d", device);
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return set_error(KD_ERR_DEVICE, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_ptr[device] = p;
  }
  *out = g_scratch_ptr[device];
  return KD_OK;
}
}  // namespace
}  // namespace kd

// ====================================================================== C ABI ========

struct kd_layout {
  kd::HostLayout L;
};

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

}  // extern "C"

namespace kd {
namespace {
// A position in a Donnelly<bh> stem array (kiddo src/stem_strategy/donnelly/core.rs): the
// array is cut into lines of 2^bh slots, each holding a triangle of bh levels (slot 0 of a
// line's triangle at offset 0, the children of offset i at 2i + 1 and 2i + 2); the child of
// a triangle's bottom row is the root of another line. `row` is the level inside the
// triangle. The step is the reference's child-index rule (core.rs:254-289), in u32 with
// wrap-around.
struct DonnellyAt {
  uint32_t slot, row;
};
DonnellyAt donnelly_child(uint32_t bh, DonnellyAt at, uint32_t right) {
  const uint32_t in_line = at.slot & ((1u << bh) - 1u);
  const uint32_t base = (at.slot & ~((1u << bh) - 1u)) + 1u;
  if (at.row + 1u < bh) return {base + right + 2u * in_line, at.row + 1u};
  const uint32_t column = in_line - at.row - 1u;
  return {(base + 2u * column + right) << bh, 0u};
}

// Kiddo pads a block layout's stem depth to a multiple of the block height with levels
// above the real root that hold no pivot and are always taken to the left
// (construction.rs:431-444), so the real root sits `pad` levels down. In Eytzinger order
// that is node 2^pad at level pad, and every real node keeps the level (hence split
// dimension) it has in the Donnelly tree.
template <class A> int from_donnelly_typed(const kd_tree_desc* d, uint32_t bh, HostLayout& L) {
  const int D = d->max_stem_level + 1;
  const uint64_t real = d->leaf_count;
  if (real == 0 || (real & (real - 1)) != 0)
    return set_error(KD_ERR_INVALID, "a Donnelly tree holds a power-of-two number of leaves, got %llu",
                     (unsigned long long)real);
  int dr = 0;
  while ((uint64_t(1) << dr) < real) ++dr;
  const int pad = D - dr;
  if (D % int(bh) != 0 || pad < 0 || pad >= int(bh))
    return set_error(KD_ERR_INVALID,
                     "stem depth %d with %llu leaves is not a Donnelly<%u> layout (depth a multiple of %u, "
                     "fewer than %u padding levels)", D, (unsigned long long)real, bh, bh, bh);
  if (D > kMaxStemLevels) return set_error(KD_ERR_TOO_DEEP, "max_stem_level %d too deep", D - 1);
  const uint64_t leaves = uint64_t(1) << D;
  const uint64_t stems_len = ((uint64_t(2) << D) + 7) / 8 * 8;
  const A* src = static_cast<const A*>(d->stems);
  std::vector<A> eyt(stems_len, AxisHost<A>::max_value());
  DonnellyAt root{0u, 0u};
  for (int i = 0; i < pad; ++i) root = donnelly_child(bh, root, 0u);
  struct Frame {
    uint64_t node;
    int level;
    DonnellyAt at;
  };
  std::vector<Frame> stack{{uint64_t(1) << pad, pad, root}};
  while (!stack.empty()) {
    const Frame f = stack.back();
    stack.pop_back();
    if (f.at.slot >= d->stems_len)
      return set_error(KD_ERR_INVALID, "stem slot %u of a level-%d node lies past the %llu stems", f.at.slot,
                       f.level, (unsigned long long)d->stems_len);
    eyt[f.node] = src[f.at.slot];
    if (f.level + 1 < D) {
      stack.push_back({2 * f.node, f.level + 1, donnelly_child(bh, f.at, 0u)});
      stack.push_back({2 * f.node + 1, f.level + 1, donnelly_child(bh, f.at, 1u)});
    }
  }
  L.dtype = d->dtype;
  L.dims = d->dims;
  L.bucket = d->bucket;
  L.size = d->size;
  L.max_stem_level = D - 1;
  L.stems_len = stems_len;
  L.stems.assign(reinterpret_cast<const uint8_t*>(eyt.data()),
                 reinterpret_cast<const uint8_t*>(eyt.data()) + stems_len * sizeof(A));
  // the arena, grown to a multiple of 8 bytes, where the empty leaves start (as an arena
  // that appended them would place them, vec_of_arenas.rs:292-323)
  const uint64_t end = (d->leaf_bytes_len + 7) / 8 * 8;
  L.arena.assign(end, 0);
  if (d->leaf_bytes_len) std::memcpy(L.arena.data(), d->leaf_bytes, d->leaf_bytes_len);
  L.extents.assign(2 * leaves, 0);
  L.max_leaf_len = 0;
  for (uint64_t l = 0; l < leaves; ++l) {
    const bool r = l < real;
    L.extents[2 * l] = r ? d->leaf_extents[2 * l] : end;
    L.extents[2 * l + 1] = r ? d->leaf_extents[2 * l + 1] : 0;
    L.max_leaf_len = std::max<uint64_t>(L.max_leaf_len, L.extents[2 * l + 1]);
  }
  return KD_OK;
}
}  // namespace
}  // namespace kd

extern "C" {

int kd_layout_from_donnelly(const kd_tree_desc* donnelly, int32_t block_height, kd_layout** out) {
  if (!donnelly || !out) return kd::set_error(KD_ERR_INVALID, "null argument");
  // the bottom row of a triangle starts at slot 2^(bh-1) - 1, which the child rule's
  // column (slot - row - 1) only respects from bh = 3 up
  if (block_height < 3 || block_height > 16)
    return kd::set_error(KD_ERR_INVALID, "Donnelly block height %d outside 3..16", block_height);
  if (donnelly->leaf_count && (!donnelly->leaf_extents || !donnelly->leaf_bytes))
    return kd::set_error(KD_ERR_INVALID, "leaf buffers are null");
  if (donnelly->stems_len && !donnelly->stems) return kd::set_error(KD_ERR_INVALID, "stems is null");
  kd_layout* l = nullptr;
  try {
    l = new kd_layout;
    int rc = KD_OK;
    if (donnelly->max_stem_level < 0) {  // no stems: the layouts are the same
      kd::HostLayout& L = l->L;
      L.dtype = donnelly->dtype;
      L.dims = donnelly->dims;
      L.bucket = donnelly->bucket;
      L.size = donnelly->size;
      L.max_stem_level = -1;
      L.stems_len = 0;
      L.extents.assign(donnelly->leaf_extents, donnelly->leaf_extents + 2 * donnelly->leaf_count);
      L.arena.assign(donnelly->leaf_bytes, donnelly->leaf_bytes + donnelly->leaf_bytes_len);
      for (uint64_t i = 0; i < donnelly->leaf_count; ++i)
        L.max_leaf_len = std::max<uint64_t>(L.max_leaf_len, L.extents[2 * i + 1]);
    } else {
      switch (donnelly->dtype) {
        case KD_F32: rc = kd::from_donnelly_typed<float>(donnelly, uint32_t(block_height), l->L); break;
        case KD_F64: rc = kd::from_donnelly_typed<double>(donnelly, uint32_t(block_height), l->L); break;
        case KD_U16: rc = kd::from_donnelly_typed<uint16_t>(donnelly, uint32_t(block_height), l->L); break;
        case KD_U32: rc = kd::from_donnelly_typed<uint32_t>(donnelly, uint32_t(block_height), l->L); break;
        default: rc = kd::set_error(KD_ERR_UNSUPPORTED, "unsupported dtype %d", donnelly->dtype);
      }
    }
    if (rc != KD_OK) {
      delete l;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete l;
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed while re-laying out");
  }
  *out = l;
  return KD_OK;
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

int kd_tree_build_device(int32_t dtype, int32_t dims, int32_t bucket, uint64_t n, const void* points,
                         const uint32_t* items, uint32_t tree_flags, kd_tree** out, void* stream) {
  if (!out) return kd::set_error(KD_ERR_INVALID, "out is null");
  kd_tree* t = nullptr;
  try {
    t = new kd_tree;
  } catch (const std::bad_alloc&) {
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed");
  }
  const int rc = kd::build_device(t, dtype, dims, bucket, n, points, items, tree_flags,
                                  static_cast<hipStream_t>(stream));
  if (rc != KD_OK) {
    kd::free_tree(t);
    return rc;
  }
  *out = t;
  return KD_OK;
}

int kd_tree_export_layout(const kd_tree* t, kd_layout** out) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!out) return kd::set_error(KD_ERR_INVALID, "out is null");
  kd_layout* l = nullptr;
  try {
    l = new kd_layout;
    int rc = KD_OK;
    KD_DISPATCH_AK(t->dtype, t->dims, rc = kd::export_typed<A, KK>(t, l->L));
    if (rc != KD_OK) {
      delete l;
      return rc;
    }
  } catch (const std::bad_alloc&) {
    delete l;
    return kd::set_error(KD_ERR_NO_MEMORY, "host allocation failed while exporting");
  }
  *out = l;
  return KD_OK;
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

int kd_tree_tune(kd_tree* t, int32_t knn_variant, int32_t group_size, int32_t lds_levels) {
  return kd_tree_tune2(t, knn_variant, group_size, lds_levels, 0);
}

int kd_tree_tune2(kd_tree* t, int32_t knn_variant, int32_t group_size, int32_t lds_levels, int32_t waves) {
  if (waves != 0 && (waves < 3 || waves > 6)) return kd::set_error(KD_ERR_INVALID, "waves must be 0 or 3..6");
  if (int rc = kd::check_tree(t)) return rc;
  if (knn_variant < 0 || knn_variant > 2) return kd::set_error(KD_ERR_INVALID, "knn_variant must be 0, 1 or 2");
  if (group_size != 0 && group_size != 4 && group_size != 8 && group_size != 16 && group_size != 32 &&
      group_size != 64)
    return kd::set_error(KD_ERR_INVALID, "group_size must be 0, 4, 8, 16, 32 or 64");
  if (lds_levels < -1 || lds_levels > 14) return kd::set_error(KD_ERR_INVALID, "lds_levels must be in [-1, 14]");
  t->knn_variant = knn_variant;
  t->group_size = group_size;
  t->lds_levels = lds_levels;
  t->waves = waves;
  return KD_OK;
}

int kd_tree_tune_within(kd_tree* t, int32_t count_group, int32_t fill_group) {
  if (int rc = kd::check_tree(t)) return rc;
  auto ok = [](int g) { return g == 0 || g == 1 || g == 4 || g == 8 || g == 16 || g == 32; };
  if (!ok(count_group) || !ok(fill_group))
    return kd::set_error(KD_ERR_INVALID, "within launch shapes must be 0, 1, 4, 8, 16 or 32");
  t->count_group = count_group;
  t->fill_group = fill_group;
  return KD_OK;
}

int kd_tree_tune_best_n_within(kd_tree* t, int32_t group) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!(group == 0 || group == 1 || group == 4 || group == 8 || group == 16 || group == 32))
    return kd::set_error(KD_ERR_INVALID, "best_n_within launch shape must be 0, 1, 4, 8, 16 or 32");
  t->best_group = group;
  return KD_OK;
}

namespace kd {
namespace {
// Metric field of the flags (KD_METRIC_*): kind 0..3; Minkowski exponents 1 and 2 are
// refused as in minkowski/mod.rs:36-45 (use Manhattan / SquaredEuclidean); the FP32 screen
// is derived for SquaredEuclidean only.
int check_metric(uint32_t flags) {
  const int kind = int((flags >> 8) & 0xFu), p = int((flags >> 12) & 0xFFu);
  if (kind > 3) return set_error(KD_ERR_INVALID, "unknown metric %d", kind);
  if (kind != 3 && p != 0) return set_error(KD_ERR_INVALID, "a Minkowski exponent needs KD_METRIC_MINKOWSKI");
  if (kind == 3 && (p == 1 || p == 2))
    return set_error(KD_ERR_INVALID, "Minkowski<%d>: use Manhattan (P = 1) or SquaredEuclidean (P = 2)", p);
  if (kind != 0 && (flags & KD_SCREEN_F32))
    return set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 applies to SquaredEuclidean only");
  return KD_OK;
}
}  // namespace
}  // namespace kd

#define KD_ORDERED(a, queries_, m_, flags_, stream_)                                  \
  if (int mrc_ = kd::check_metric(flags_)) return mrc_;                               \
  hipStream_t s_ = static_cast<hipStream_t>(stream_);                                 \
  kd::Order order_;                                                                   \
  {                                                                                   \
    int orc_ = KD_OK;                                                                 \
    KD_DISPATCH_AK(t->dtype, t->dims,                                                 \
                   orc_ = (kd::make_order<A, KK>(t, static_cast<const A*>(queries_), m_, flags_, s_, order_))); \
    if (orc_) return orc_;                                                            \
  }                                                                                   \
  const kd::QueryArgs a{t, queries_, m_, order_.perm, flags_, s_};

int kd_nearest_one(kd_tree* t, const void* queries, uint64_t m, uint32_t flags, uint32_t* items,
                   void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_nearest_one(a, items, dists);
}

int kd_approx_nearest_one(kd_tree* t, const void* queries, uint64_t m, uint32_t flags, uint32_t* items,
                          void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (t->dtype != KD_F32 && t->dtype != KD_F64)
    return kd::set_error(KD_ERR_UNSUPPORTED, "approx nearest_one needs a float tree (Output == A)");
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_approx_nearest_one(a, items, dists);
}

int kd_nearest_n(kd_tree* t, const void* queries, uint64_t m, uint32_t k, double radius, uint32_t flags,
                 uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1 (NonZeroUsize)");
  if (int rc = kd::check_radius(radius)) return rc;
  if ((flags & KD_SCREEN_F32) && !t->shadow)
    return kd::set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 needs a tree created with KD_TREE_SHADOW_F32");
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_nearest_n(a, int(k), radius, items, dists, counts);
}

int kd_nearest_n_noalloc(kd_tree* t, const void* queries, uint64_t m, uint32_t k, double radius,
                         uint32_t flags, uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1 (NonZeroUsize)");
  if (int rc = kd::check_radius(radius)) return rc;
  if ((flags & KD_SCREEN_F32) && !t->shadow)
    return kd::set_error(KD_ERR_UNSUPPORTED, "KD_SCREEN_F32 needs a tree created with KD_TREE_SHADOW_F32");
  if ((flags & KD_MORTON) && m > uint64_t(KD_NOALLOC_MAX_QUERIES))
    return kd::set_error(KD_ERR_CAPACITY, "kd_nearest_n_noalloc orders at most %u queries (KD_MORTON), got %llu",
                         unsigned(KD_NOALLOC_MAX_QUERIES), (unsigned long long)m);
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  if (int mrc = kd::check_metric(flags)) return mrc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  kd::Order order;
  if ((flags & KD_MORTON) && m >= 2) {
    int dev = 0;
    KD_HIP(hipGetDevice(&dev));
    if (dev != t->device)
      return kd::set_error(KD_ERR_INVALID, "the tree lives on device %d, device %d is current", t->device, dev);
    void* scr = nullptr;
    if (int rc = kd::device_scratch(dev, &scr)) return rc;
    int orc = KD_OK;
    KD_DISPATCH_AK(t->dtype, t->dims,
                   orc = (kd::make_order<A, KK>(t, static_cast<const A*>(queries), m, flags, s, order, scr,
                                                kd::kScratchBytes)));
    if (orc) return orc;
  }
  const kd::QueryArgs a{t, queries, m, order.perm, flags, s};
  return kd::launch_nearest_n(a, int(k), radius, items, dists, counts);
}

__global__ void kd_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * kd::dev::kBlock + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

int kd_morton_order(kd_tree* t, const void* queries, uint64_t m, uint32_t* perm, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !perm) return kd::set_error(KD_ERR_INVALID, "null buffer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  kd::Order order;
  int orc = KD_OK;
  KD_DISPATCH_AK(t->dtype, t->dims,
                 orc = (kd::make_order<A, KK>(t, static_cast<const A*>(queries), m, KD_MORTON, s, order)));
  if (orc) return orc;
  if (!order.perm) {  // a single query
    const uint32_t zero = 0;
    KD_HIP(hipMemcpyAsync(perm, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    KD_HIP(hipStreamSynchronize(s));
    return KD_OK;
  }
  hipLaunchKernelGGL(kd_copy_u32, dim3(kd::grid_for(m)), dim3(kd::dev::kBlock), 0, s, order.perm, perm, m);
  KD_HIP(hipGetLastError());
  return KD_OK;
}

int kd_within_count(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t flags,
                    uint64_t* offsets, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!offsets) return kd::set_error(KD_ERR_INVALID, "offsets is null");
  if (int rc = kd::check_radius(radius)) return rc;
  if (m == 0) {
    hipLaunchKernelGGL(kd::dev::k_set_u64, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), offsets,
                       uint64_t(0));
    KD_HIP(hipGetLastError());
    return KD_OK;
  }
  if (!queries) return kd::set_error(KD_ERR_INVALID, "queries is null");
  if (m >= 0x7FFFFFFFull) return kd::set_error(KD_ERR_UNSUPPORTED, "within batches hold < 2^31 queries");
  KD_ORDERED(a, queries, m, flags, stream);
  // counts into a scratch array, then an exclusive scan of m+1 entries into offsets
  void* mem = nullptr;
  size_t temp = 0;
  uint64_t* cnt = nullptr;
  KD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, cnt, offsets, int(m + 1), s_));
  const size_t arr = ((m + 1) * 8 + 255) / 256 * 256;
  KD_HIP(hipMallocAsync(&mem, arr + temp, s_));
  cnt = static_cast<uint64_t*>(mem);
  int rc = kd::launch_within_count(a, radius, cnt);
  if (rc == KD_OK) {
    hipLaunchKernelGGL(kd::dev::k_set_u64, dim3(1), dim3(1), 0, s_, cnt + m, uint64_t(0));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(static_cast<uint8_t*>(mem) + arr, temp, cnt, offsets,
                                                    int(m + 1), s_);
    if (e != hipSuccess) rc = kd::set_error(KD_ERR_DEVICE, "scan failed: %s", hipGetErrorString(e));
  }
  (void)hipFreeAsync(mem, s_);
  return rc;
}

int kd_within_fill(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t flags,
                   const uint64_t* offsets, uint32_t* items, void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !offsets) return kd::set_error(KD_ERR_INVALID, "null buffer");
  if (int rc = kd::check_radius(radius)) return rc;
  KD_ORDERED(a, queries, m, flags, stream);
  // dense outputs (>= 32 results per query on average) take the group-per-query fill,
  // whose lanes write consecutive results
  uint64_t total = 0;
  KD_HIP(hipMemcpyAsync(&total, offsets + m, sizeof(uint64_t), hipMemcpyDeviceToHost, s_));
  KD_HIP(hipStreamSynchronize(s_));
  kd::QueryArgs ad = a;
  ad.dense = total >= 32 * m;
  return kd::launch_within_fill(ad, radius, offsets, items, dists);
}

int kd_best_n_within(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t n, uint32_t flags,
                     uint32_t* items, void* dists, uint32_t* counts, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (n == 0) return kd::set_error(KD_ERR_INVALID, "n must be >= 1 (NonZero)");
  if (int rc = kd::check_radius(radius)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_best_n_within(a, int(n), radius, items, dists, counts);
}

namespace kd {
namespace {
// builder.rs:29-35: every periodic box length must be > 0 (NaN fails); float trees only
int check_periodic(const kd_tree* t, const void* box) {
  if (t->dtype != KD_F32 && t->dtype != KD_F64)
    return set_error(KD_ERR_UNSUPPORTED, "periodic queries need a float tree (PeriodicAxis is f32/f64)");
  if (!box) return set_error(KD_ERR_INVALID, "box is null");
  for (int d = 0; d < t->dims; ++d) {
    const double L = t->dtype == KD_F32 ? double(static_cast<const float*>(box)[d])
                                        : static_cast<const double*>(box)[d];
    if (!(L > 0.0)) return set_error(KD_ERR_INVALID, "periodic box sizes must be strictly positive");
  }
  return KD_OK;
}
}  // namespace
}  // namespace kd

int kd_periodic_nearest_one(kd_tree* t, const void* queries, uint64_t m, const void* box, uint32_t flags,
                            uint32_t* items, void* dists, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (int rc = kd::check_periodic(t, box)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_periodic_nearest_one(a, box, items, dists);
}

int kd_periodic_nearest_n(kd_tree* t, const void* queries, uint64_t m, const void* box, uint32_t k,
                          double radius, uint32_t flags, uint32_t* items, void* dists, uint32_t* counts,
                          void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (int rc = kd::check_periodic(t, box)) return rc;
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1 (NonZeroUsize)");
  if (int rc = kd::check_radius(radius)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !items || !dists || !counts) return kd::set_error(KD_ERR_INVALID, "null buffer");
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_periodic_nearest_n(a, box, int(k), radius, items, dists, counts);
}

int kd_periodic_within_count(kd_tree* t, const void* queries, uint64_t m, const void* box, double radius,
                             uint32_t flags, uint64_t* offsets, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (int rc = kd::check_periodic(t, box)) return rc;
  if (!offsets) return kd::set_error(KD_ERR_INVALID, "offsets is null");
  if (int rc = kd::check_radius(radius)) return rc;
  if (m == 0) {
    hipLaunchKernelGGL(kd::dev::k_set_u64, dim3(1), dim3(1), 0, static_cast<hipStream_t>(stream), offsets,
                       uint64_t(0));
    KD_HIP(hipGetLastError());
    return KD_OK;
  }
  if (!queries) return kd::set_error(KD_ERR_INVALID, "queries is null");
  if (m >= 0x7FFFFFFFull) return kd::set_error(KD_ERR_UNSUPPORTED, "within batches hold < 2^31 queries");
  KD_ORDERED(a, queries, m, flags, stream);
  void* mem = nullptr;
  size_t temp = 0;
  uint64_t* cnt = nullptr;
  KD_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, temp, cnt, offsets, int(m + 1), s_));
  const size_t arr = ((m + 1) * 8 + 255) / 256 * 256;
  KD_HIP(hipMallocAsync(&mem, arr + temp, s_));
  cnt = static_cast<uint64_t*>(mem);
  int rc = kd::launch_periodic_within_count(a, box, radius, cnt);
  if (rc == KD_OK) {
    hipLaunchKernelGGL(kd::dev::k_set_u64, dim3(1), dim3(1), 0, s_, cnt + m, uint64_t(0));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(static_cast<uint8_t*>(mem) + arr, temp, cnt, offsets,
                                                    int(m + 1), s_);
    if (e != hipSuccess) rc = kd::set_error(KD_ERR_DEVICE, "scan failed: %s", hipGetErrorString(e));
  }
  (void)hipFreeAsync(mem, s_);
  return rc;
}

int kd_periodic_within_fill(kd_tree* t, const void* queries, uint64_t m, const void* box, double radius,
                            uint32_t flags, const uint64_t* offsets, uint32_t* items, void* dists,
                            void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (int rc = kd::check_periodic(t, box)) return rc;
  if (m == 0) return KD_OK;
  if (!queries || !offsets) return kd::set_error(KD_ERR_INVALID, "null buffer");
  if (int rc = kd::check_radius(radius)) return rc;
  KD_ORDERED(a, queries, m, flags, stream);
  return kd::launch_periodic_within_fill(a, box, radius, offsets, items, dists);
}

}  // extern "C"

namespace kd {
namespace {
// The autotuners' sample of ms queries of a batch: the first ms queries, or, for a
// Morton-ordered batch, 16 evenly spaced runs of its own Morton order (its first queries in
// input order are spatially sparse and tune for the wrong shape). *mem (freed by the
// caller with hipFree) holds a gathered sample.
int tuning_sample(kd_tree* t, const void* queries, uint64_t m, uint64_t ms, uint32_t flags, hipStream_t s,
                  void** mem, const void** sample) {
  constexpr uint64_t kRuns = 16;
  *mem = nullptr;
  *sample = queries;
  if (!((flags & KD_MORTON) && m > ms && ms % kRuns == 0)) return KD_OK;
  int src = KD_OK;
  KD_DISPATCH_AK(t->dtype, t->dims, {
    Order order;
    src = make_order<A, KK>(t, static_cast<const A*>(queries), m, flags, s, order);
    if (src == KD_OK && order.perm) {
      if (hipMalloc(mem, ms * KK * sizeof(A)) != hipSuccess) {
        *mem = nullptr;
        src = set_error(KD_ERR_NO_MEMORY, "hipMalloc failed");
      } else {
        hipLaunchKernelGGL((dev::k_gather_sample<A, KK>), dim3(grid_for(ms)), dim3(dev::kBlock), 0, s,
                           static_cast<const A*>(queries), order.perm, m, ms, kRuns, static_cast<A*>(*mem));
        if (hipGetLastError() != hipSuccess) src = set_error(KD_ERR_DEVICE, "sample gather failed");
        (void)hipStreamSynchronize(s);
        *sample = *mem;
      }
    }
  });
  if (src != KD_OK) {
    (void)hipFree(*mem);
    *mem = nullptr;
  }
  return src;
}
}  // namespace
}  // namespace kd

extern "C" {

int kd_autotune_nearest_n(kd_tree* t, const void* queries, uint64_t m, uint32_t k, double radius,
                          uint32_t flags, int32_t* chosen4, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1");
  if (!queries) return kd::set_error(KD_ERR_INVALID, "queries is null");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // 2^20 queries: a thread-per-query candidate then still fills the device (256 CUs x 2048
  // lanes); a smaller sample under-occupies it and favours the group forms
  const uint64_t ms = std::min<uint64_t>(m, uint64_t(1) << 20);
  if (ms == 0) return KD_OK;
  // candidate launch shapes: thread per query, and groups of the two smallest sizes
  // holding k, each with and without LDS-pinned stems
  int g0 = k <= 4 ? 4 : k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
  struct Cand { int variant, group, lds, waves; };
  std::vector<Cand> cands = {{1, 0, 0, 0}, {1, 0, -1, 0}, {2, g0, -1, 0}};
  if (flags & KD_SCREEN_F32)
    for (int w = 4; w <= 6; ++w) cands.push_back({1, 0, -1, w});
  if (g0 < 64) cands.push_back({2, 2 * g0, -1, 0});
  void* sample_mem = nullptr;
  const void* sample = queries;
  if (int src = kd::tuning_sample(t, queries, m, ms, flags, s, &sample_mem, &sample)) return src;
  void* mem = nullptr;
  const size_t osz = ms * k * (sizeof(uint32_t) + sizeof(double)) + ms * sizeof(uint32_t) + 1024;
  if (hipMalloc(&mem, osz) != hipSuccess) {
    (void)hipFree(sample_mem);
    return kd::set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed", osz);
  }
  uint32_t* items = static_cast<uint32_t*>(mem);
  double* dists = reinterpret_cast<double*>(static_cast<uint8_t*>(mem) + ((ms * k * 4 + 255) / 256) * 256);
  uint32_t* counts = reinterpret_cast<uint32_t*>(dists + ms * k);
  const int save_variant = t->knn_variant, save_group = t->group_size, save_lds = t->lds_levels,
            save_waves = t->waves;
  const bool save_timing = t->timing;
  int rc = kd_tree_set_timing(t, 1);
  float best_ms = 1e30f;
  Cand best = cands[0];
  for (const Cand& c : cands) {
    if (rc) break;
    t->knn_variant = c.variant;
    t->group_size = c.group;
    t->lds_levels = c.lds;
    t->waves = c.waves;
    float ms_c = 0.0f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      rc = kd_nearest_n(t, sample, ms, k, radius, flags, items, dists, counts, stream);
      if (rc == KD_OK) rc = kd_last_kernel_ms(t, &ms_c);
    }
    if (rc == KD_OK && ms_c < best_ms) {
      best_ms = ms_c;
      best = c;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(mem);
  (void)hipFree(sample_mem);
  t->timing = save_timing;
  if (rc) {
    t->knn_variant = save_variant;
    t->group_size = save_group;
    t->lds_levels = save_lds;
    t->waves = save_waves;
    return rc;
  }
  t->knn_variant = best.variant;
  t->group_size = best.group;
  t->lds_levels = best.lds;
  t->waves = best.waves;
  if (chosen4) {
    chosen4[0] = best.variant;
    chosen4[1] = best.group;
    chosen4[2] = best.lds;
    chosen4[3] = best.waves;
  }
  return KD_OK;
}

// Bucket size by measurement: the tree is built on the device at each candidate bucket (the
// canonical answer does not depend on it), its launch shape autotuned, and nearest_n timed on
// the tuning sample of the batch (the same sample kd_autotune_nearest_n uses, up to 2^20
// queries); the fastest bucket wins.
int kd_autotune_bucket(int32_t dtype, int32_t dims, uint64_t n, const void* points, const uint32_t* items,
                       uint32_t tree_flags, const void* queries, uint64_t m, uint32_t k, double radius,
                       uint32_t flags, const int32_t* buckets, uint32_t nbuckets, int32_t* best_bucket,
                       float* ms_out, void* stream) {
  if (!points || !queries || !buckets || !best_bucket) return kd::set_error(KD_ERR_INVALID, "null argument");
  if (nbuckets == 0) return kd::set_error(KD_ERR_INVALID, "no candidate bucket");
  if (k == 0) return kd::set_error(KD_ERR_INVALID, "k must be >= 1");
  if (m == 0) return kd::set_error(KD_ERR_INVALID, "no queries");
  for (uint32_t i = 0; i < nbuckets; ++i)
    if (buckets[i] < 1 || buckets[i] > 1024)
      return kd::set_error(KD_ERR_INVALID, "bucket %d is not in 1..1024", buckets[i]);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t ms = std::min<uint64_t>(m, uint64_t(1) << 20);
  const size_t osz = ms * k * (sizeof(uint32_t) + sizeof(double)) + ms * sizeof(uint32_t) + 1024;
  void* mem = nullptr;
  if (hipMalloc(&mem, osz) != hipSuccess) return kd::set_error(KD_ERR_NO_MEMORY, "hipMalloc(%zu) failed", osz);
  uint32_t* oitems = static_cast<uint32_t*>(mem);
  double* odists = reinterpret_cast<double*>(static_cast<uint8_t*>(mem) + ((ms * k * 4 + 255) / 256) * 256);
  uint32_t* ocounts = reinterpret_cast<uint32_t*>(odists + ms * k);
  float best_ms = 1e30f;
  int rc = KD_OK;
  *best_bucket = buckets[0];
  for (uint32_t i = 0; i < nbuckets && rc == KD_OK; ++i) {
    kd_tree* t = nullptr;
    rc = kd_tree_build_device(dtype, dims, buckets[i], n, points, items, tree_flags, &t, stream);
    if (rc) break;
    void* sample_mem = nullptr;
    const void* sample = queries;
    rc = kd::tuning_sample(t, queries, m, ms, flags, s, &sample_mem, &sample);
    if (rc == KD_OK) rc = kd_autotune_nearest_n(t, sample, ms, k, radius, flags, nullptr, stream);
    if (rc == KD_OK) rc = kd_tree_set_timing(t, 1);
    float tm = 1e30f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      float e = 0.0f;
      rc = kd_nearest_n(t, sample, ms, k, radius, flags, oitems, odists, ocounts, stream);
      if (rc == KD_OK) rc = kd_last_kernel_ms(t, &e);
      tm = std::min(tm, e);
    }
    (void)hipStreamSynchronize(s);
    (void)hipFree(sample_mem);
    kd_tree_destroy(t);
    if (rc) break;
    if (ms_out) ms_out[i] = tm;
    if (tm < best_ms) {
      best_ms = tm;
      *best_bucket = buckets[i];
    }
  }
  (void)hipFree(mem);
  return rc;
}

int kd_autotune_within(kd_tree* t, const void* queries, uint64_t m, double radius, uint32_t flags,
                       int32_t* chosen2, void* stream) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!queries) return kd::set_error(KD_ERR_INVALID, "queries is null");
  if (int rc = kd::check_radius(radius)) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const uint64_t ms = std::min<uint64_t>(m, uint64_t(1) << 20);
  if (ms == 0) return KD_OK;
  void* sample_mem = nullptr;
  const void* sample = queries;
  if (int src = kd::tuning_sample(t, queries, m, ms, flags, s, &sample_mem, &sample)) return src;
  const int save_count = t->count_group, save_fill = t->fill_group;
  const bool save_timing = t->timing;
  uint64_t* offs = nullptr;
  uint32_t* items = nullptr;
  void* dists = nullptr;
  int rc = kd_tree_set_timing(t, 1);
  if (rc == KD_OK && hipMalloc(&offs, (ms + 1) * sizeof(uint64_t)) != hipSuccess) {
    offs = nullptr;
    rc = kd::set_error(KD_ERR_NO_MEMORY, "hipMalloc failed");
  }
  const int shapes[] = {1, 4, 8, 16, 32};
  // the count pass: each shape twice, the second timed
  float best = 1e30f;
  int best_count = 1;
  for (int g : shapes) {
    if (rc) break;
    t->count_group = g;
    float ms_c = 0.0f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      rc = kd_within_count(t, sample, ms, radius, flags, offs, stream);
      if (rc == KD_OK) rc = kd_last_kernel_ms(t, &ms_c);
    }
    if (rc == KD_OK && ms_c < best) {
      best = ms_c;
      best_count = g;
    }
  }
  // the fill pass, on as many of the sample's queries as 2^29 results allow (with 2^28, a
  // dense workload's fill sample was ~150k queries, too small a launch to rank the shapes:
  // the pick flipped between 8 and 16 lanes, 178 against 277 ms at full size)
  int best_fill = 1;
  uint64_t mf = ms, total = 0;
  std::vector<uint64_t> h(ms + 1);
  if (rc == KD_OK) {
    t->count_group = best_count;
    if (hipMemcpy(h.data(), offs, (ms + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess)
      rc = kd::set_error(KD_ERR_DEVICE, "offsets copy failed");
  }
  if (rc == KD_OK) {
    constexpr uint64_t kBudget = uint64_t(1) << 29;
    mf = uint64_t(std::upper_bound(h.begin(), h.end(), kBudget) - h.begin()) - 1;
    mf = std::max<uint64_t>(std::min<uint64_t>(mf, ms), 1);
    total = h[mf];
    const size_t osz = t->dtype == KD_F32 ? 4 : 8;
    if (hipMalloc(&items, std::max<uint64_t>(total, 1) * 4) != hipSuccess ||
        hipMalloc(&dists, std::max<uint64_t>(total, 1) * osz) != hipSuccess)
      rc = kd::set_error(KD_ERR_NO_MEMORY, "hipMalloc failed");
  }
  // A shape with more lanes per query is taken only when it is more than 10% faster on the
  // sample: 8 and 16 lanes timed within 5% of each other on samples of a dense workload
  // (10M-point Gaussian mixture, r = 0.005) while at full size 16 lanes took 284 ms and 8
  // lanes 195 ms.
  best = 1e30f;
  for (int g : shapes) {
    if (rc) break;
    t->fill_group = g;
    float ms_c = 0.0f;
    for (int rep = 0; rep < 2 && rc == KD_OK; ++rep) {
      rc = kd_within_fill(t, sample, mf, radius, flags, offs, items, dists, stream);
      if (rc == KD_OK) rc = kd_last_kernel_ms(t, &ms_c);
    }
    if (rc == KD_OK && ms_c < 0.9f * best) {
      best = ms_c;
      best_fill = g;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipFree(offs);
  (void)hipFree(items);
  (void)hipFree(dists);
  (void)hipFree(sample_mem);
  t->timing = save_timing;
  if (rc) {
    t->count_group = save_count;
    t->fill_group = save_fill;
    return rc;
  }
  t->count_group = best_count;
  t->fill_group = best_fill;
  if (chosen2) {
    chosen2[0] = best_count;
    chosen2[1] = best_fill;
  }
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
  std::vector<unsigned long long> h(4 * kd::dev::kStatSlots);
  KD_HIP(hipDeviceSynchronize());
  KD_HIP(hipMemcpy(h.data(), t->counters, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost));
  for (int j = 0; j < 4; ++j) out4[j] = 0;
  for (size_t i = 0; i < h.size(); ++i) out4[i % 4] += h[i];
  return KD_OK;
}

int kd_query_redos(kd_tree* t, uint64_t* out) {
  if (int rc = kd::check_tree(t)) return rc;
  if (!out) return kd::set_error(KD_ERR_INVALID, "out is null");
  unsigned long long h = 0;
  KD_HIP(hipDeviceSynchronize());
  KD_HIP(hipMemcpy(&h, t->counters + kd::dev::kRedoWord, sizeof(h), hipMemcpyDeviceToHost));
  *out = h;
  return KD_OK;
}

}  // extern "C"
