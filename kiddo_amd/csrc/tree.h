// tree.h — host-side view of a device-resident tree and the kernel launchers that the
// C ABI (api.hip) dispatches to; each family's kernels live in their own translation unit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"
#include "kiddo_amd_internal.h"

#define KD_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ::kd::set_error(KD_ERR_DEVICE, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

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
  double* d_max_abs = nullptr;  // per-dimension max |coordinate| (FP32 screen bound)
  double* d_bbox = nullptr;     // lo[4], scale[4] for Morton keys
  unsigned long long* counters = nullptr;
  double bbox_min[4] = {0, 0, 0, 0}, bbox_max[4] = {0, 0, 0, 0};
  uint64_t device_bytes = 0;
  // launch tuning (kd_tree_tune): k-NN kernel form and LDS-pinned stem levels
  int knn_variant = 0;  // 0 auto, 1 thread per query, 2 group per query
  int group_size = 0;   // 0 auto (smallest of 4/8/16/32/64 holding k)
  int lds_levels = -1;  // -1 auto
  int waves = 0;        // minimum waves per SIMD the k-NN kernels are compiled for (0 = 3)
  int fill_group = 0;   // within fill: 0 automatic, 1 thread per query, else lanes per query
  int count_group = 0;  // within count: 0 or 1 thread per query, else lanes per query
  int best_group = 0;   // best_n_within: 0 automatic, 1 thread per query, else lanes per query
  // optional HIP-event timing of the main query kernel of each call
  bool timing = false;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  bool ev_pending = false;
};

namespace kd {

uint32_t grid_for(uint64_t threads);
void reset_counters(kd_tree* t, hipStream_t s);

// Number of stems (top levels) copied to LDS for a launch.
uint32_t lds_stems_for(const kd_tree* t, size_t a_size);

template <class A, int K> dev::DevTree<A, K> dev_tree(const kd_tree* t) {
  dev::DevTree<A, K> d;
  d.stems = static_cast<const A*>(t->stems);
  d.leaf_start = t->leaf_start;
  d.coords = static_cast<const A*>(t->coords);
  d.items = t->items;
  d.shadow = t->shadow;
  d.lds = nullptr;
  d.lds_n = lds_stems_for(t, sizeof(A));
  d.plane = t->plane;
  d.n = uint32_t(t->size);
  d.nlev = t->max_stem_level + 1;
  return d;
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

struct QueryArgs {
  kd_tree* t;
  const void* queries;
  uint64_t m;
  const uint32_t* order;  // Morton permutation or null
  uint32_t flags;
  hipStream_t s;
  bool dense = false;  // within fill: many results per query (group-per-query fill)
};

constexpr int kMaxK = 64;

// k_build.hip: fills t (dtype, dims, bucket, size set here) from device points
int build_device(kd_tree* t, int dtype, int dims, int bucket, uint64_t n, const void* points,
                 const uint32_t* items, uint32_t flags, hipStream_t s);

// k_knn.hip
int launch_nearest_n(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists,
                     uint32_t* counts);
// k_families.hip
int launch_nearest_one(const QueryArgs& a, uint32_t* items, void* dists);
int launch_approx_nearest_one(const QueryArgs& a, uint32_t* items, void* dists);
int launch_best_n_within(const QueryArgs& a, int n, double radius, uint32_t* items, void* dists,
                         uint32_t* counts);
int launch_within_count(const QueryArgs& a, double radius, uint64_t* counts);
int launch_within_fill(const QueryArgs& a, double radius, const uint64_t* offsets, uint32_t* items,
                       void* dists);
// k_periodic.hip (box: K lengths of the tree's coordinate type, host memory)
int launch_periodic_nearest_one(const QueryArgs& a, const void* box, uint32_t* items, void* dists);
int launch_periodic_nearest_n(const QueryArgs& a, const void* box, int k, double radius, uint32_t* items,
                              void* dists, uint32_t* counts);
int launch_periodic_within_count(const QueryArgs& a, const void* box, double radius, uint64_t* counts);
int launch_periodic_within_fill(const QueryArgs& a, const void* box, double radius, const uint64_t* offsets,
                                uint32_t* items, void* dists);

}  // namespace kd
