// k_knn.hip — nearest_n / nearest_n().within entry (kiddo query/nearest_n.rs:21-38,
// nearest_n_within.rs:127-179, :441-479): other metrics, f32 and integer trees here; f64
// trees in k_knn_f64.hip (kernels and launchers: knn_launch.h). The split keeps the
// translation units small enough to compile in parallel.

#include "knn_launch.h"

namespace kd {

int launch_nearest_n_f64(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists,
                         uint32_t* counts);

int launch_nearest_n(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists,
                     uint32_t* counts) {
  if (metric_kind(a.flags) != 0)  // other metrics: the thread-per-query kernel (families.h)
    KD_WITH_METRIC(a.flags, mlaunch_nearest_n(a, met, k, radius, items, dists, counts));
  if (a.t->dtype == KD_F64) return launch_nearest_n_f64(a, k, radius, items, dists, counts);
  switch (a.t->dtype * 8 + a.t->dims) {
    case KD_F32 * 8 + 2: return knn::launch_typed<float, 2>(a, k, radius, items, dists, counts);
    case KD_F32 * 8 + 3: return knn::launch_typed<float, 3>(a, k, radius, items, dists, counts);
    case KD_F32 * 8 + 4: return knn::launch_typed<float, 4>(a, k, radius, items, dists, counts);
    case KD_U16 * 8 + 2: return knn::launch_typed<uint16_t, 2>(a, k, radius, items, dists, counts);
    case KD_U16 * 8 + 3: return knn::launch_typed<uint16_t, 3>(a, k, radius, items, dists, counts);
    case KD_U16 * 8 + 4: return knn::launch_typed<uint16_t, 4>(a, k, radius, items, dists, counts);
    case KD_U32 * 8 + 2: return knn::launch_typed<uint32_t, 2>(a, k, radius, items, dists, counts);
    case KD_U32 * 8 + 3: return knn::launch_typed<uint32_t, 3>(a, k, radius, items, dists, counts);
    case KD_U32 * 8 + 4: return knn::launch_typed<uint32_t, 4>(a, k, radius, items, dists, counts);
    default:
      return set_error(KD_ERR_UNSUPPORTED, "unsupported dtype %d / dims %d", a.t->dtype, a.t->dims);
  }
}

}  // namespace kd
