// k_knn_f64.hip — nearest_n / nearest_n().within on f64 trees (plain FP64 and FP32-screened
// kernels; knn_launch.h), dispatched from k_knn.hip.

#include "knn_launch.h"

namespace kd {

int launch_nearest_n_f64(const QueryArgs& a, int k, double radius, uint32_t* items, void* dists,
                         uint32_t* counts) {
  switch (a.t->dims) {
    case 2: return knn::launch_typed<double, 2>(a, k, radius, items, dists, counts);
    case 3: return knn::launch_typed<double, 3>(a, k, radius, items, dists, counts);
    case 4: return knn::launch_typed<double, 4>(a, k, radius, items, dists, counts);
    default: return set_error(KD_ERR_UNSUPPORTED, "unsupported dims %d", a.t->dims);
  }
}

}  // namespace kd
