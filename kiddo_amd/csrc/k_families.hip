// k_families.hip — entry launchers of nearest_one, nearest_one().approx(), best_n_within
// and the two passes of within (count, then fill), one thread per query. The kernels are in
// families.h; the flags' metric selects the instantiation (SquaredEuclidean here, the
// others in k_metric_*.hip).

#include "metric_launch.h"

namespace kd {

KD_METRIC_LAUNCHERS(, dev::SqEuclidean)

int launch_nearest_one(const QueryArgs& a, uint32_t* items, void* dists) {
  KD_WITH_METRIC(a.flags, mlaunch_nearest_one(a, met, items, dists));
}

int launch_approx_nearest_one(const QueryArgs& a, uint32_t* items, void* dists) {
  KD_WITH_METRIC(a.flags, mlaunch_approx_nearest_one(a, met, items, dists));
}

int launch_best_n_within(const QueryArgs& a, int n, double radius, uint32_t* items, void* dists,
                         uint32_t* counts) {
  KD_WITH_METRIC(a.flags, mlaunch_best_n_within(a, met, n, radius, items, dists, counts));
}

int launch_within_count(const QueryArgs& a, double radius, uint64_t* counts) {
  KD_WITH_METRIC(a.flags, mlaunch_within_count(a, met, radius, counts));
}

int launch_within_fill(const QueryArgs& a, double radius, const uint64_t* offsets, uint32_t* items,
                       void* dists) {
  KD_WITH_METRIC(a.flags, mlaunch_within_fill(a, met, radius, offsets, items, dists));
}

// periodic boundary conditions (query/periodic.rs): kernels in metric_launch.h
int launch_periodic_nearest_one(const QueryArgs& a, const void* box, uint32_t* items, void* dists) {
  KD_WITH_METRIC(a.flags, mlaunch_periodic_nearest_one(a, met, box, items, dists));
}

int launch_periodic_nearest_n(const QueryArgs& a, const void* box, int k, double radius, uint32_t* items,
                              void* dists, uint32_t* counts) {
  KD_WITH_METRIC(a.flags, mlaunch_periodic_nearest_n(a, met, box, k, radius, items, dists, counts));
}

int launch_periodic_within_count(const QueryArgs& a, const void* box, double radius, uint64_t* counts) {
  KD_WITH_METRIC(a.flags, mlaunch_periodic_within_count(a, met, box, radius, counts));
}

int launch_periodic_within_fill(const QueryArgs& a, const void* box, double radius, const uint64_t* offsets,
                                uint32_t* items, void* dists) {
  KD_WITH_METRIC(a.flags, mlaunch_periodic_within_fill(a, met, box, radius, offsets, items, dists));
}

}  // namespace kd
