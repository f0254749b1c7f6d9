// k_metric_manhattan.hip — the thread-per-query and periodic kernels of every family under the
// Manhattan metric (kiddo src/dist/manhattan/mod.rs), instantiated in a translation unit of their own.

#include "metric_launch.h"

namespace kd {

KD_METRIC_LAUNCHERS(, dev::Manhattan)

}  // namespace kd
