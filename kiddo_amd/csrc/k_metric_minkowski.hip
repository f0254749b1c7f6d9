// k_metric_minkowski.hip — the thread-per-query and periodic kernels of every family under the
// Minkowski metric (kiddo src/dist/minkowski/mod.rs), instantiated in a translation unit of their own.

#include "metric_launch.h"

namespace kd {

KD_METRIC_LAUNCHERS(, dev::Minkowski)

}  // namespace kd
