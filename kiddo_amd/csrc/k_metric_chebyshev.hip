// k_metric_chebyshev.hip — the thread-per-query and periodic kernels of every family under the
// Chebyshev metric (kiddo src/dist/chebyshev/mod.rs), instantiated in a translation unit of their own.

#include "metric_launch.h"

namespace kd {

KD_METRIC_LAUNCHERS(, dev::Chebyshev)

}  // namespace kd
