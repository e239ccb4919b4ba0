// dem_f64.hip — FP64 DEM step, built with -ffp-contract=off (bit-identical to
// the host oracle on the same input).
#include "dem_kernels.hpp"

namespace prestige {
void dem_step_f64(Ctx& c) { dem_step_impl<kF64>(c); }
}  // namespace prestige
