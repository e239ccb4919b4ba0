// dem_fast.hip — mixed-precision DEM step: FP32 contact forces and history,
// FP64 integrator (FMA contraction on).
#include "dem_kernels.hpp"

namespace prestige {
void dem_step_mixed(Ctx& c) { dem_step_impl<kMixed>(c); }
}  // namespace prestige
