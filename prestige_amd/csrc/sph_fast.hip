// sph_fast.hip — FP32 and mixed-precision WCSPH steps (FMA contraction on).
#include "sph_kernels.hpp"

namespace prestige {
void sph_step_f32(Ctx& c) { sph_step_impl<kF32>(c); }
void sph_step_mixed(Ctx& c) { sph_step_impl<kMixed>(c); }
}  // namespace prestige
