// sph_f64.hip — FP64 WCSPH step.  Built with -ffp-contract=off: together with
// the operation order of sph_kernels.hpp this makes every FP64 result
// bit-identical to the host oracle on the same input.
#include "sph_kernels.hpp"

namespace prestige {
void sph_step_f64(Ctx& c) { sph_step_impl<kF64>(c); }
}  // namespace prestige
