// adjoint.hip — reverse mode of the WCSPH and DEM steps.
#include "common.hpp"

namespace prestige {
void tape_forward(Ctx& c, uint32_t nsteps, uint32_t every) {
  (void)c; (void)nsteps; (void)every;
  throw Error(PRESTIGE_E_UNSUPPORTED, "adjoint not built yet");
}
void adjoint_run(Ctx& c, const double* seed_x, const double* seed_v, double* grad_params,
                 uint32_t n_params, double* grad_v0) {
  (void)c; (void)seed_x; (void)seed_v; (void)grad_params; (void)n_params; (void)grad_v0;
  throw Error(PRESTIGE_E_UNSUPPORTED, "adjoint not built yet");
}
}  // namespace prestige
