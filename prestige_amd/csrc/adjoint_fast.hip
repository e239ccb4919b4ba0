// adjoint_fast.hip — mixed-precision instantiation of the WCSPH tape and
// reverse sweep (FP32 pair geometry, FP64 adjoint sums).
#include "adjoint_dem.hpp"
#include "adjoint_sph.hpp"

namespace prestige {
void sph_tape_forward_mixed(Ctx& c, uint32_t nsteps, uint32_t every) { sph_tape_forward<kMixed>(c, nsteps, every); }
void sph_adjoint_mixed(Ctx& c, const double* sx, const double* sv, double* gp, double* gv) {
  sph_adjoint<kMixed>(c, sx, sv, gp, gv);
}
void dem_tape_forward_mixed(Ctx& c, uint32_t nsteps, uint32_t every) { dem_tape_forward<kMixed>(c, nsteps, every); }
void dem_adjoint_mixed(Ctx& c, const double* sx, const double* sv, double* gp, double* gv) {
  dem_adjoint<kMixed, float>(c, sx, sv, gp, gv);
}
}  // namespace prestige
