// sph_f64.hip — FP64 WCSPH step.  Built with -ffp-contract=off: together with
// the operation order of sph_kernels.hpp this makes every FP64 result
// bit-identical to the host oracle on the same input.
#include "sph_kernels.hpp"

namespace prestige {
void sph_step_f64(Ctx& c) { sph_step_impl<kF64>(c); }

// FP32 copies of the FP64 density outputs for the mixed force pass.
__global__ __launch_bounds__(kBlock) void k_narrow_density(uint32_t n, const double2* __restrict__ rp64,
                                                           const double* __restrict__ p64,
                                                           float2* __restrict__ rp32, float* __restrict__ p32) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double2 r = rp64[i];
    rp32[i] = make_float2((float)r.x, (float)r.y);
    p32[i] = (float)p64[i];
  }
}

// Density pass of the mixed form with PRESTIGE_FLAG_DENSITY_F64: the F64
// form's kernel on the mixed context's FP64 state (the same positions, the
// same operations, so the same bits as the F64 form), then FP32 copies of
// (rho, p/rho^2) and p for the mixed force pass.
void sph_density_f64_for_mixed(Ctx& c) {
  const SphStepIO<kF64> io = sph_step_io<kF64>(c);
  sph_density_launch<kF64>(c, io, c.rp64.as<double2>(), c.pres64.as<double>(), c.pos[c.cur].as<double4>(),
                           c.vel[c.cur].as<double4>());
  const uint32_t n = (uint32_t)c.n;
  LaunchScope ls(&c, "narrow_density", 36.0 * n);
  k_narrow_density<<<blocks_for(n), kBlock, 0, c.stream>>>(n, c.rp64.as<double2>(), c.pres64.as<double>(),
                                                           c.rp.as<float2>(), c.pres.as<float>());
  PR_HIP(hipGetLastError());
}
}  // namespace prestige
