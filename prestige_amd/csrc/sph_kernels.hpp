// sph_kernels.hpp — WCSPH gather kernels, templated on the precision form.
//
// The pairwise bodies are the textbook WCSPH set (cubic spline of Monaghan &
// Lattanzio 1985, Tait EOS gamma = 7, Monaghan 1992 artificial viscosity)
// run inside the loop shape of prestige/src/codegen/simple_cpu.rs:7-19: i
// outer (one lane per sorted slot), j inner, gather into [i] only, so no
// atomics.
//
// Two passes per step, split where the fused bodies would read rho[j] and
// p[j] that the density body writes:
//   density  (a) scans the 27-cell candidates with a compact FP32 probe of
//            the positions (16 B per particle) and appends every candidate
//            inside the cutoff plus a relative margin of 2^-10 to the
//            particle's neighbour list, in scan order;
//            (b) walks that list in the form's own precision, keeps the
//            exact r < 2h pairs and sums the density, then applies the EOS.
//   force    walks the same list: momentum + viscosity, symplectic Euler and
//            the next step's cell key.
// Scanning in FP32 halves the bytes each candidate costs, and the expensive
// pair bodies only ever run over the compact list, where every lane of a
// wavefront has work, instead of inside the 15 %-hit candidate loop.  The
// list keeps scan order, so FP64 sums are formed in exactly the order of the
// host oracle; it is also the frozen neighbour set the adjoint reuses.
//
// Forms:
//   kF64   state and pair arithmetic in FP64, positions absolute.  The
//          including translation unit is compiled with -ffp-contract=off
//          and every expression keeps the operation order of the host
//          oracle, so the results are bit-identical to it.
//   kF32   state and pair arithmetic in FP32, positions absolute.
//   kMixed FP64 state (positions, velocities, integrator); pair arithmetic
//          in FP32 on the probe itself: each position is an offset from the
//          FP64 origin of its cell, in units of h, and the pair vector is
//          (o_i - o_j) - (c_j - c_i) * cs/h with c_j - c_i in {-1,0,1}^3, so
//          FP32 never holds an absolute coordinate; with cs = 2h every
//          spline constant is exact in FP32.  Density and force sums use
//          Kahan compensation; the EOS and the final scaling of both sums
//          are evaluated in FP64 from the compensated (sum, carry) pair.
#pragma once
#include "common.hpp"

namespace prestige {

enum Form : int { kF64 = 0, kF32 = 1, kMixed = 2 };

template <int F> struct SphForm;
template <> struct SphForm<kF64> {
  using real = double; using G4 = double4; using R2 = double2; using S4 = double4; using sreal = double;
};
template <> struct SphForm<kF32> {
  using real = float; using G4 = float4; using R2 = float2; using S4 = float4; using sreal = float;
};
template <> struct SphForm<kMixed> {
  using real = float; using G4 = float4; using R2 = float2; using S4 = double4; using sreal = double;
};

// Constants of one form, in the units its pair arithmetic uses.
struct SphUnits {
  double H2;       // cutoff^2: (2h)^2, or 4 in units of h
  double inv_h;    // q = r * inv_h (1 in units of h)
  double eps2;     // 0.01 h^2 (0.01 in units of h)
  double hv;       // viscosity mu = hv * (v.r) / (r^2 + eps2): h, or 1 in units of h
  float cs_u;      // cell size in probe units (cs / h)
  float H2_probe;  // probe cutoff^2 in units of h, with margin for forms that recheck
};

// Plain or compensated running sum.
template <class T, bool KAHAN> struct Sum {
  T s = T(0), c = T(0);
  __device__ inline void add(T x) {
    if constexpr (KAHAN) {
      T y = x - c;
      T t = s + y;
      c = (t - s) - y;
      s = t;
    } else {
      s = s + x;
    }
  }
  // value as FP64: s - c carries the compensation into the wider type
  __device__ inline double wide() const {
    if constexpr (KAHAN) return (double)s - (double)c;
    else return (double)s;
  }
};

struct CellRows {
  int cx, cy, cz, x0, x1;
};

__device__ inline CellRows decode_cell(uint32_t key, const Grid& g) {
  const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1];
  CellRows r;
  r.cx = (int)(key % nx);
  uint32_t rest = key / nx;
  r.cy = (int)(rest % ny);
  r.cz = (int)(rest / ny);
  r.x0 = r.cx > 0 ? r.cx - 1 : 0;
  r.x1 = r.cx < (int)nx - 1 ? r.cx + 1 : (int)nx - 1;
  return r;
}

// Visit the nine row runs of a cell: fn(j0, j1, dy, dz).
template <class Fn>
__device__ inline void for_each_run(const CellRows& cr, const Grid& g, const uint32_t* __restrict__ lb,
                                    Fn&& fn) {
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  for (int dz = -1; dz <= 1; ++dz) {
    int zz = cr.cz + dz;
    if (zz < 0 || zz >= nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int yy = cr.cy + dy;
      if (yy < 0 || yy >= ny) continue;
      uint32_t row = ((uint32_t)zz * (uint32_t)ny + (uint32_t)yy) * (uint32_t)nx;
      uint32_t j0 = lb[row + (uint32_t)cr.x0], j1 = lb[row + (uint32_t)cr.x1 + 1];
      fn(j0, j1, dy, dz);
    }
  }
}

// Mixed-form pair vector x_i - x_j from two probes (units of h).
__device__ inline void probe_pair(const float4& qi, const float4& qj, float cs_u, float& dx, float& dy,
                                  float& dz) {
  uint32_t wi = __float_as_uint(qi.w), wj = __float_as_uint(qj.w);
  dx = (qi.x - qj.x) - (float)cell_shift(wi, wj, 0) * cs_u;
  dy = (qi.y - qj.y) - (float)cell_shift(wi, wj, 1) * cs_u;
  dz = (qi.z - qj.z) - (float)cell_shift(wi, wj, 2) * cs_u;
}

template <int F>
__device__ inline void form_pair(const typename SphForm<F>::G4& pi, const typename SphForm<F>::G4& pj,
                                 float cs_u, typename SphForm<F>::real& dx, typename SphForm<F>::real& dy,
                                 typename SphForm<F>::real& dz) {
  if constexpr (F == kMixed) {
    probe_pair(pi, pj, cs_u, dx, dy, dz);
  } else {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
  }
}

// Cubic spline shape f(q) for q < 2 (W = sigma f).
template <class real>
__device__ inline real spline_f(real q) {
  if (q < real(1.0)) return real(1.0) - real(1.5) * q * q + real(0.75) * q * q * q;
  real t = real(2.0) - q;
  return real(0.25) * t * t * t;
}

// df/dq for q < 2.
template <class real>
__device__ inline real spline_df(real q) {
  if (q < real(1.0)) return real(-3.0) * q + real(2.25) * q * q;
  real t = real(2.0) - q;
  return real(-0.75) * t * t;
}

// Scan of the 27-cell candidates of slot i with the FP32 probe; appends hits
// to the k-major list (entry k of slot i at list[k * n + i]).  Returns the
// number of hits (may exceed cap: the caller flags the overflow).
__device__ inline uint32_t probe_scan(uint32_t i, uint32_t n, const float4* __restrict__ probe,
                                      const CellRows& cr, const Grid& g, const uint32_t* __restrict__ lb,
                                      float cs_u, float H2p, uint32_t cap, uint32_t* __restrict__ list) {
  const float4 qi = probe[i];
  const uint32_t wi = __float_as_uint(qi.w);
  uint32_t cnt = 0;
  for_each_run(cr, g, lb, [&](uint32_t j0, uint32_t j1, int ddy, int ddz) {
    const float yi = qi.y - (float)ddy * cs_u;   // row shifts are known per run
    const float zi = qi.z - (float)ddz * cs_u;
    for (uint32_t j = j0; j < j1; ++j) {
      const float4 qj = probe[j];
      float dx = (qi.x - qj.x) - (float)cell_shift(wi, __float_as_uint(qj.w), 0) * cs_u;
      float dy = yi - qj.y;
      float dz = zi - qj.z;
      float r2 = dx * dx + dy * dy + dz * dz;
      if (r2 < H2p) {
        if (cnt < cap) list[(size_t)cnt * n + i] = j;
        ++cnt;
      }
    }
  });
  return cnt;
}

// ---------------------------------------------------------------------------
// density + EOS: rho_i = m sigma sum_j f(|x_ij|/h) over j with |x_ij| < 2h
// (j == i included), p_i = max(B((rho_i/rho0)^7 - 1), 0), pr_i = p_i / rho_i^2.
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(kBlock) void k_sph_density(
    uint32_t n, const float4* __restrict__ probe, const typename SphForm<F>::G4* __restrict__ pos,
    const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ lb, Grid g, SphConst kc, SphUnits u,
    uint32_t cap, uint32_t* __restrict__ list, uint32_t* __restrict__ count,
    typename SphForm<F>::R2* __restrict__ rp, typename SphForm<F>::real* __restrict__ pres,
    uint32_t* flags) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const CellRows cr = decode_cell(skeys[i], g);
  uint32_t cnt = probe_scan(i, n, probe, cr, g, lb, u.cs_u, u.H2_probe, cap, list);
  if (cnt > cap) {
    atomicOr(&flags[FLAG_CAPACITY], 1u);
    cnt = cap;
  }
  count[i] = cnt;
  const G4 pi = pos[i];
  const real H2 = (real)u.H2, inv_h = (real)u.inv_h;
  Sum<real, F == kMixed> acc;
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t j = list[(size_t)k * n + i];
    const G4 pj = pos[j];
    real dx, dy, dz;
    form_pair<F>(pi, pj, u.cs_u, dx, dy, dz);
    real r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < H2) {
      real q = sqrt(r2) * inv_h;
      acc.add(spline_f(q));
    }
  }
  if constexpr (F == kF64) {
    double rho = kc.msig * acc.s;
    double ratio = rho * kc.inv_rho0;
    double r2_ = ratio * ratio;
    double r4 = r2_ * r2_;
    double r7 = r4 * r2_ * ratio;
    double p = kc.B * (r7 - 1.0);
    if (p < 0.0) p = 0.0;
    rp[i] = make_double2(rho, p / (rho * rho));
    pres[i] = p;
  } else if constexpr (F == kF32) {
    float rho = (float)kc.msig * acc.s;
    float ratio = rho * (float)kc.inv_rho0;
    float r2_ = ratio * ratio;
    float r4 = r2_ * r2_;
    float r7 = r4 * r2_ * ratio;
    float p = (float)kc.B * (r7 - 1.0f);
    if (p < 0.0f) p = 0.0f;
    rp[i] = make_float2(rho, p / (rho * rho));
    pres[i] = p;
  } else {
    // EOS in FP64 from the compensated sum: gamma = 7 amplifies the relative
    // error of rho by 7 rho^7 / (rho^7 - rho0^7), so rho is never rounded to
    // FP32 before p is formed.
    double rho = kc.msig * acc.wide();
    double ratio = rho * kc.inv_rho0;
    double r2_ = ratio * ratio;
    double r4 = r2_ * r2_;
    double r7 = r4 * r2_ * ratio;
    double p = kc.B * (r7 - 1.0);
    if (p < 0.0) p = 0.0;
    rp[i] = make_float2((float)rho, (float)(p / (rho * rho)));
    pres[i] = (float)p;
  }
}

// ---------------------------------------------------------------------------
// momentum + artificial viscosity + symplectic Euler + next cell key.
//   a_i = g - m sigma/h sum_j (pr_i + pr_j + Pi_ij) f'(q)/r x_ij,
//   Pi_ij = -alpha c0 mu_ij / (0.5 (rho_i + rho_j)) if v_ij.x_ij < 0,
//   mu_ij = h v_ij.x_ij / (r^2 + 0.01 h^2);
//   v' = v + dt a, x' = x + dt v'.  Fixed particles (id >= n_moving) copy
//   through.  State is read from (spos, svel) and written to (opos, ovel):
//   neighbours keep reading the pre-step state.
// ---------------------------------------------------------------------------
template <int F>
__global__ __launch_bounds__(kBlock) void k_sph_force(
    uint32_t n, uint64_t n_moving, const typename SphForm<F>::G4* __restrict__ gpos,
    const typename SphForm<F>::G4* __restrict__ gvel, const typename SphForm<F>::R2* __restrict__ rp,
    const uint32_t* __restrict__ ids, const uint32_t* __restrict__ skeys,
    const uint32_t* __restrict__ list, const uint32_t* __restrict__ count, Grid g, SphConst kc,
    SphUnits u, const typename SphForm<F>::S4* __restrict__ spos,
    const typename SphForm<F>::S4* __restrict__ svel, typename SphForm<F>::S4* __restrict__ opos,
    typename SphForm<F>::S4* __restrict__ ovel, uint32_t* __restrict__ okeys,
    typename SphForm<F>::S4* __restrict__ oacc, uint32_t* flags) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  using R2 = typename SphForm<F>::R2;
  using S4 = typename SphForm<F>::S4;
  using sreal = typename SphForm<F>::sreal;
  const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if ((uint64_t)ids[i] >= n_moving) {
    opos[i] = spos[i];
    ovel[i] = svel[i];
    okeys[i] = skeys[i];
    if (oacc) oacc[i] = S4{0, 0, 0, 0};
    return;
  }
  const G4 pi = gpos[i];
  const G4 vi = gvel[i];
  const R2 rpi = rp[i];
  const real pri = rpi.y, rhoi = rpi.x;
  const real H2 = (real)u.H2, inv_h = (real)u.inv_h, eps2 = (real)u.eps2, hv = (real)u.hv;
  const real alpha_c0 = (real)kc.alpha_c0;
  Sum<real, F == kMixed> sx, sy, sz;
  const uint32_t cnt = count[i];
  for (uint32_t k = 0; k < cnt; ++k) {
    const uint32_t j = list[(size_t)k * n + i];
    const G4 pj = gpos[j];
    real dx, dy, dz;
    form_pair<F>(pi, pj, u.cs_u, dx, dy, dz);
    real r2 = dx * dx + dy * dy + dz * dz;
    if (r2 < H2 && r2 > real(0.0)) {
      real rr = sqrt(r2);
      real q = rr * inv_h;
      real dw = spline_df(q);
      const G4 vj = gvel[j];
      const R2 rpj = rp[j];
      real dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
      real vdotr = dvx * dx + dvy * dy + dvz * dz;
      real pij = real(0.0);
      if (vdotr < real(0.0)) {
        real muij = hv * vdotr / (r2 + eps2);
        pij = -alpha_c0 * muij / (real(0.5) * (rhoi + rpj.x));
      }
      real sc = (pri + rpj.y + pij) * (dw / rr);
      sx.add(sc * dx);
      sy.add(sc * dy);
      sz.add(sc * dz);
    }
  }
  const S4 xs = spos[i];
  const S4 vs = svel[i];
  sreal ax, ay, az;
  if constexpr (F == kF32) {
    const float msh = (float)kc.msh;
    ax = (float)kc.g[0] - msh * sx.s;
    ay = (float)kc.g[1] - msh * sy.s;
    az = (float)kc.g[2] - msh * sz.s;
  } else {
    ax = kc.g[0] - kc.msh * sx.wide();
    ay = kc.g[1] - kc.msh * sy.wide();
    az = kc.g[2] - kc.msh * sz.wide();
  }
  const sreal dt = (sreal)kc.dt;
  S4 v2 = vs, x2 = xs;
  v2.x = vs.x + dt * ax;
  v2.y = vs.y + dt * ay;
  v2.z = vs.z + dt * az;
  x2.x = xs.x + dt * v2.x;
  x2.y = xs.y + dt * v2.y;
  x2.z = xs.z + dt * v2.z;
  opos[i] = x2;
  ovel[i] = v2;
  if (oacc) oacc[i] = S4{ax, ay, az, 0};
  if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(&flags[FLAG_NONFINITE], 1u);
  // next step's cell key, same FP64 floor as the key kernel
  double fx = floor(((double)x2.x - g.lo[0]) * g.inv_cs);
  double fy = floor(((double)x2.y - g.lo[1]) * g.inv_cs);
  double fz = floor(((double)x2.z - g.lo[2]) * g.inv_cs);
  bool ok = fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)g.dims[0] &&
            fy < (double)g.dims[1] && fz < (double)g.dims[2];
  uint32_t key = 0;
  if (ok)
    key = ((uint32_t)fz * (uint32_t)g.dims[1] + (uint32_t)fy) * (uint32_t)g.dims[0] + (uint32_t)fx;
  else
    atomicOr(&flags[FLAG_DOMAIN], 1u);
  okeys[i] = key;
}

template <int F>
SphUnits sph_units(const Ctx& c) {
  SphUnits u;
  // The scan's r^2 rounds differently from the pair arithmetic that defines
  // the neighbour set (another order for the mixed form, FP32 vs FP64 or
  // absolute FP32 coordinates for the others; all within ~1e-6 relative), so
  // it keeps a 2^-10 relative margin and the list walk applies the exact test.
  const float margin = 1.0f + 1.0f / 1024.0f;
  const double cs_h = c.grid.cs * c.sph.inv_h;
  u.cs_u = (float)cs_h;
  u.H2_probe = 4.0f * margin;
  if constexpr (F == kMixed) {
    u.H2 = 4.0;
    u.inv_h = 1.0;
    u.eps2 = 0.01;
    u.hv = 1.0;
  } else {
    u.H2 = c.sph.H2;
    u.inv_h = c.sph.inv_h;
    u.eps2 = c.sph.eps2;
    u.hv = c.sph.h;
  }
  return u;
}

// Host driver of one WCSPH step in form F: rebuild, density, force.
template <int F>
void sph_step_impl(Ctx& c) {
  using G4 = typename SphForm<F>::G4;
  using R2 = typename SphForm<F>::R2;
  using S4 = typename SphForm<F>::S4;
  using real = typename SphForm<F>::real;
  neighbors_build(c);
  const uint32_t n = (uint32_t)c.n;
  if (!n) return;
  const SphUnits u = sph_units<F>(c);
  const float4* probe = c.off32.as<float4>();
  const G4* gpos = (F == kMixed) ? c.off32.as<G4>() : c.pos[c.cur].as<G4>();
  const G4* gvel = (F == kMixed) ? c.vel32.as<G4>() : c.vel[c.cur].as<G4>();
  const uint32_t* skeys = c.keys[c.cur_key].as<uint32_t>();
  const uint32_t grid = xcd_grid(blocks_for(n));
  const double ge = (double)sizeof(real);   // bytes per scalar of the pair arithmetic
  const double se = (F == kF32) ? 4.0 : 8.0;
  const double lb_bytes = 4.0 * (double)c.grid.n_cells / (double)n;
  const double list_bytes = 4.0 * c.avg_neighbours;   // one u32 per listed neighbour
  {
    // probe 16, key 4, cell table, own position, list write + count, (rho, pr, p) out
    double bytes = (16.0 + 4.0 + lb_bytes + (F == kMixed ? 0.0 : 3.0 * ge) + list_bytes + 4.0 + 3.0 * ge) * n;
    LaunchScope ls(&c, F == kF64 ? "sph_density_f64" : (F == kF32 ? "sph_density_f32" : "sph_density_mixed"),
                   bytes);
    k_sph_density<F><<<grid, kBlock, 0, c.stream>>>(n, probe, gpos, skeys, c.lb.as<uint32_t>(), c.grid, c.sph, u,
                                                    c.list_cap, c.nlist.as<uint32_t>(), c.ncount.as<uint32_t>(),
                                                    c.rp.as<R2>(), c.pres.as<real>(), c.flags.as<uint32_t>());
    PR_HIP(hipGetLastError());
  }
  {
    // per moving particle: list + count, pos 3, vel 3, (rho, pr) 2 of the pair
    // arithmetic, id, state in/out 6 + 6 (+ a 3), key out
    const double nm = (double)c.n_moving;
    double bytes = nm * (list_bytes + 4.0 + 8.0 * ge + 4.0 + (F == kMixed ? 6.0 * se : 0.0) + 6.0 * se + 4.0 +
                         (c.keep_acc ? 3.0 * se : 0.0)) +
                   ((double)n - nm) * (4.0 + 2.0 * 6.0 * se + 8.0);
    LaunchScope ls(&c, F == kF64 ? "sph_force_f64" : (F == kF32 ? "sph_force_f32" : "sph_force_mixed"), bytes);
    const int dst = c.cur ^ 1;
    const int kdst = c.cur_key ^ 1;
    k_sph_force<F><<<grid, kBlock, 0, c.stream>>>(
        n, c.n_moving, gpos, gvel, c.rp.as<R2>(), c.ids[c.cur_ids].as<uint32_t>(), skeys,
        c.nlist.as<uint32_t>(), c.ncount.as<uint32_t>(), c.grid, c.sph, u, c.pos[c.cur].as<S4>(),
        c.vel[c.cur].as<S4>(), c.pos[dst].as<S4>(), c.vel[dst].as<S4>(), c.keys[kdst].as<uint32_t>(),
        c.keep_acc ? c.acc.as<S4>() : nullptr, c.flags.as<uint32_t>());
    PR_HIP(hipGetLastError());
    c.cur = dst;
    c.cur_key = kdst;
    c.sorted = false;
    c.keys_valid = true;
  }
}

}  // namespace prestige
