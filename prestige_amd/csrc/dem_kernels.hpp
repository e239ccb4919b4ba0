// dem_kernels.hpp — DEM contact gather kernel, templated on the precision form.
//
// Bodies: linear spring-dashpot normal force with cn = 2 zeta sqrt(m_eff kn),
// zeta = -ln e / sqrt(pi^2 + ln^2 e); tangential spring xi (projected onto the
// current tangent plane, advanced by v_t dt, Coulomb-capped |Ft| <= mu |Fn|,
// reset when the contact opens); torque (-r_i n) x Ft; half-space walls with
// m_eff = m (Cundall & Strack 1979).  Loop shape of
// prestige/src/codegen/simple_cpu.rs:7-19 over the 27-cell iterator, j != i,
// gather into [i].  The tangential history of particle i lives at index
// id(i) of SoA slot arrays, so the sort never moves it; each step reads the
// previous buffer and writes the next (double-buffered), which is why a
// gather needs no atomics.
//
// kF64: all FP64 (bit-identical to the host oracle; TU built with
// -ffp-contract=off).  kMixed: pair geometry from FP32 offsets to the FP64
// cell origin (metres), forces and history in FP32, walls' distances and the
// integrator in FP64.
#pragma once
#include "sph_kernels.hpp"

namespace prestige {

template <int F> struct DemForm;
template <> struct DemForm<kF64> { using real = double; using G4 = double4; };
template <> struct DemForm<kMixed> { using real = float; using G4 = float4; };

// Tangential spring update (shared by pair and wall contacts); true when the
// spring is at the Coulomb cap (sliding).
template <class real>
__device__ inline bool dem_tangential(real fn, real nx, real ny, real nz, real vtx, real vty, real vtz,
                                      real xox, real xoy, real xoz, real kt, real mu, real neg_inv_kt,
                                      real dt, real* ft, real* xn) {
  real xdn = xox * nx + xoy * ny + xoz * nz;
  real xpx = xox - xdn * nx, xpy = xoy - xdn * ny, xpz = xoz - xdn * nz;
  real xnx = xpx + vtx * dt, xny = xpy + vty * dt, xnz = xpz + vtz * dt;
  real ftx = -kt * xnx, fty = -kt * xny, ftz = -kt * xnz;
  real ftn = sqrt(ftx * ftx + fty * fty + ftz * ftz);
  real fcap = mu * (fn < real(0.0) ? -fn : fn);
  const bool slip = ftn > fcap;
  if (slip) {
    real sc = fcap / ftn;
    ftx = ftx * sc; fty = fty * sc; ftz = ftz * sc;
    xnx = ftx * neg_inv_kt; xny = fty * neg_inv_kt; xnz = ftz * neg_inv_kt;
  }
  ft[0] = ftx; ft[1] = fty; ft[2] = ftz;
  xn[0] = xnx; xn[1] = xny; xn[2] = xnz;
  return slip;
}

struct DemHist {
  const uint32_t* cnt_in; const uint32_t* id_in; const void* xi_in; const void* wxi_in;
  uint32_t* cnt_out; uint32_t* id_out; void* xi_out; void* wxi_out;
};

template <int F>
__global__ __launch_bounds__(kBlock) void k_dem_force(
    uint32_t n, uint64_t n_moving, const typename DemForm<F>::G4* __restrict__ gpos,
    const typename DemForm<F>::G4* __restrict__ gvel, const typename DemForm<F>::G4* __restrict__ gom,
    const uint32_t* __restrict__ ids, const uint32_t* __restrict__ skeys,
    const uint32_t* __restrict__ lb, Grid g, DemConst kc, const double4* __restrict__ spos,
    const double4* __restrict__ svel, const double4* __restrict__ som, double4* __restrict__ opos,
    double4* __restrict__ ovel, double4* __restrict__ oom, uint32_t* __restrict__ okeys,
    double4* __restrict__ oforce, DemHist hist, MovedOut moved, uint32_t* flags,
    uint32_t* __restrict__ events, uint32_t ev_weight) {
  using real = typename DemForm<F>::real;
  using G4 = typename DemForm<F>::G4;
  const uint32_t i = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t me = ids[i];
  const uint32_t mc = kc.max_contacts;
  const real* xi_in = (const real*)hist.xi_in;
  const real* wxi_in = (const real*)hist.wxi_in;
  real* xi_out = (real*)hist.xi_out;
  real* wxi_out = (real*)hist.wxi_out;
  const uint32_t key_i = skeys[i];
  const uint64_t live_mask = __ballot((uint64_t)me < n_moving);
  if ((uint64_t)me >= n_moving) {
    opos[i] = spos[i]; ovel[i] = svel[i]; oom[i] = som[i]; okeys[i] = key_i;
    if (oforce) oforce[i] = make_double4(0, 0, 0, 0);
    hist.cnt_out[me] = 0;
    for (int c = 0; c < 18; ++c) wxi_out[(size_t)c * n + me] = real(0.0);
    moved_word_fixed(moved, i, live_mask);
    return;
  }
  const CellRows cr = decode_cell(key_i, g);
  const G4 pi = gpos[i], vi = gvel[i], wi = gom[i];
  // radius / mass: FP64 form keeps them in pos.w / vel.w; mixed in om.w / vel.w
  const real ri = (F == kF64) ? (real)pi.w : (real)wi.w;
  const real mi = (real)vi.w;
  const real kn = (real)kc.kn, kt = (real)kc.kt, mu = (real)kc.mu, two_zeta = (real)kc.two_zeta,
             neg_inv_kt = (real)kc.neg_inv_kt, dt = (real)kc.dt;
  const float cs_u = (float)g.cs;
  real Fx = 0, Fy = 0, Fz = 0, Tx = 0, Ty = 0, Tz = 0;
  uint32_t nc = 0, nwall = 0, nslip = 0;
  const uint32_t ocnt = hist.cnt_in[me];
  // Candidates of a row run in slot order; their positions are loaded four at
  // a time ahead of the tests, so the loads of a batch overlap (a lane has
  // few contacts, and its time is the candidate loads' latency).
  auto contact = [&](uint32_t j, const G4& pj, const G4& wjb, int ddy, int ddz) {
    {
      real dx, dy, dz;
      if constexpr (F == kMixed) {
        dx = (pi.x - pj.x) - (float)cell_shift(__float_as_uint(pi.w), __float_as_uint(pj.w), 0) * cs_u;
        dy = (pi.y - pj.y) - (float)ddy * cs_u;
        dz = (pi.z - pj.z) - (float)ddz * cs_u;
      } else {
        dx = pi.x - pj.x; dy = pi.y - pj.y; dz = pi.z - pj.z;
      }
      real r2 = dx * dx + dy * dy + dz * dz;
      // radius: FP64 form in pos.w, mixed form in om.w (loaded with the batch)
      const real rj = (F == kF64) ? (real)pj.w : (real)wjb.w;
      real R = ri + rj;
      if (!(r2 < R * R && r2 > real(0.0))) return;
      const G4 wj = (F == kF64) ? gom[j] : wjb;
      const G4 vj = gvel[j];
      real d = sqrt(r2);
      real invd = real(1.0) / d;
      real nx = dx * invd, ny = dy * invd, nz = dz * invd;
      real delta = R - d;
      real wsx = ri * wi.x + rj * wj.x;
      real wsy = ri * wi.y + rj * wj.y;
      real wsz = ri * wi.z + rj * wj.z;
      real cx = wsy * nz - wsz * ny, cy = wsz * nx - wsx * nz, cz = wsx * ny - wsy * nx;
      real vrx = (vi.x - vj.x) - cx, vry = (vi.y - vj.y) - cy, vrz = (vi.z - vj.z) - cz;
      real vn = vrx * nx + vry * ny + vrz * nz;
      real vtx = vrx - vn * nx, vty = vry - vn * ny, vtz = vrz - vn * nz;
      real mj = (real)vj.w;
      real meff = mi * mj / (mi + mj);
      real cn = two_zeta * sqrt(meff * kn);
      real fn = kn * delta - cn * vn;
      const uint32_t pid = ids[j];
      real xox = 0, xoy = 0, xoz = 0;
      for (uint32_t c = 0; c < ocnt; ++c) {
        if (hist.id_in[(size_t)c * n + me] == pid) {
          xox = xi_in[(size_t)(c * 3 + 0) * n + me];
          xoy = xi_in[(size_t)(c * 3 + 1) * n + me];
          xoz = xi_in[(size_t)(c * 3 + 2) * n + me];
          break;
        }
      }
      real ft[3], xn[3];
      nslip += dem_tangential<real>(fn, nx, ny, nz, vtx, vty, vtz, xox, xoy, xoz, kt, mu, neg_inv_kt, dt, ft, xn);
      Fx = Fx + (fn * nx + ft[0]);
      Fy = Fy + (fn * ny + ft[1]);
      Fz = Fz + (fn * nz + ft[2]);
      real tx = ny * ft[2] - nz * ft[1], ty = nz * ft[0] - nx * ft[2], tz = nx * ft[1] - ny * ft[0];
      Tx = Tx - ri * tx; Ty = Ty - ri * ty; Tz = Tz - ri * tz;
      if (nc < mc) {
        hist.id_out[(size_t)nc * n + me] = pid;
        xi_out[(size_t)(nc * 3 + 0) * n + me] = xn[0];
        xi_out[(size_t)(nc * 3 + 1) * n + me] = xn[1];
        xi_out[(size_t)(nc * 3 + 2) * n + me] = xn[2];
      }
      ++nc;
    }
  };
  for_each_run(cr, g, lb, [&](uint32_t j0, uint32_t j1, int ddy, int ddz) {
    for (uint32_t jb = j0; jb < j1; jb += 4u) {
      G4 pjs[4], wjs[4];
#pragma unroll
      for (uint32_t u = 0; u < 4u; ++u) {
        const uint32_t jc = jb + u < j1 ? jb + u : j1 - 1u;
        pjs[u] = gpos[jc];
        if constexpr (F == kMixed) wjs[u] = gom[jc];
        else wjs[u] = pjs[u];   // unused: the FP64 radius is pos.w
      }
#pragma unroll
      for (uint32_t u = 0; u < 4u; ++u) {
        const uint32_t j = jb + u;
        if (j >= j1) break;
        if (j != i) contact(j, pjs[u], wjs[u], ddy, ddz);
      }
    }
  });
  // walls: bit 2a = lower wall of axis a (normal +e_a), bit 2a+1 = upper (-e_a)
  const double4 xs = spos[i];
  for (int w = 0; w < 6; ++w) {
    real delta = real(-1.0);
    if (kc.wall_mask & (1u << w)) {
      int a = w >> 1;
      if constexpr (F == kF64) {
        double pa = a == 0 ? xs.x : (a == 1 ? xs.y : xs.z);
        double dist = (w & 1) ? (kc.wall_hi[a] - pa) : (pa - kc.wall_lo[a]);
        delta = ri - dist;
      } else {
        double pa = a == 0 ? xs.x : (a == 1 ? xs.y : xs.z);
        double dist = (w & 1) ? (kc.wall_hi[a] - pa) : (pa - kc.wall_lo[a]);
        delta = (real)(xs.w - dist);
      }
    }
    if (!(delta > real(0.0))) {
      wxi_out[(size_t)(w * 3 + 0) * n + me] = real(0.0);
      wxi_out[(size_t)(w * 3 + 1) * n + me] = real(0.0);
      wxi_out[(size_t)(w * 3 + 2) * n + me] = real(0.0);
      continue;
    }
    int a = w >> 1;
    real sg = (w & 1) ? real(-1.0) : real(1.0);
    real nx = a == 0 ? sg : real(0.0), ny = a == 1 ? sg : real(0.0), nz = a == 2 ? sg : real(0.0);
    real wsx = ri * wi.x, wsy = ri * wi.y, wsz = ri * wi.z;
    real cx = wsy * nz - wsz * ny, cy = wsz * nx - wsx * nz, cz = wsx * ny - wsy * nx;
    real vrx = vi.x - cx, vry = vi.y - cy, vrz = vi.z - cz;
    real vn = vrx * nx + vry * ny + vrz * nz;
    real vtx = vrx - vn * nx, vty = vry - vn * ny, vtz = vrz - vn * nz;
    real cn = two_zeta * sqrt(mi * kn);
    real fn = kn * delta - cn * vn;
    real xox = wxi_in[(size_t)(w * 3 + 0) * n + me];
    real xoy = wxi_in[(size_t)(w * 3 + 1) * n + me];
    real xoz = wxi_in[(size_t)(w * 3 + 2) * n + me];
    real ft[3], xn[3];
    nslip += dem_tangential<real>(fn, nx, ny, nz, vtx, vty, vtz, xox, xoy, xoz, kt, mu, neg_inv_kt, dt, ft, xn);
    ++nwall;
    Fx = Fx + (fn * nx + ft[0]);
    Fy = Fy + (fn * ny + ft[1]);
    Fz = Fz + (fn * nz + ft[2]);
    real tx = ny * ft[2] - nz * ft[1], ty = nz * ft[0] - nx * ft[2], tz = nx * ft[1] - ny * ft[0];
    Tx = Tx - ri * tx; Ty = Ty - ri * ty; Tz = Tz - ri * tz;
    wxi_out[(size_t)(w * 3 + 0) * n + me] = xn[0];
    wxi_out[(size_t)(w * 3 + 1) * n + me] = xn[1];
    wxi_out[(size_t)(w * 3 + 2) * n + me] = xn[2];
  }
  if (nc > mc) atomicOr(&flags[FLAG_CAPACITY], 1u);
  hist.cnt_out[me] = nc < mc ? nc : mc;
  if (events) {   // step-weighted fingerprints of this particle's contact and slip events
    events[me] += ev_weight * (nc + nwall);
    events[(size_t)n + me] += ev_weight * nslip;
  }
  // symplectic Euler in FP64
  const double4 vs = svel[i], ws = som[i];
  const double m64 = vs.w, r64 = xs.w;
  double axx = (double)Fx / m64 + kc.g[0];
  double ayy = (double)Fy / m64 + kc.g[1];
  double azz = (double)Fz / m64 + kc.g[2];
  double4 v2 = vs, x2 = xs, w2 = ws;
  v2.x = vs.x + kc.dt * axx;
  v2.y = vs.y + kc.dt * ayy;
  v2.z = vs.z + kc.dt * azz;
  x2.x = xs.x + kc.dt * v2.x;
  x2.y = xs.y + kc.dt * v2.y;
  x2.z = xs.z + kc.dt * v2.z;
  double inertia = 0.4 * m64 * r64 * r64;
  w2.x = ws.x + kc.dt * ((double)Tx / inertia);
  w2.y = ws.y + kc.dt * ((double)Ty / inertia);
  w2.z = ws.z + kc.dt * ((double)Tz / inertia);
  opos[i] = x2;
  ovel[i] = v2;
  oom[i] = w2;
  if (oforce) oforce[i] = make_double4(axx, ayy, azz, 0.0);
  if (!isfinite(axx) || !isfinite(ayy) || !isfinite(azz)) atomicOr(&flags[FLAG_NONFINITE], 1u);
  if ((v2.x * v2.x + v2.y * v2.y + v2.z * v2.z) * (kc.dt * kc.dt) > 0.0625 * (r64 * r64))
    atomicOr(&flags[FLAG_CFL], 1u);
  double fx = floor((x2.x - g.lo[0]) * g.inv_cs);
  double fy = floor((x2.y - g.lo[1]) * g.inv_cs);
  double fz = floor((x2.z - g.lo[2]) * g.inv_cs);
  bool ok = fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)g.dims[0] &&
            fy < (double)g.dims[1] && fz < (double)g.dims[2];
  uint32_t key = 0;
  if (ok)
    key = ((uint32_t)fz * (uint32_t)g.dims[1] + (uint32_t)fy) * (uint32_t)g.dims[0] + (uint32_t)fx;
  else
    atomicOr(&flags[FLAG_DOMAIN], 1u);
  okeys[i] = key;
  moved_word_live(moved, i, key != key_i, live_mask);
}

template <int F>
void dem_step_impl(Ctx& c) {
  using G4 = typename DemForm<F>::G4;
  using real = typename DemForm<F>::real;
  neighbors_build(c);
  const uint32_t n = (uint32_t)c.n;
  if (!n) return;
  const G4* gpos = (F == kMixed) ? c.off32.as<G4>() : c.pos[c.cur].as<G4>();
  const G4* gvel = (F == kMixed) ? c.vel32.as<G4>() : c.vel[c.cur].as<G4>();
  const G4* gom = (F == kMixed) ? c.om32.as<G4>() : c.omega[c.cur].as<G4>();
  const int h = c.cur_hist;
  DemHist hist{c.hcnt[h].as<uint32_t>(), c.hid[h].as<uint32_t>(), c.hxi[h].p, c.wxi[h].p,
               c.hcnt[h ^ 1].as<uint32_t>(), c.hid[h ^ 1].as<uint32_t>(), c.hxi[h ^ 1].p, c.wxi[h ^ 1].p};
  // one wavefront per block: small scenes (the 10k lattice) spread over more CUs
  const uint32_t grid = xcd_grid(blocks_for(n, 64));
  const double ge = (double)sizeof(real);
  const double lb_bytes = 4.0 * (double)c.grid.n_cells / (double)n;
  // gathered: pos 3 + r, vel 3 + m, omega 3; ids + key; state in/out (pos, vel, omega) 3 x 24 x 2;
  // history in/out: count + (typical) 6 slots x (id + 3 reals) + walls 18 reals
  double hist_bytes = 2.0 * (4.0 + 6.0 * (4.0 + 3.0 * ge) + 18.0 * ge);
  double bytes = (11.0 * ge + 8.0 + lb_bytes + 2.0 * 3.0 * 32.0 + 4.0 + hist_bytes) * n;
  LaunchScope ls(&c, F == kF64 ? "dem_force_f64" : "dem_force_mixed", bytes);
  const int dst = c.cur ^ 1, kdst = c.cur_key ^ 1;
  k_dem_force<F><<<grid, 64, 0, c.stream>>>(
      n, c.n_moving, gpos, gvel, gom, c.ids[c.cur_ids].as<uint32_t>(), c.keys[c.cur_key].as<uint32_t>(),
      c.lb.as<uint32_t>(), c.grid, c.dem, c.pos[c.cur].as<double4>(), c.vel[c.cur].as<double4>(),
      c.omega[c.cur].as<double4>(), c.pos[dst].as<double4>(), c.vel[dst].as<double4>(),
      c.omega[dst].as<double4>(), c.keys[kdst].as<uint32_t>(), c.keep_acc ? c.acc.as<double4>() : nullptr,
      hist, MovedOut{c.inc_mask.as<uint64_t>(), c.inc_tiles.as<uint32_t>()}, c.flags.as<uint32_t>(),
      c.ev_weight ? c.events.as<uint32_t>() : nullptr, c.ev_weight);
  PR_HIP(hipGetLastError());
  c.cur = dst;
  c.cur_key = kdst;
  c.cur_hist ^= 1;
  c.sorted = false;
  c.keys_valid = true;
  c.inc_ready = true;   // keys[cur_key ^ 1] and lb still describe the slots the step kept
}

}  // namespace prestige
