// adjoint_sph.hpp — reverse mode of one WCSPH step, gather form.
//
// Forward step (sph_kernels.hpp), per moving particle k with d = x_k - x_j,
// e = v_k - v_j, r = |d|, q = r/h over its frozen neighbour masks:
//   rho_k = m sigma sum_j f(q),  p_k = max(B((rho_k/rho0)^7 - 1), 0),
//   P_k = p_k / rho_k^2,  B = rho0 c0^2 / 7,
//   a_k = g - (m sigma/h) sum_j S_kj d,  S = Q g(r),  Q = P_k + P_j + Pi,
//   g(r) = f'(q)/r,  Pi = -alpha c0 mu / rhobar if w = e.d < 0 (else 0),
//   mu = h w / D,  D = r^2 + 0.01 h^2,  rhobar = (rho_k + rho_j)/2,
//   v' = v + dt a,  x' = x + dt v'.
// With c_k = -(m sigma/h) abar_k, the loss depends on the pair (k, j) through
// c_k . S d (row k) and c_j . S (-d) (row j).  S, g, Pi, mu and w are
// symmetric in the pair, so both rows collapse into one body per neighbour
// with C = c_k - c_j and U = C . d:
//   xbar_k += S C + U Q g'(r) d/r + mubar h (e/D - 2 w d/D^2),
//   vbar_k += mubar h d/D,   mubar = U g (-alpha c0 / rhobar)      (w < 0)
//   Pbar_k += U g,           rhobar_k += U g alpha c0 mu / (2 rhobar^2)
// and each particle gathers only from its own neighbour set: no atomics.
// Parameter sums take row k only (each ordered pair once):
//   dalpha += (c_k.d) g (-c0 mu/rhobar),  dc0 += (c_k.d) g (-alpha mu/rhobar).
// EOS: rhobar_k += Pbar_k dP/drho,  dc0 += Pbar_k dP/dc0 (zero where p was
// clamped).  Density: xbar_k += (m sigma/h) sum_j (rhobar_k + rhobar_j) f'(q) d/r.
// Integrator: vbt = vbar' + dt xbar',  abar = dt vbt,  xbar = xbar' + ...,
// vbar = vbt + ... .  dt is a fixed input (not differentiated through c0).
//
// The reverse kernels use the same pair geometry as the forward form: FP64
// absolute positions for kF64, FP32 probe offsets in units of h for kMixed
// (position adjoints are accumulated in those units and scaled by 1/h).
// Adjoint sums are FP64 in every form.
#pragma once
#include "sph_kernels.hpp"

namespace prestige {

// What one recorded forward step left behind (slot order of that step).
struct SphStepRecord {
  const void* gpos;       // pair geometry: double4 positions (F64) / float4 probes (mixed)
  const void* gvel;       // double4 / float4 velocities
  const void* rp;         // (rho, p/rho^2): double2 / float2
  const uint32_t* ids;
  const uint32_t* perm;   // slot k of this step <- slot perm[k] of the previous order
  const uint2* meta;       // row-run masks of the step
  const uint32_t* words;
};

template <class real>
__device__ inline real spline_d2f(real q) {
  if (q < real(1.0)) return real(-3.0) + real(4.5) * q;
  return real(1.5) * (real(2.0) - q);
}

// Element types of the reverse sweep per form: pair scalars (ar) and the
// per-particle seeds gathered from neighbours (C4 for c, ar for rhobar).  The
// mixed form keeps them in FP32 like its forward; every sum is FP64.
template <int F> struct AdjForm { using ar = double; using C4 = double4; };
template <> struct AdjForm<kMixed> { using ar = float; using C4 = float4; };

// Reciprocal and 1/sqrt of the reverse pair scalars: hardware estimates
// (about 1 ulp) in FP32, where the mixed form's gradient tolerance dwarfs
// them; IEEE operations in FP64.
template <class T>
__device__ inline T adj_rcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return T(1.0) / x;
}

template <class T>
__device__ inline void adj_len(T r2, T& rr, T& irr) {
  if constexpr (sizeof(T) == 4) {
    irr = __builtin_amdgcn_rsqf(r2);
    rr = r2 * irr;
  } else {
    rr = sqrt(r2);
    irr = T(1.0) / rr;
  }
}

// Reverse of the symplectic-Euler update for one slot: vbt = vbar' + dt xbar',
// c = -(m sigma/h) dt vbt; fixed particles get zero.
template <int F>
__global__ __launch_bounds__(kBlock) void k_adj_prep(uint32_t n, uint64_t n_moving,
                                                     const uint32_t* __restrict__ ids,
                                                     const double4* __restrict__ xb,
                                                     double4* __restrict__ vb,
                                                     typename AdjForm<F>::C4* __restrict__ cvec,
                                                     double dt, double msh) {
  using C4 = typename AdjForm<F>::C4;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if ((uint64_t)ids[k] >= n_moving) {
    vb[k] = make_double4(0, 0, 0, 0);
    cvec[k] = C4{0, 0, 0, 0};
    return;
  }
  double4 x = xb[k], v = vb[k];
  double4 vt = make_double4(v.x + dt * x.x, v.y + dt * x.y, v.z + dt * x.z, 0.0);
  vb[k] = vt;
  double s = -msh * dt;
  cvec[k] = C4{(decltype(C4{}.x))(s * vt.x), (decltype(C4{}.x))(s * vt.y), (decltype(C4{}.x))(s * vt.z), 0};
}

// Block sum of two doubles into partial[slot] (deterministic); W wavefronts
// per block.
template <uint32_t W>
__device__ inline void block_sum2(double a, double b, double2* __restrict__ partial, uint32_t slot) {
  __shared__ double sa[W], sb[W];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, (unsigned)o, 64);
    b += __shfl_down(b, (unsigned)o, 64);
  }
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    sa[w] = a;
    sb[w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0.0, tb = 0.0;
    for (uint32_t k = 0; k < W; ++k) {
      ta += sa[k];
      tb += sb[k];
    }
    partial[slot] = make_double2(ta, tb);
  }
}

// Force + EOS adjoint (all slots; fixed particles collect Pbar and rhobar as
// neighbours of moving ones).  Writes xbar of the force rows (metres), adds
// the viscosity part of vbar, writes rhobar, and block partials of
// (dalpha, dc0).  One wavefront per block; each row run's neighbour data are
// staged in LDS as in the forward force pass.
constexpr int kAdjBlock = 64;
// Wavefronts per SIMD the register budgets are held to: each kernel's
// occupancy before the tile loop, but for the mixed density adjoint, whose
// measured best is one fewer (256 against 263 us at 1M; 8: 273).
template <int F>
constexpr int kAdjForceWaves = F == kF64 ? 2 : 4;
template <int F>
constexpr int kAdjDensityWaves = F == kF64 ? 5 : 6;
// Mask words loaded one run group ahead (walk_runs_staged): in the mixed
// reverse passes (force 380 -> 363, density 256 -> 247 us at 1M); the FP64
// ones, at 2 and 5 wavefronts per SIMD, run slower with it (759 -> 795 and
// 302 -> 309 us).
template <int F>
constexpr bool kAdjPrefetch = F == kMixed;

template <int F>
__global__ __launch_bounds__(kAdjBlock, kAdjForceWaves<F>) void k_adj_force(
    uint32_t n, uint64_t n_moving, SphStepRecord rec, Grid g, SphConst kc, SphUnits u, double alpha, double c0,
    double rho0, const typename AdjForm<F>::C4* __restrict__ cvec, double4* __restrict__ xbf,
    double4* __restrict__ vb, typename AdjForm<F>::ar* __restrict__ rhob, double2* __restrict__ partial,
    TileQueue tq) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  using R2 = typename SphForm<F>::R2;
  using ar = typename AdjForm<F>::ar;
  using C4 = typename AdjForm<F>::C4;
  const G4* gpos = (const G4*)rec.gpos;
  const G4* gvel = (const G4*)rec.gvel;
  const R2* rp = (const R2*)rec.rp;
  __shared__ StageArr<G4, kForceCap> s_p, s_v;
  __shared__ R2 s_rp[kForceCap];
  __shared__ StageArr<C4, kForceCap> s_c;
  const uint32_t lane = threadIdx.x;
  for_each_tile(tq, [&](uint32_t tile) {
    const uint32_t i = tile * kAdjBlock + lane;
    const bool live = i < n;
    double pa = 0.0, pc = 0.0;
    G4 pi{}, vi{};
    R2 rpi{};
    C4 ci{};
    if (live) {
      pi = gpos[i];
      vi = gvel[i];
      rpi = rp[i];
      ci = cvec[i];
    }
    {
      const real H2 = (real)u.H2;
      const ar inv_h = (ar)u.inv_h, eps2 = (ar)u.eps2, hv = (ar)u.hv, ac0 = (ar)kc.alpha_c0;
      const ar rhoi = (ar)rpi.x, pri = (ar)rpi.y;
      const ar cix = (ar)ci.x, ciy = (ar)ci.y, ciz = (ar)ci.z;
      // sums: per row run in the pair precision (a few terms each), then FP64
      double xx = 0, xy = 0, xz = 0, vx = 0, vy = 0, vz = 0, pb = 0, rb = 0, sa = 0;
      ar qxx = 0, qxy = 0, qxz = 0, qvx = 0, qvy = 0, qvz = 0, qpb = 0, qrb = 0, qsa = 0;
      auto end_run = [&] {
        xx += (double)qxx; xy += (double)qxy; xz += (double)qxz;
        vx += (double)qvx; vy += (double)qvy; vz += (double)qvz;
        pb += (double)qpb; rb += (double)qrb; sa += (double)qsa;
        qxx = qxy = qxz = qvx = qvy = qvz = qpb = qrb = qsa = ar(0.0);
      };
      ProbeRunBases pb_;   // mixed form: the forward force pass's pair vector
      // pair rows over a batch of neighbours fetched by get(j, p, v, rp, c)
      auto body = [&](const uint32_t* jj, const uint32_t* tt, auto&& get) {
        G4 pj[kUnr], vj[kUnr];
        R2 rpj[kUnr];
        C4 cj[kUnr];
  #pragma unroll
        for (int v = 0; v < kUnr; ++v) get(jj[v] == kPad ? jj[0] : jj[v], pj[v], vj[v], rpj[v], cj[v]);
  #pragma unroll
        for (int v = 0; v < kUnr; ++v) {
          if (jj[v] == kPad) continue;
          real dxr, dyr, dzr;
          if constexpr (F == kMixed) pb_.pair(pi, pj[v], tt[v], u.cs_u, dxr, dyr, dzr);
          else form_pair<F>(pi, pj[v], u.cs_u, dxr, dyr, dzr);
          real r2r = dot3(dxr, dyr, dzr, dxr, dyr, dzr);
          if (!(r2r < H2 && r2r > real(0.0))) continue;
          const ar dx = (ar)dxr, dy = (ar)dyr, dz = (ar)dzr, r2 = (ar)r2r;
          ar rr, irr;
          adj_len(r2, rr, irr);
          const ar q = rr * inv_h;
          const ar f1 = spline_df(q), f2 = spline_d2f(q);
          const ar gr = f1 * irr;
          const ar gp = (f2 * inv_h - gr) * irr;   // d(f'(q)/r)/dr
          const ar ex = (ar)vi.x - (ar)vj[v].x, ey = (ar)vi.y - (ar)vj[v].y, ez = (ar)vi.z - (ar)vj[v].z;
          // the forward pass's branch: v.r in its operation order
          const ar w = dot3(ex, ey, ez, dx, dy, dz);
          const ar D = r2 + eps2, iD = adj_rcp(D);
          const ar rbar = ar(0.5) * (rhoi + (ar)rpj[v].x), irb = adj_rcp(rbar);
          ar mu = ar(0.0), Pi = ar(0.0);
          if (w < ar(0.0)) {
            mu = hv * w * iD;
            Pi = -ac0 * mu * irb;
          }
          const ar Q = pri + (ar)rpj[v].y + Pi;
          const ar S = Q * gr;
          const ar Cx = cix - (ar)cj[v].x, Cy = ciy - (ar)cj[v].y, Cz = ciz - (ar)cj[v].z;
          const ar U = Cx * dx + Cy * dy + Cz * dz;
          const ar Ug = U * gr;
          const ar t = U * Q * gp * irr;
          qxx += S * Cx + t * dx;
          qxy += S * Cy + t * dy;
          qxz += S * Cz + t * dz;
          qpb += Ug;
          if (w < ar(0.0)) {
            const ar mub = -Ug * ac0 * irb;
            const ar hD = hv * iD, k2 = ar(2.0) * hv * w * iD * iD;
            qxx += mub * (hD * ex - k2 * dx);
            qxy += mub * (hD * ey - k2 * dy);
            qxz += mub * (hD * ez - k2 * dz);
            qvx += mub * hD * dx;
            qvy += mub * hD * dy;
            qvz += mub * hD * dz;
            qrb += Ug * ac0 * mu * ar(0.5) * irb * irb;
            // d/dalpha = c0 sum, d/dc0 = alpha sum (Pi is linear in alpha c0)
            qsa -= (cix * dx + ciy * dy + ciz * dz) * gr * mu * irb;
          }
        }
      };
      walk_runs_staged<kUnr, kForceCap, kAdjPrefetch<F>>(
          rec.meta, rec.words, n, i, live, lane,
          [&](uint32_t grp) {
            if constexpr (F == kMixed) pb_.begin(pi, grp, u.cs_u);
          },
          [&](uint32_t k, uint32_t j) {
            s_p.put(k, gpos[j]);
            s_v.put(k, gvel[j]);
            s_rp[k] = rp[j];
            s_c.put(k, cvec[j]);
          },
          [&](const uint32_t* jj, const uint32_t* tt) {
            body(jj, tt, [&](uint32_t k, G4& p, G4& v, R2& r, C4& cc) {
              p = s_p.get(k);
              v = s_v.get(k);
              r = s_rp[k];
              cc = s_c.get(k);
            });
          },
          [&](const uint32_t* jj, const uint32_t* tt) {
            body(jj, tt, [&](uint32_t j, G4& p, G4& v, R2& r, C4& cc) {
              p = gpos[j];
              v = gvel[j];
              r = rp[j];
              cc = cvec[j];
            });
          },
          end_run);
      if (live) {
        pa = sa * c0;
        pc = sa * alpha;
        // EOS: P = p / rho^2, p = B((rho/rho0)^7 - 1) unless the forward clamped
        // it to 0 (then P = 0 and the EOS has no derivative)
        const double rho = (double)rpi.x;
        const double ratio = rho * kc.inv_rho0;
        const double r6 = ratio * ratio * ratio * ratio * ratio * ratio;
        const double p = (double)rpi.y * rho * rho;
        if ((double)rpi.y > 0.0) {
          const double dpdrho = 7.0 * kc.B * r6 * kc.inv_rho0;
          const double dPdrho = dpdrho / (rho * rho) - 2.0 * p / (rho * rho * rho);
          rb += pb * dPdrho;
          const double dPdc0 = (r6 * ratio - 1.0) * (2.0 * rho0 * c0 / 7.0) / (rho * rho);
          pc += pb * dPdc0;
        }
        rhob[i] = (ar)rb;
        const double us = u.unit_scale;   // pair-geometry units -> metres for position adjoints
        xbf[i] = make_double4(us * xx, us * xy, us * xz, 0.0);
        if ((uint64_t)rec.ids[i] < n_moving) {
          double4 v0 = vb[i];
          vb[i] = make_double4(v0.x + vx, v0.y + vy, v0.z + vz, 0.0);
        }
      }
    }
    block_sum2<kAdjBlock / 64>(pa, pc, partial, tile);   // one partial per tile: a fixed sum order
  });
}

// Density adjoint + assembly: xbar_k = xbar'_k + force part + density part
// (moving particles; fixed ones 0), then both adjoints move to the previous
// step's slot order through the frozen permutation.  One wavefront per
// block, neighbour data staged per row run.
template <int F>
__global__ __launch_bounds__(kAdjBlock, kAdjDensityWaves<F>) void k_adj_density(
    uint32_t n, uint64_t n_moving, SphStepRecord rec, Grid g, SphConst kc, SphUnits u,
    const typename AdjForm<F>::ar* __restrict__ rhob, const double4* __restrict__ xb_in,
    const double4* __restrict__ xbf, const double4* __restrict__ vb_in, double4* __restrict__ xb_out,
    double4* __restrict__ vb_out, TileQueue tq) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  using ar = typename AdjForm<F>::ar;
  const G4* gpos = (const G4*)rec.gpos;
  __shared__ StageArr<G4, kForceCap> s_p;
  __shared__ ar s_rb[kForceCap];
  const uint32_t lane = threadIdx.x;
  for_each_tile(tq, [&](uint32_t tile) {
    const uint32_t i = tile * kAdjBlock + lane;
    const bool live = i < n && (uint64_t)rec.ids[i] < n_moving;
    G4 pi{};
    ar rbi = ar(0.0);
    if (live) {
      pi = gpos[i];
      rbi = rhob[i];
    }
    const real H2 = (real)u.H2;
    const ar inv_h = (ar)u.inv_h;
    double sx = 0, sy = 0, sz = 0;
    ar qx = 0, qy = 0, qz = 0;  // per row run, then FP64
    ProbeRunBases pb_;   // mixed form: the forward force pass's pair vector
    auto body = [&](const uint32_t* jj, const uint32_t* tt, auto&& get) {
      G4 pj[kUnr];
      ar rbj[kUnr];
  #pragma unroll
      for (int v = 0; v < kUnr; ++v) get(jj[v] == kPad ? jj[0] : jj[v], pj[v], rbj[v]);
  #pragma unroll
      for (int v = 0; v < kUnr; ++v) {
        if (jj[v] == kPad) continue;
        real dxr, dyr, dzr;
        if constexpr (F == kMixed) pb_.pair(pi, pj[v], tt[v], u.cs_u, dxr, dyr, dzr);
        else form_pair<F>(pi, pj[v], u.cs_u, dxr, dyr, dzr);
        real r2r = dot3(dxr, dyr, dzr, dxr, dyr, dzr);
        if (!(r2r < H2 && r2r > real(0.0))) continue;
        ar rr, irr;
        adj_len((ar)r2r, rr, irr);
        const ar s = (rbi + rbj[v]) * spline_df(rr * inv_h) * irr;
        qx += s * (ar)dxr;
        qy += s * (ar)dyr;
        qz += s * (ar)dzr;
      }
    };
    walk_runs_staged<kUnr, kForceCap, kAdjPrefetch<F>>(
        rec.meta, rec.words, n, i, live, lane,
        [&](uint32_t grp) {
          if constexpr (F == kMixed) pb_.begin(pi, grp, u.cs_u);
        },
        [&](uint32_t k, uint32_t j) {
          s_p.put(k, gpos[j]);
          s_rb[k] = rhob[j];
        },
        [&](const uint32_t* jj, const uint32_t* tt) {
          body(jj, tt, [&](uint32_t k, G4& p, ar& rb) {
            p = s_p.get(k);
            rb = s_rb[k];
          });
        },
        [&](const uint32_t* jj, const uint32_t* tt) {
          body(jj, tt, [&](uint32_t j, G4& p, ar& rb) {
            p = gpos[j];
            rb = rhob[j];
          });
        },
        [&] {
          sx += (double)qx;
          sy += (double)qy;
          sz += (double)qz;
          qx = qy = qz = ar(0.0);
        });
    if (i >= n) return;
    const uint32_t dst = rec.perm[i];
    if (!live) {
      xb_out[dst] = make_double4(0, 0, 0, 0);
      vb_out[dst] = make_double4(0, 0, 0, 0);
      return;
    }
    const double k = kc.msig * u.inv_h * u.unit_scale;
    const double4 a = xb_in[i], b = xbf[i];
    xb_out[dst] = make_double4(a.x + b.x + k * sx, a.y + b.y + k * sy, a.z + b.z + k * sz, 0.0);
    vb_out[dst] = vb_in[i];
  });
}

// Sum the block partials of one reverse step into acc[0..1] (one block).
constexpr uint32_t kAdjReduceBlock = 1024;
// One block, fixed order: thread t sums partials t + 1024 (4 q + s) into its
// accumulator s (four independent load chains, so the loads overlap), then the
// four and the block's threads are combined in a fixed tree.
static __global__ __launch_bounds__(kAdjReduceBlock) void k_adj_reduce(uint32_t nparts,
                                                                       const double2* __restrict__ partial,
                                                                       double* __restrict__ acc) {
  __shared__ double sa[kAdjReduceBlock], sb[kAdjReduceBlock];
  double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
  for (uint32_t k0 = 0; k0 < nparts; k0 += 4u * kAdjReduceBlock) {
    double2 v[4];
#pragma unroll
    for (uint32_t s = 0; s < 4u; ++s) {
      const uint32_t k = k0 + s * kAdjReduceBlock + threadIdx.x;
      v[s] = k < nparts ? partial[k] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (uint32_t s = 0; s < 4u; ++s) {
      a[s] += v[s].x;
      b[s] += v[s].y;
    }
  }
  sa[threadIdx.x] = (a[0] + a[1]) + (a[2] + a[3]);
  sb[threadIdx.x] = (b[0] + b[1]) + (b[2] + b[3]);
  __syncthreads();
  for (uint32_t s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sb[threadIdx.x] += sb[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[0] += sa[0];
    acc[1] += sb[0];
  }
}

// Device buffers of the reverse sweep.
struct AdjBufs {
  double4 *xb, *vb, *xb2, *vb2;
  void* cvec;   // AdjForm<F>::C4 per slot
  double4* xbf;
  void* rhob;   // AdjForm<F>::ar per slot
  double2* partial;
  double* acc;
};

template <int F>
void sph_reverse_step(Ctx& c, const SphStepRecord& rec, AdjBufs& A) {
  const uint32_t n = (uint32_t)c.n;
  const SphUnits u = sph_units<F>(c);
  const uint32_t n_tiles = blocks_for(n, kAdjBlock);
  const TileQueue tq{c.tile_q.as<uint32_t>(), n_tiles, (n_tiles + 7u) / 8u};
  {
    LaunchScope ls(&c, "adj_prep", 2.0 * 32.0 * n + 32.0 * n);
    k_adj_prep<F><<<blocks_for(n), kBlock, 0, c.stream>>>(n, c.n_moving, rec.ids, A.xb, A.vb,
                                                          (typename AdjForm<F>::C4*)A.cvec, c.sph.dt, c.sph.msh);
    PR_HIP(hipGetLastError());
  }
  {
    const double ge = (double)sizeof(typename SphForm<F>::real);
    LaunchScope ls(&c, F == kF64 ? "adj_force_f64" : "adj_force_mixed", (8.0 * ge + 24.0 + 24.0 + 8.0 + 48.0) * n);
    static const int per_cu = blocks_per_cu(k_adj_force<F>, kAdjBlock);
    k_adj_force<F><<<queue_grid(per_cu, c.n_cus, n_tiles), kAdjBlock, 0, c.stream>>>(
        n, c.n_moving, rec, c.grid, c.sph, u, c.desc.alpha, c.desc.c0, c.desc.rho0,
        (const typename AdjForm<F>::C4*)A.cvec, A.xbf, A.vb, (typename AdjForm<F>::ar*)A.rhob, A.partial, tq);
    PR_HIP(hipGetLastError());
    k_adj_reduce<<<1, kAdjReduceBlock, 0, c.stream>>>(n_tiles, A.partial, A.acc);
    PR_HIP(hipGetLastError());
  }
  {
    const double ge = (double)sizeof(typename SphForm<F>::real);
    LaunchScope ls(&c, F == kF64 ? "adj_density_f64" : "adj_density_mixed",
                   (3.0 * ge + 8.0 + 24.0 * 3 + 24.0 * 2 + 8.0) * n);
    static const int per_cu = blocks_per_cu(k_adj_density<F>, kAdjBlock);
    k_adj_density<F><<<queue_grid(per_cu, c.n_cus, n_tiles), kAdjBlock, 0, c.stream>>>(
        n, c.n_moving, rec, c.grid, c.sph, u, (const typename AdjForm<F>::ar*)A.rhob, A.xb, A.xbf, A.vb, A.xb2,
        A.vb2, tq);
    PR_HIP(hipGetLastError());
  }
  std::swap(A.xb, A.xb2);
  std::swap(A.vb, A.vb2);
}

}  // namespace prestige

namespace prestige {

// Host helpers shared by the forms (adjoint.hip).
void snapshot_state(Ctx& c, Snapshot& s);
void restore_state(Ctx& c, const Snapshot& s);
void adj_seed_gather(Ctx& c, const uint32_t* ids, const double* seed_x, const double* seed_v);
void adj_grad_scatter(Ctx& c, const uint32_t* ids, const double4* vb, double* grad_v0_host);
void ensure_adj_buffers(Ctx& c, uint32_t grid_blocks);

// Keep what the reverse sweep needs from the step that just ran.  Taken by
// swap: the pair geometry (sorted pre-step state or FP32 probes), the
// permutation and the row-run masks, all of which the next step
// regenerates.  Copied: ids (the next reorder reads them) and (rho, p/rho^2),
// which stays downloadable.
template <int F>
void record_sph_step(Ctx& c, StepRecord& r) {
  const size_t n = c.n ? c.n : 1;
  const size_t r2 = (F == kF64) ? sizeof(double2) : sizeof(float2);
  const int prev = c.cur ^ 1;   // the step left its sorted input in the other buffer
  if (F == kMixed) {
    take_buffer(c.off32, r.gpos);
    take_buffer(c.vel32, r.gvel);
  } else {
    take_buffer(c.pos[prev], r.gpos);
    take_buffer(c.vel[prev], r.gvel);
  }
  take_buffer(c.vals[c.perm_buf], r.perm);
  take_buffer(c.nmeta, r.nmeta);
  take_buffer(c.nword, r.nword);
  r.rp.alloc(n * r2);
  r.ids.alloc(n * 4);
  auto cp = [&](DevBuf& d, const void* s, size_t b) {
    PR_HIP(hipMemcpyAsync(d.p, s, b, hipMemcpyDeviceToDevice, c.stream));
  };
  cp(r.rp, c.rp.p, c.n * r2);
  cp(r.ids, c.ids[c.cur_ids].p, c.n * 4);
}

// Steps taken while this is alive leave every particle usable row masks
// (the adjoint force pass walks those of fixed particles too).
struct RecordAll {
  Ctx& c;
  bool old;
  explicit RecordAll(Ctx& ctx) : c(ctx), old(ctx.record_all) { c.record_all = true; }
  ~RecordAll() { c.record_all = old; }
};

template <int F>
void sph_tape_forward(Ctx& c, uint32_t nsteps, uint32_t every) {
  RecordAll keep_masks(c);   // switched off below for the steps not recorded
  Tape& T = c.tape;
  T.clear();
  T.nsteps = nsteps;
  T.every = every;
  // the last segment is recorded on the way, so the reverse sweep starts
  // without recomputing it
  const uint32_t last0 = ((nsteps - 1) / every) * every;
  while (T.rec.size() < every) T.rec.emplace_back(new StepRecord());
  for (uint32_t s = 0; s < nsteps; ++s) {
    if (s % every == 0) {
      snapshot_state(c, T.next_ckpt());
    }
    c.record_all = s >= last0;   // only recorded steps need usable masks for every particle
    sph_step_impl<F>(c);
    if (s >= last0) record_sph_step<F>(c, *T.rec[s - last0]);
  }
  T.last_recorded = true;
  if (!c.tape.final_state) c.tape.final_state.reset(new Snapshot());
  snapshot_state(c, *c.tape.final_state);
  c.check_flags();
  c.tape.valid = true;
}

template <int F>
void sph_adjoint(Ctx& c, const double* seed_x, const double* seed_v, double* grad_params, double* grad_v0) {
  Tape& T = c.tape;
  const uint32_t n = (uint32_t)c.n;
  ensure_adj_buffers(c, blocks_for(n, kAdjBlock));   // one force partial per tile
  AdjBufs A{T.xb.as<double4>(), T.vb.as<double4>(), T.xb2.as<double4>(), T.vb2.as<double4>(),
            T.cvec.p, T.xbf.as<double4>(), T.rhob.p, T.partial.as<double2>(),
            T.acc.as<double>()};
  PR_HIP(hipMemsetAsync(A.acc, 0, 2 * sizeof(double), c.stream));
  adj_seed_gather(c, T.final_state->ids.as<uint32_t>(), seed_x, seed_v);
  const uint32_t nseg = (T.nsteps + T.every - 1) / T.every;
  while (T.rec.size() < T.every) T.rec.emplace_back(new StepRecord());
  for (uint32_t seg = nseg; seg-- > 0;) {
    const uint32_t s0 = seg * T.every;
    const uint32_t len = (T.nsteps - s0) < T.every ? (T.nsteps - s0) : T.every;
    if (!(seg == nseg - 1 && T.last_recorded)) {
      restore_state(c, *T.ckpt[seg]);
      RecordAll keep_masks(c);
      for (uint32_t t = 0; t < len; ++t) {
        sph_step_impl<F>(c);
        record_sph_step<F>(c, *T.rec[t]);
      }
    }
    T.last_recorded = false;   // records are overwritten by the recompute of earlier segments
    for (uint32_t t = len; t-- > 0;) {
      const StepRecord& r = *T.rec[t];
      SphStepRecord sr{r.gpos.p, r.gvel.p, r.rp.p, r.ids.as<uint32_t>(), r.perm.as<uint32_t>(),
                       r.nmeta.as<uint2>(), r.nword.as<uint32_t>()};
      sph_reverse_step<F>(c, sr, A);
    }
  }
  // A.xb / A.vb now hold the adjoints of the state the tape started from
  if (grad_params) {
    double acc[2];
    PR_HIP(hipMemcpyAsync(acc, A.acc, sizeof(acc), hipMemcpyDeviceToHost, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    grad_params[0] = acc[0];
    grad_params[1] = acc[1];
  }
  if (grad_v0) adj_grad_scatter(c, T.ckpt[0]->ids.as<uint32_t>(), A.vb, grad_v0);
  restore_state(c, *T.final_state);
  c.check_flags();
}

}  // namespace prestige
