// adjoint_dem.hpp — reverse mode of one DEM step, gather form.
//
// Forward row (i, j) of a contact, evaluated by particle i (dem_kernels.hpp):
//   d = x_i - x_j, r = |d|, n = d/r, delta = R - r,
//   ws = r_i w_i + r_j w_j, vr = (v_i - v_j) - ws x n, vn = vr.n,
//   vt = vr - vn n, cn = tz sqrt(meff kn), fn = kn delta - cn vn,
//   xp = xo - (xo.n) n, x0 = xp + dt vt, ft0 = -kt x0,
//   slip (|ft0| > mu |fn|): ft = ft0 mu|fn|/|ft0|, xn = -ft/kt;  else ft = ft0, xn = x0,
//   F_i += fn n + ft,  T_i += -r_i (n x ft),  history xn stored at i for j.
// Walls: the same row with the partner terms removed, n = +-e_a and
// delta = r_i - dist, m_eff = m_i.  Integrator: a = F/m + g, v' = v + dt a,
// x' = x + dt v', w' = w + dt T/I.
//
// Reverse: each particle k walks its own frozen contact list of the step
// (the history it wrote, in forward order; ids -> slots through the inverse
// permutation) and, for every partner j, runs the vector-Jacobian product of
// its own row (k, j) with seeds (Fbar_k, Tbar_k, xibar of its history copy)
// and of the partner's row (j, k) with the partner's seeds, keeping only the
// terms that belong to k.  Both history copies of a contact therefore get
// their own adjoint from their owner, the parameter sums take own rows only,
// and nothing is scattered: no atomics.  kt = (kt0/kn0) kn moves with kn;
// tz = 2 zeta(e) with dtz/de from the host.  Arithmetic is FP64 on the FP64
// state for both forms; the contact set and the history slots are the
// forward's own.
#pragma once
#include "dem_kernels.hpp"

namespace prestige {

struct DemAdjConst {
  double kn, kt, mu, tz, dtz_de, dt, kt_per_kn;
  double wall_lo[3], wall_hi[3];
  uint32_t wall_mask, max_contacts;
};

struct V3 {
  double x, y, z;
};
__device__ inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
__device__ inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ inline V3 operator*(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline V3 cross(V3 a, V3 b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ inline V3 xyz(const double4& a) { return v3(a.x, a.y, a.z); }

// Gradients of one contact row.
struct RowBar {
  V3 d;                  // w.r.t. the pair vector d (pairs) / position of i (walls)
  V3 vi, vj, wi, wj;     // w.r.t. velocities and angular velocities
  V3 xo;                 // w.r.t. the incoming tangential spring
  double kn, tz, mu;     // w.r.t. parameters (kt folded into kn)
};

// VJP of one row.  wall = true: n is the wall normal (no dependence on d),
// delta = ri - dist, and the partner terms vanish.
__device__ inline RowBar row_vjp(bool wall, V3 d, V3 nwall, double delta_wall, V3 vi, V3 vj, V3 wi, V3 wj,
                                 double ri, double rj, double mi, double mj, V3 xo, V3 Fb, V3 Tb, V3 xnb,
                                 const DemAdjConst& k) {
  // ---- forward recompute ----
  V3 n;
  double r = 0.0, invr = 0.0, delta;
  if (wall) {
    n = nwall;
    delta = delta_wall;
  } else {
    r = sqrt(dot(d, d));
    invr = 1.0 / r;
    n = invr * d;
    delta = (ri + rj) - r;
  }
  const V3 ws = wall ? ri * wi : ri * wi + rj * wj;
  const V3 cr = cross(ws, n);
  const V3 vr = (wall ? vi : vi - vj) - cr;
  const double vn = dot(vr, n);
  const V3 vt = vr - vn * n;
  const double meff = wall ? mi : mi * mj / (mi + mj);
  const double sq = sqrt(meff * k.kn);
  const double cn = k.tz * sq;
  const double fn = k.kn * delta - cn * vn;
  const double s = dot(xo, n);
  const V3 xp = xo - s * n;
  const V3 x0 = xp + k.dt * vt;
  const V3 ft0 = (-k.kt) * x0;
  const double ftn = sqrt(dot(ft0, ft0));
  const double sg = fn < 0.0 ? -1.0 : 1.0;
  const double fcap = k.mu * sg * fn;
  const bool slip = ftn > fcap;
  const double sc = slip ? fcap / ftn : 1.0;
  const V3 ft = sc * ft0;
  // ---- reverse ----
  RowBar b{};
  V3 nb = v3(0, 0, 0), ftb = v3(0, 0, 0);
  // T = -ri (n x ft)
  const V3 tqb = (-ri) * Tb;
  nb = nb + cross(ft, tqb);
  ftb = ftb + cross(tqb, n);
  // F = fn n + ft
  double fnb = dot(Fb, n);
  nb = nb + fn * Fb;
  ftb = ftb + Fb;
  V3 ft0b, x0b;
  double ktb = 0.0;
  if (slip) {
    // xn = ft * (-1/kt)
    ftb = ftb + (-1.0 / k.kt) * xnb;
    ktb += dot(xnb, ft) / (k.kt * k.kt);
    ft0b = sc * ftb;
    const double scb = dot(ftb, ft0);
    const double fcapb = scb / ftn;
    const double ftnb = -scb * fcap / (ftn * ftn);
    ft0b = ft0b + (ftnb / ftn) * ft0;
    b.mu += fcapb * sg * fn;
    fnb += fcapb * k.mu * sg;
    x0b = v3(0, 0, 0);
  } else {
    ft0b = ftb;
    x0b = xnb;
  }
  // ft0 = -kt x0
  x0b = x0b + (-k.kt) * ft0b;
  ktb += -dot(ft0b, x0);
  // x0 = xp + dt vt
  const V3 xpb = x0b;
  V3 vtb = k.dt * x0b;
  // xp = xo - s n, s = xo.n
  b.xo = xpb;
  const double sb = -dot(xpb, n);
  nb = nb + (-s) * xpb;
  b.xo = b.xo + sb * n;
  nb = nb + sb * xo;
  // fn = kn delta - cn vn
  b.kn += fnb * delta;
  const double deltab = fnb * k.kn;
  const double cnb = -fnb * vn;
  double vnb = -fnb * cn;
  // cn = tz sqrt(meff kn)
  b.tz += cnb * sq;
  if (k.kn > 0.0) b.kn += cnb * cn / (2.0 * k.kn);
  // vt = vr - vn n
  V3 vrb = vtb;
  vnb += -dot(vtb, n);
  nb = nb + (-vn) * vtb;
  // vn = vr . n
  vrb = vrb + vnb * n;
  nb = nb + vnb * vr;
  // vr = (vi - vj) - cr
  b.vi = vrb;
  b.vj = (-1.0) * vrb;
  const V3 crb = (-1.0) * vrb;
  // cr = ws x n
  const V3 wsb = cross(n, crb);
  nb = nb + cross(crb, ws);
  b.wi = ri * wsb;
  b.wj = rj * wsb;
  // geometry
  if (wall) {
    // delta = ri - dist with d(dist)/dx_i = n: ddelta/dx_i = -n
    b.d = (-deltab) * n;
    b.vj = v3(0, 0, 0);
    b.wj = v3(0, 0, 0);
  } else {
    // n = d / r, delta = R - r
    const V3 dn = invr * (nb - dot(nb, n) * n);
    b.d = dn + (-deltab) * n;
  }
  b.kn += ktb * k.kt_per_kn;
  return b;
}

template <class H>
__device__ inline int find_slot(const uint32_t* __restrict__ hid, uint32_t cnt, uint32_t n, uint32_t owner,
                                uint32_t partner) {
  for (uint32_t c = 0; c < cnt; ++c)
    if (hid[(size_t)c * n + owner] == partner) return (int)c;
  return -1;
}

template <class H>
__device__ inline V3 load_xi(const H* __restrict__ xi, uint32_t slot, uint32_t n, uint32_t owner) {
  return v3((double)xi[(size_t)(slot * 3 + 0) * n + owner], (double)xi[(size_t)(slot * 3 + 1) * n + owner],
            (double)xi[(size_t)(slot * 3 + 2) * n + owner]);
}

__device__ inline void store3(double* __restrict__ a, uint32_t slot, uint32_t n, uint32_t owner, V3 v) {
  a[(size_t)(slot * 3 + 0) * n + owner] = v.x;
  a[(size_t)(slot * 3 + 1) * n + owner] = v.y;
  a[(size_t)(slot * 3 + 2) * n + owner] = v.z;
}

__device__ inline V3 load3(const double* __restrict__ a, uint32_t slot, uint32_t n, uint32_t owner) {
  return v3(a[(size_t)(slot * 3 + 0) * n + owner], a[(size_t)(slot * 3 + 1) * n + owner],
            a[(size_t)(slot * 3 + 2) * n + owner]);
}

struct DemStepView {
  const double4 *pos, *vel, *om;       // FP64 sorted pre-step state (pos.w radius, vel.w mass)
  const uint32_t *ids, *perm, *slot_of;
  const uint32_t *cnt_in, *hid_in;     // history the step read
  const void* xi_in;
  const void* wxi_in;
  const uint32_t *cnt_out, *hid_out;   // contacts the step found (its frozen list)
};

static __global__ __launch_bounds__(kBlock) void k_dem_slot_of(uint32_t n, const uint32_t* __restrict__ ids,
                                                                uint32_t* __restrict__ slot_of) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) slot_of[ids[k]] = k;
}

// vbt = vbar' + dt xbar', Fbar = dt vbt / m, Tbar = dt wbar' / I.
static __global__ __launch_bounds__(kBlock) void k_dem_adj_prep(uint32_t n, uint64_t n_moving, DemStepView s,
                                                                 const double4* __restrict__ xb,
                                                                 double4* __restrict__ vb,
                                                                 const double4* __restrict__ wb,
                                                                 double4* __restrict__ Fb, double4* __restrict__ Tb,
                                                                 double dt) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if ((uint64_t)s.ids[k] >= n_moving) {
    vb[k] = make_double4(0, 0, 0, 0);
    Fb[k] = make_double4(0, 0, 0, 0);
    Tb[k] = make_double4(0, 0, 0, 0);
    return;
  }
  const double m = s.vel[k].w, r = s.pos[k].w;
  const double4 x = xb[k], v = vb[k], w = wb[k];
  const double4 vt = make_double4(v.x + dt * x.x, v.y + dt * x.y, v.z + dt * x.z, 0.0);
  vb[k] = vt;
  const double fm = dt / m;
  Fb[k] = make_double4(fm * vt.x, fm * vt.y, fm * vt.z, 0.0);
  const double ti = dt / (0.4 * m * r * r);
  Tb[k] = make_double4(ti * w.x, ti * w.y, ti * w.z, 0.0);
}

__device__ inline void block_sum3(double a, double b, double c, double4* __restrict__ partial) {
  __shared__ double s[3][kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_down(a, (unsigned)o, 64);
    b += __shfl_down(b, (unsigned)o, 64);
    c += __shfl_down(c, (unsigned)o, 64);
  }
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
    s[0][w] = a;
    s[1][w] = b;
    s[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t0 = 0, t1 = 0, t2 = 0;
    for (uint32_t q = 0; q < kBlock / 64; ++q) {
      t0 += s[0][q];
      t1 += s[1][q];
      t2 += s[2][q];
    }
    partial[blockIdx.x] = make_double4(t0, t1, t2, 0.0);
  }
}

// Contact adjoint + integrator assembly for one reverse step.
template <class H>
__global__ __launch_bounds__(kBlock) void k_dem_adj(uint32_t n, uint64_t n_moving, DemStepView s, DemAdjConst kc,
                                                    const double4* __restrict__ xb_in,
                                                    const double4* __restrict__ vb_in,
                                                    const double4* __restrict__ wb_in,
                                                    const double4* __restrict__ Fb, const double4* __restrict__ Tb,
                                                    const double* __restrict__ hb_next,
                                                    const double* __restrict__ wallb_next,
                                                    double* __restrict__ hb_prev, double* __restrict__ wallb_prev,
                                                    double4* __restrict__ xb_out, double4* __restrict__ vb_out,
                                                    double4* __restrict__ wb_out, double4* __restrict__ partial) {
  const H* xi_in = (const H*)s.xi_in;
  const H* wxi_in = (const H*)s.wxi_in;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  double pk = 0.0, pt = 0.0, pm = 0.0;
  if (k < n) {
    const uint32_t me = s.ids[k];
    const uint32_t dst = s.perm[k];
    const uint32_t cin = s.cnt_in[me];
    for (uint32_t c = 0; c < cin; ++c) store3(hb_prev, c, n, me, v3(0, 0, 0));
    if ((uint64_t)me >= n_moving) {
      for (int w = 0; w < 6; ++w) store3(wallb_prev, (uint32_t)w, n, me, v3(0, 0, 0));
      xb_out[dst] = make_double4(0, 0, 0, 0);
      vb_out[dst] = make_double4(0, 0, 0, 0);
      wb_out[dst] = make_double4(0, 0, 0, 0);
    } else {
      const double4 pk4 = s.pos[k], vk4 = s.vel[k], wk4 = s.om[k];
      const V3 xk = xyz(pk4), vk = xyz(vk4), wk = xyz(wk4);
      const double rk = pk4.w, mk = vk4.w;
      const V3 Fbk = xyz(Fb[k]), Tbk = xyz(Tb[k]);
      V3 X = v3(0, 0, 0), V = v3(0, 0, 0), W = v3(0, 0, 0);
      const uint32_t cout = s.cnt_out[me];
      for (uint32_t c = 0; c < cout; ++c) {
        const uint32_t pid = s.hid_out[(size_t)c * n + me];
        const uint32_t j = s.slot_of[pid];
        const double4 pj4 = s.pos[j], vj4 = s.vel[j], wj4 = s.om[j];
        const V3 xj = xyz(pj4), vj = xyz(vj4), wj = xyz(wj4);
        const double rj = pj4.w, mj = vj4.w;
        // own row (k, j)
        const int ci = find_slot<H>(s.hid_in, cin, n, me, pid);
        const V3 xo = ci >= 0 ? load_xi(xi_in, (uint32_t)ci, n, me) : v3(0, 0, 0);
        const V3 xnb = load3(hb_next, c, n, me);
        RowBar a = row_vjp(false, xk - xj, v3(0, 0, 0), 0.0, vk, vj, wk, wj, rk, rj, mk, mj, xo, Fbk, Tbk, xnb, kc);
        X = X + a.d;
        V = V + a.vi;
        W = W + a.wi;
        if (ci >= 0) store3(hb_prev, (uint32_t)ci, n, me, a.xo);
        pk += a.kn;
        pt += a.tz;
        pm += a.mu;
        // partner row (j, k): keep the terms of k
        const uint32_t cjin = s.cnt_in[pid], cjout = s.cnt_out[pid];
        const int cj = find_slot<H>(s.hid_in, cjin, n, pid, me);
        const V3 xoj = cj >= 0 ? load_xi(xi_in, (uint32_t)cj, n, pid) : v3(0, 0, 0);
        const int co = find_slot<H>(s.hid_out, cjout, n, pid, me);
        const V3 xnbj = co >= 0 ? load3(hb_next, (uint32_t)co, n, pid) : v3(0, 0, 0);
        const bool jm = (uint64_t)pid < n_moving;
        const V3 Fbj = jm ? xyz(Fb[j]) : v3(0, 0, 0), Tbj = jm ? xyz(Tb[j]) : v3(0, 0, 0);
        RowBar b = row_vjp(false, xj - xk, v3(0, 0, 0), 0.0, vj, vk, wj, wk, rj, rk, mj, mk, xoj, Fbj, Tbj, xnbj, kc);
        X = X - b.d;
        V = V + b.vj;
        W = W + b.wj;
      }
      // walls
      for (int w = 0; w < 6; ++w) {
        bool on = false;
        double delta = 0.0;
        V3 nw = v3(0, 0, 0);
        if (kc.wall_mask & (1u << w)) {
          const int ax = w >> 1;
          const double pa = ax == 0 ? xk.x : (ax == 1 ? xk.y : xk.z);
          const double dist = (w & 1) ? (kc.wall_hi[ax] - pa) : (pa - kc.wall_lo[ax]);
          delta = rk - dist;
          on = delta > 0.0;
          const double sgn = (w & 1) ? -1.0 : 1.0;
          nw = v3(ax == 0 ? sgn : 0.0, ax == 1 ? sgn : 0.0, ax == 2 ? sgn : 0.0);
        }
        if (!on) {
          store3(wallb_prev, (uint32_t)w, n, me, v3(0, 0, 0));
          continue;
        }
        const V3 xo = load_xi(wxi_in, (uint32_t)w, n, me);
        const V3 xnb = load3(wallb_next, (uint32_t)w, n, me);
        RowBar a = row_vjp(true, v3(0, 0, 0), nw, delta, vk, v3(0, 0, 0), wk, v3(0, 0, 0), rk, 0.0, mk, 0.0, xo,
                           Fbk, Tbk, xnb, kc);
        X = X + a.d;
        V = V + a.vi;
        W = W + a.wi;
        store3(wallb_prev, (uint32_t)w, n, me, a.xo);
        pk += a.kn;
        pt += a.tz;
        pm += a.mu;
      }
      const double4 x0 = xb_in[k], v0 = vb_in[k], w0 = wb_in[k];
      xb_out[dst] = make_double4(x0.x + X.x, x0.y + X.y, x0.z + X.z, 0.0);
      vb_out[dst] = make_double4(v0.x + V.x, v0.y + V.y, v0.z + V.z, 0.0);
      wb_out[dst] = make_double4(w0.x + W.x, w0.y + W.y, w0.z + W.z, 0.0);
    }
  }
  block_sum3(pk, pt * kc.dtz_de, pm, partial);
}

static __global__ void k_dem_adj_reduce(uint32_t nparts, const double4* __restrict__ partial,
                                        double* __restrict__ acc) {
  __shared__ double s[3][256];
  double a = 0.0, b = 0.0, c = 0.0;
  for (uint32_t q = threadIdx.x; q < nparts; q += blockDim.x) {
    a += partial[q].x;
    b += partial[q].y;
    c += partial[q].z;
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  s[2][threadIdx.x] = c;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
      s[0][threadIdx.x] += s[0][threadIdx.x + st];
      s[1][threadIdx.x] += s[1][threadIdx.x + st];
      s[2][threadIdx.x] += s[2][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    acc[0] += s[0][0];
    acc[1] += s[1][0];
    acc[2] += s[2][0];
  }
}

}  // namespace prestige

namespace prestige {

void snapshot_state(Ctx& c, Snapshot& s);
void restore_state(Ctx& c, const Snapshot& s);
void adj_seed_gather(Ctx& c, const uint32_t* ids, const double* seed_x, const double* seed_v);
void adj_grad_scatter(Ctx& c, const uint32_t* ids, const double4* vb, double* grad_v0_host);
void ensure_adj_buffers(Ctx& c, uint32_t grid_blocks);

// Keep what the reverse sweep needs from the step that just ran.  Taken by
// swap (the next step overwrites them before reading): the sorted pre-step
// state, the permutation and the contact history the step read.  Copied:
// ids (the next reorder reads them) and the history the step wrote (the next
// step's input).
template <class H>
void record_dem_step(Ctx& c, StepRecord& r) {
  const size_t n = c.n ? c.n : 1, mc = c.desc.max_contacts;
  const int prev = c.cur ^ 1, hin = c.cur_hist ^ 1, hout = c.cur_hist;
  auto cp = [&](DevBuf& d, const DevBuf& s, size_t b) {
    d.alloc(b);
    PR_HIP(hipMemcpyAsync(d.p, s.p, b, hipMemcpyDeviceToDevice, c.stream));
  };
  take_buffer(c.pos[prev], r.state_pos);
  take_buffer(c.vel[prev], r.state_vel);
  take_buffer(c.omega[prev], r.state_om);
  take_buffer(c.vals[c.perm_buf], r.perm);
  take_buffer(c.hcnt[hin], r.hcnt);
  take_buffer(c.hid[hin], r.hid);
  take_buffer(c.hxi[hin], r.hxi);
  take_buffer(c.wxi[hin], r.wxi);
  cp(r.ids, c.ids[c.cur_ids], n * 4);
  cp(r.hcnt_out, c.hcnt[hout], n * 4);
  cp(r.hid_out, c.hid[hout], n * mc * 4);
}

inline DemAdjConst dem_adj_const(const Ctx& c) {
  DemAdjConst k{};
  k.kn = c.dem.kn;
  k.kt = c.dem.kt;
  k.mu = c.dem.mu;
  k.tz = c.dem.two_zeta;
  const double L = std::log(c.desc.restitution);
  const double s = M_PI * M_PI + L * L;
  k.dtz_de = 2.0 * (-M_PI * M_PI / (s * std::sqrt(s) * c.desc.restitution));
  k.dt = c.dem.dt;
  k.kt_per_kn = c.dem.kn > 0.0 ? c.dem.kt / c.dem.kn : 0.0;
  for (int a = 0; a < 3; ++a) {
    k.wall_lo[a] = c.dem.wall_lo[a];
    k.wall_hi[a] = c.dem.wall_hi[a];
  }
  k.wall_mask = c.dem.wall_mask;
  k.max_contacts = c.dem.max_contacts;
  return k;
}

template <int F>
void dem_tape_forward(Ctx& c, uint32_t nsteps, uint32_t every) {
  using H = typename std::conditional<F == kF64, double, float>::type;
  Tape& T = c.tape;
  T.clear();
  T.nsteps = nsteps;
  T.every = every;
  // the last segment is recorded on the way, so the reverse sweep starts
  // without recomputing it
  const uint32_t last0 = nsteps ? ((nsteps - 1) / every) * every : 0;
  while (T.rec.size() < every) T.rec.emplace_back(new StepRecord());
  for (uint32_t s = 0; s < nsteps; ++s) {
    if (s % every == 0) {
      snapshot_state(c, T.next_ckpt());
    }
    dem_step_impl<F>(c);
    if (s >= last0) record_dem_step<H>(c, *T.rec[s - last0]);
  }
  T.last_recorded = nsteps > 0;
  if (!c.tape.final_state) c.tape.final_state.reset(new Snapshot());
  snapshot_state(c, *c.tape.final_state);
  c.check_flags();
  c.tape.valid = true;
}

template <int F, class H>
void dem_adjoint(Ctx& c, const double* seed_x, const double* seed_v, double* grad_params, double* grad_v0) {
  Tape& T = c.tape;
  const uint32_t n = (uint32_t)c.n;
  const size_t nn = n ? n : 1, mc = c.desc.max_contacts;
  const uint32_t nb = blocks_for(n);
  ensure_adj_buffers(c, nb);
  T.partial.alloc((size_t)nb * sizeof(double4) + 64);
  T.hb[0].alloc(nn * mc * 3 * sizeof(double));
  T.hb[1].alloc(nn * mc * 3 * sizeof(double));
  T.wallb[0].alloc(nn * 18 * sizeof(double));
  T.wallb[1].alloc(nn * 18 * sizeof(double));
  T.slot_of.alloc(nn * 4);
  PR_HIP(hipMemsetAsync(T.acc.p, 0, 3 * sizeof(double), c.stream));
  PR_HIP(hipMemsetAsync(T.wb.p, 0, nn * sizeof(double4), c.stream));
  PR_HIP(hipMemsetAsync(T.hb[0].p, 0, nn * mc * 3 * sizeof(double), c.stream));
  PR_HIP(hipMemsetAsync(T.wallb[0].p, 0, nn * 18 * sizeof(double), c.stream));
  adj_seed_gather(c, T.final_state->ids.as<uint32_t>(), seed_x, seed_v);
  double4 *xb = T.xb.as<double4>(), *vb = T.vb.as<double4>(), *wb = T.wb.as<double4>();
  double4 *xb2 = T.xb2.as<double4>(), *vb2 = T.vb2.as<double4>(), *wb2 = T.wb2.as<double4>();
  double *hn = T.hb[0].as<double>(), *hp = T.hb[1].as<double>();
  double *wn = T.wallb[0].as<double>(), *wp = T.wallb[1].as<double>();
  const DemAdjConst kc = dem_adj_const(c);
  const uint32_t nseg = (T.nsteps + T.every - 1) / T.every;
  while (T.rec.size() < T.every) T.rec.emplace_back(new StepRecord());
  for (uint32_t seg = nseg; seg-- > 0;) {
    const uint32_t s0 = seg * T.every;
    const uint32_t len = (T.nsteps - s0) < T.every ? (T.nsteps - s0) : T.every;
    if (!(seg == nseg - 1 && T.last_recorded)) {
      restore_state(c, *T.ckpt[seg]);
      for (uint32_t t = 0; t < len; ++t) {
        dem_step_impl<F>(c);
        record_dem_step<H>(c, *T.rec[t]);
      }
    }
    T.last_recorded = false;   // records are overwritten by the recompute of earlier segments
    for (uint32_t t = len; t-- > 0;) {
      StepRecord& r = *T.rec[t];
      DemStepView v{r.state_pos.as<double4>(), r.state_vel.as<double4>(), r.state_om.as<double4>(),
                    r.ids.as<uint32_t>(), r.perm.as<uint32_t>(), T.slot_of.as<uint32_t>(),
                    r.hcnt.as<uint32_t>(), r.hid.as<uint32_t>(), r.hxi.p, r.wxi.p,
                    r.hcnt_out.as<uint32_t>(), r.hid_out.as<uint32_t>()};
      if (!n) continue;
      LaunchScope ls(&c, F == kF64 ? "dem_adjoint_f64" : "dem_adjoint_mixed", 0.0);
      k_dem_slot_of<<<nb, kBlock, 0, c.stream>>>(n, v.ids, T.slot_of.as<uint32_t>());
      k_dem_adj_prep<<<nb, kBlock, 0, c.stream>>>(n, c.n_moving, v, xb, vb, wb, T.cvec.as<double4>(),
                                                  T.xbf.as<double4>(), c.dem.dt);
      k_dem_adj<H><<<nb, kBlock, 0, c.stream>>>(n, c.n_moving, v, kc, xb, vb, wb, T.cvec.as<double4>(),
                                                T.xbf.as<double4>(), hn, wn, hp, wp, xb2, vb2, wb2,
                                                T.partial.as<double4>());
      k_dem_adj_reduce<<<1, 256, 0, c.stream>>>(nb, T.partial.as<double4>(), T.acc.as<double>());
      PR_HIP(hipGetLastError());
      std::swap(xb, xb2);
      std::swap(vb, vb2);
      std::swap(wb, wb2);
      std::swap(hn, hp);
      std::swap(wn, wp);
    }
  }
  if (grad_params) {
    double acc[3];
    PR_HIP(hipMemcpyAsync(acc, T.acc.p, sizeof(acc), hipMemcpyDeviceToHost, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    for (int q = 0; q < 3; ++q) grad_params[q] = acc[q];
  }
  if (grad_v0) adj_grad_scatter(c, T.ckpt[0]->ids.as<uint32_t>(), vb, grad_v0);
  restore_state(c, *T.final_state);
  c.check_flags();
}

}  // namespace prestige
