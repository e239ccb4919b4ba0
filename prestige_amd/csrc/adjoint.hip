// adjoint.hip — reverse mode: checkpoints, seeds, gradient collection, and
// the FP64 instantiation of the WCSPH tape (mixed precision in
// adjoint_fast.hip).  Built with -ffp-contract=off like every FP64 unit.
#include "adjoint_dem.hpp"
#include "adjoint_sph.hpp"

namespace prestige {

static void copy_dev(Ctx& c, DevBuf& dst, const DevBuf& src, size_t bytes) {
  dst.alloc(bytes ? bytes : 1);
  if (bytes) PR_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, c.stream));
}

static void copy_back(Ctx& c, DevBuf& dst, const DevBuf& src, size_t bytes) {
  if (bytes) PR_HIP(hipMemcpyAsync(dst.p, src.p, bytes, hipMemcpyDeviceToDevice, c.stream));
}

static size_t elem4(const Ctx& c) { return c.desc.precision == PRESTIGE_F32 ? sizeof(float4) : sizeof(double4); }

void snapshot_state(Ctx& c, Snapshot& s) {
  const size_t n = c.n, e4 = elem4(c);
  s.cur = c.cur; s.cur_ids = c.cur_ids; s.cur_key = c.cur_key; s.cur_hist = c.cur_hist;
  s.sorted = c.sorted; s.keys_valid = c.keys_valid; s.built = c.built;
  copy_dev(c, s.pos, c.pos[c.cur], n * e4);
  copy_dev(c, s.vel, c.vel[c.cur], n * e4);
  copy_dev(c, s.ids, c.ids[c.cur_ids], n * 4);
  copy_dev(c, s.keys, c.keys[c.cur_key], n * 4);
  if (c.desc.model == PRESTIGE_MODEL_DEM) {
    const size_t mc = c.desc.max_contacts, hr = c.desc.precision == PRESTIGE_MIXED ? 4 : 8;
    copy_dev(c, s.omega, c.omega[c.cur], n * e4);
    copy_dev(c, s.hcnt, c.hcnt[c.cur_hist], n * 4);
    copy_dev(c, s.hid, c.hid[c.cur_hist], n * mc * 4);
    copy_dev(c, s.hxi, c.hxi[c.cur_hist], n * mc * 3 * hr);
    copy_dev(c, s.wxi, c.wxi[c.cur_hist], n * 18 * hr);
  }
}

void restore_state(Ctx& c, const Snapshot& s) {
  const size_t n = c.n, e4 = elem4(c);
  c.cur = s.cur; c.cur_ids = s.cur_ids; c.cur_key = s.cur_key; c.cur_hist = s.cur_hist;
  c.sorted = s.sorted; c.keys_valid = s.keys_valid; c.built = s.built;
  c.inc_ready = false;   // the previous build's keys and cell table are not restored
  c.forget_dry();
  copy_back(c, c.pos[c.cur], s.pos, n * e4);
  copy_back(c, c.vel[c.cur], s.vel, n * e4);
  copy_back(c, c.ids[c.cur_ids], s.ids, n * 4);
  copy_back(c, c.keys[c.cur_key], s.keys, n * 4);
  if (c.desc.model == PRESTIGE_MODEL_DEM) {
    const size_t mc = c.desc.max_contacts, hr = c.desc.precision == PRESTIGE_MIXED ? 4 : 8;
    copy_back(c, c.omega[c.cur], s.omega, n * e4);
    copy_back(c, c.hcnt[c.cur_hist], s.hcnt, n * 4);
    copy_back(c, c.hid[c.cur_hist], s.hid, n * mc * 4);
    copy_back(c, c.hxi[c.cur_hist], s.hxi, n * mc * 3 * hr);
    copy_back(c, c.wxi[c.cur_hist], s.wxi, n * 18 * hr);
  }
}

__global__ void k_seed_gather(uint32_t n, const uint32_t* __restrict__ ids, const double* __restrict__ sx,
                              const double* __restrict__ sv, double4* __restrict__ xb, double4* __restrict__ vb) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = ids[k];
  xb[k] = sx ? make_double4(sx[id], sx[(size_t)n + id], sx[2 * (size_t)n + id], 0.0) : make_double4(0, 0, 0, 0);
  vb[k] = sv ? make_double4(sv[id], sv[(size_t)n + id], sv[2 * (size_t)n + id], 0.0) : make_double4(0, 0, 0, 0);
}

__global__ void k_grad_scatter(uint32_t n, uint64_t n_moving, const uint32_t* __restrict__ ids,
                               const double4* __restrict__ vb, double* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const uint32_t id = ids[k];
  const bool mv = (uint64_t)id < n_moving;
  const double4 v = vb[k];
  out[id] = mv ? v.x : 0.0;
  out[(size_t)n + id] = mv ? v.y : 0.0;
  out[2 * (size_t)n + id] = mv ? v.z : 0.0;
}

void ensure_adj_buffers(Ctx& c, uint32_t grid_blocks) {
  Tape& T = c.tape;
  const size_t n = c.n ? c.n : 1;
  for (DevBuf* b : {&T.xb, &T.vb, &T.xb2, &T.vb2, &T.cvec, &T.xbf, &T.wb, &T.wb2}) b->alloc(n * sizeof(double4));
  T.rhob.alloc(n * sizeof(double));
  T.partial.alloc((size_t)grid_blocks * sizeof(double2) + 64);
  T.acc.alloc(8 * sizeof(double));
  T.seeds.alloc(6 * n * sizeof(double));
}

void adj_seed_gather(Ctx& c, const uint32_t* ids, const double* seed_x, const double* seed_v) {
  Tape& T = c.tape;
  const size_t n = c.n;
  double* dsx = nullptr;
  double* dsv = nullptr;
  if (seed_x) {
    dsx = T.seeds.as<double>();
    PR_HIP(hipMemcpyAsync(dsx, seed_x, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  if (seed_v) {
    dsv = T.seeds.as<double>() + 3 * n;
    PR_HIP(hipMemcpyAsync(dsv, seed_v, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  }
  if (n) {
    k_seed_gather<<<blocks_for(n), kBlock, 0, c.stream>>>((uint32_t)n, ids, dsx, dsv, T.xb.as<double4>(),
                                                          T.vb.as<double4>());
    PR_HIP(hipGetLastError());
  }
  PR_HIP(hipStreamSynchronize(c.stream));   // host seed buffers are borrowed for this call only
}

void adj_grad_scatter(Ctx& c, const uint32_t* ids, const double4* vb, double* grad_v0_host) {
  Tape& T = c.tape;
  const size_t n = c.n;
  if (!n) return;
  double* d = T.seeds.as<double>();
  k_grad_scatter<<<blocks_for(n), kBlock, 0, c.stream>>>((uint32_t)n, c.n_moving, ids, vb, d);
  PR_HIP(hipGetLastError());
  PR_HIP(hipMemcpyAsync(grad_v0_host, d, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  PR_HIP(hipStreamSynchronize(c.stream));
}

void sph_tape_forward_mixed(Ctx& c, uint32_t nsteps, uint32_t every);
void sph_adjoint_mixed(Ctx& c, const double* sx, const double* sv, double* gp, double* gv);
void dem_tape_forward_mixed(Ctx& c, uint32_t nsteps, uint32_t every);
void dem_adjoint_mixed(Ctx& c, const double* sx, const double* sv, double* gp, double* gv);

void tape_forward(Ctx& c, uint32_t nsteps, uint32_t every) {
  PR_CHECK(every >= 1, PRESTIGE_E_INVALID, "checkpoint_every must be >= 1");
  PR_CHECK(nsteps >= 1, PRESTIGE_E_INVALID, "nsteps must be >= 1");
  if (c.desc.model == PRESTIGE_MODEL_WCSPH) {
    PR_CHECK(!(c.desc.flags & PRESTIGE_FLAG_DENSITY_F64), PRESTIGE_E_UNSUPPORTED,
             "the adjoint does not cover PRESTIGE_FLAG_DENSITY_F64");
    if (c.desc.precision == PRESTIGE_F64) sph_tape_forward<kF64>(c, nsteps, every);
    else if (c.desc.precision == PRESTIGE_MIXED) sph_tape_forward_mixed(c, nsteps, every);
    else throw Error(PRESTIGE_E_UNSUPPORTED, "the WCSPH adjoint runs in F64 or MIXED precision");
    return;
  }
  if (c.desc.model == PRESTIGE_MODEL_DEM) {
    if (c.desc.precision == PRESTIGE_F64) dem_tape_forward<kF64>(c, nsteps, every);
    else dem_tape_forward_mixed(c, nsteps, every);
    return;
  }
  throw Error(PRESTIGE_E_UNSUPPORTED, "no adjoint for this model");
}

void adjoint_run(Ctx& c, const double* seed_x, const double* seed_v, double* grad_params, uint32_t n_params,
                 double* grad_v0) {
  PR_CHECK(c.tape.valid, PRESTIGE_E_STATE, "call prestige_forward_tape first");
  if (c.desc.model == PRESTIGE_MODEL_WCSPH) {
    PR_CHECK(grad_params == nullptr || n_params >= 2, PRESTIGE_E_INVALID, "WCSPH has 2 parameters (alpha, c0)");
    if (c.desc.precision == PRESTIGE_F64) sph_adjoint<kF64>(c, seed_x, seed_v, grad_params, grad_v0);
    else sph_adjoint_mixed(c, seed_x, seed_v, grad_params, grad_v0);
    return;
  }
  if (c.desc.model == PRESTIGE_MODEL_DEM) {
    PR_CHECK(grad_params == nullptr || n_params >= 3, PRESTIGE_E_INVALID, "DEM has 3 parameters (kn, e, mu)");
    if (c.desc.precision == PRESTIGE_F64) dem_adjoint<kF64, double>(c, seed_x, seed_v, grad_params, grad_v0);
    else dem_adjoint_mixed(c, seed_x, seed_v, grad_params, grad_v0);
    return;
  }
  throw Error(PRESTIGE_E_UNSUPPORTED, "no adjoint for this model");
}

}  // namespace prestige
