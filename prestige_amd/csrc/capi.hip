// capi.hip — the extern "C" entry points of include/prestige_hip.h: context
// lifetime, SoA field transfer by the reference's slice names, step and
// neighbour-build dispatch, timing and error reporting.  No C++ exception
// crosses this boundary.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "common.hpp"

namespace prestige {

// ---------------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------------
hipEvent_t Timing::get_event() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  PR_HIP(hipEventCreate(&e));
  return e;
}

size_t Timing::stat_index(const char* name, double bytes) {
  for (size_t k = 0; k < stats.size(); ++k)
    if (std::strcmp(stats[k].name, name) == 0) {
      stats[k].bytes_per_launch = bytes;
      return k;
    }
  KernelStat s;
  s.name = name;
  s.bytes_per_launch = bytes;
  stats.push_back(s);
  return stats.size() - 1;
}

void Timing::flush() {
  for (auto& p : pending) {
    PR_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    PR_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    stats[p.stat].total_ms += ms;
    stats[p.stat].launches += 1;
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
}

void Timing::clear() {
  for (auto& p : pending) {
    pool.push_back(p.a);
    pool.push_back(p.b);
  }
  pending.clear();
  stats.clear();
}

Timing::~Timing() {
  for (auto& p : pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : pool) (void)hipEventDestroy(e);
}

LaunchScope::LaunchScope(Ctx* c, const char* name, double bytes) : ctx(c), stat(0) {
  if (!c || !c->timing.enabled) return;
  stat = c->timing.stat_index(name, bytes);
  a = c->timing.get_event();
  b = c->timing.get_event();
  PR_HIP(hipEventRecord(a, c->stream));
}

LaunchScope::~LaunchScope() noexcept(false) {
  if (!a) return;
  hipError_t e = hipEventRecord(b, ctx->stream);
  ctx->timing.pending.push_back({a, b, stat});
  if (e != hipSuccess && !std::uncaught_exceptions())
    throw Error(PRESTIGE_E_HIP, std::string("hipEventRecord: ") + hipGetErrorString(e));
  // bound the number of outstanding event pairs
  if (ctx->timing.pending.size() > 65536 && !std::uncaught_exceptions()) ctx->timing.flush();
}

// ---------------------------------------------------------------------------
// flags
// ---------------------------------------------------------------------------
void Ctx::check_flags(const std::string& where) {
  uint32_t h[FLAG_COUNT] = {0, 0, 0, 0};
  PR_HIP(hipMemcpyAsync(h, flags.p, sizeof(h), hipMemcpyDeviceToHost, stream));
  PR_HIP(hipStreamSynchronize(stream));
  if (timing.enabled) timing.flush();
  bool any = h[0] | h[1] | h[2] | h[3];
  if (any) PR_HIP(hipMemsetAsync(flags.p, 0, sizeof(h), stream));
  const std::string at = where.empty() ? std::string() : " (" + where + ")";
  // a non-finite or runaway particle usually also leaves the grid: name the cause first
  if (h[FLAG_NONFINITE]) throw Error(PRESTIGE_E_NONFINITE, "non-finite acceleration" + at);
  if (h[FLAG_CFL])
    throw Error(PRESTIGE_E_CFL, "a particle outran the time step: WCSPH dt (c0 + |v|) > h / 2, DEM dt |v| > r / 4" + at);
  if (h[FLAG_CAPACITY])
    throw Error(PRESTIGE_E_CAPACITY,
                "a row of three cells holds more than 128 candidates, or a particle has more DEM "
                "contacts than max_contacts" + at);
  if (h[FLAG_DOMAIN]) throw Error(PRESTIGE_E_DOMAIN, "a particle left the cell grid" + at);
}

// ---------------------------------------------------------------------------
// context set-up
// ---------------------------------------------------------------------------
static void derive_constants(Ctx& c) {
  const prestige_desc& d = c.desc;
  Grid& g = c.grid;
  g = make_grid(d.grid_lo, d.grid_hi, d.cell_size);

  SphConst& s = c.sph;
  s.dt = d.dt;
  for (int a = 0; a < 3; ++a) s.g[a] = d.gravity[a];
  if (d.model == PRESTIGE_MODEL_WCSPH) {
    const double h = d.h;
    PR_CHECK(h > 0.0 && d.rho0 > 0.0 && d.mass > 0.0, PRESTIGE_E_INVALID, "WCSPH needs h, rho0, mass > 0");
    PR_CHECK(d.cell_size >= 2.0 * h, PRESTIGE_E_INVALID, "cell_size must be >= 2h");
    // same expressions as the host oracle: bit-identical constants
    s.h = h;
    s.inv_h = 1.0 / h;
    s.H2 = (2.0 * h) * (2.0 * h);
    s.eps2 = (0.01 * h) * h;
    s.inv_rho0 = 1.0 / d.rho0;
    double sigma = 1.0 / (((M_PI * h) * h) * h);
    s.msig = d.mass * sigma;
    s.msh = (d.mass * sigma) * s.inv_h;
    s.alpha_c0 = d.alpha * d.c0;
    s.B = ((d.rho0 * d.c0) * d.c0) / 7.0;
    const double vmax = 0.5 * h / d.dt - d.c0;
    s.cfl_v2 = vmax > 0.0 ? vmax * vmax : -1.0;
    g.off_scale = s.inv_h;
  }
  DemConst& k = c.dem;
  k.dt = d.dt;
  for (int a = 0; a < 3; ++a) {
    k.g[a] = d.gravity[a];
    k.wall_lo[a] = d.wall_lo[a];
    k.wall_hi[a] = d.wall_hi[a];
  }
  k.wall_mask = d.wall_mask;
  k.max_contacts = d.max_contacts;
  if (d.model == PRESTIGE_MODEL_DEM) {
    PR_CHECK(d.kn > 0.0 && d.kt >= 0.0 && d.mu >= 0.0, PRESTIGE_E_INVALID, "DEM needs kn > 0, kt >= 0, mu >= 0");
    PR_CHECK(d.restitution > 0.0 && d.restitution <= 1.0, PRESTIGE_E_INVALID, "restitution must be in (0, 1]");
    PR_CHECK(d.max_contacts >= 1 && d.max_contacts <= 32, PRESTIGE_E_INVALID, "max_contacts must be in [1, 32]");
    k.kn = d.kn;
    k.kt = d.kt;
    k.mu = d.mu;
    double le = std::log(d.restitution);
    k.two_zeta = 2.0 * (-le / std::sqrt(M_PI * M_PI + le * le));
    k.neg_inv_kt = d.kt > 0.0 ? -1.0 / d.kt : 0.0;
  }
}

__global__ void k_iota_ids(uint32_t n, uint32_t* ids) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ids[i] = i;
}

static void allocate(Ctx& c) {
  const uint64_t n = c.n;
  const bool f32 = c.desc.precision == PRESTIGE_F32;
  const bool mixed = c.desc.precision == PRESTIGE_MIXED;
  const bool dem = c.desc.model == PRESTIGE_MODEL_DEM;
  const size_t e4 = f32 ? sizeof(float4) : sizeof(double4);
  const size_t nn = n ? n : 1;
  for (int b = 0; b < 2; ++b) {
    c.pos[b].alloc(nn * e4);
    c.vel[b].alloc(nn * e4);
    PR_HIP(hipMemsetAsync(c.pos[b].p, 0, nn * e4, c.stream));
    PR_HIP(hipMemsetAsync(c.vel[b].p, 0, nn * e4, c.stream));
    if (dem) {
      c.omega[b].alloc(nn * e4);
      PR_HIP(hipMemsetAsync(c.omega[b].p, 0, nn * e4, c.stream));
    }
    c.ids[b].alloc(nn * 4);
    c.keys[b].alloc(nn * 4);
    c.vals[b].alloc(nn * 4);
  }
  const uint64_t tiles = sort_tiles(n) + 1;
  c.sort_hist.alloc(256 * tiles * 4);
  c.sort_tot.alloc(256 * 4);
  c.inc_mask.alloc(((nn + 63) / 64) * 8);
  c.inc_woff.alloc(((nn + 63) / 64) * 4);
  c.inc_tiles.alloc((2 * tiles + 1) * 4);
  PR_HIP(hipMemsetAsync(c.inc_tiles.p, 0, (2 * tiles + 1) * 4, c.stream));
  c.inc_count.alloc(4);
  c.inc_mov.alloc(nn * 8);
  c.inc_tmp.alloc(nn * 8);
  PR_HIP(hipHostMalloc((void**)&c.inc_host, sizeof(uint32_t), hipHostMallocMapped));
  *c.inc_host = 0;
  PR_HIP(hipHostGetDevicePointer((void**)&c.inc_host_dev, c.inc_host, 0));
  if (const char* dr = std::getenv("PRESTIGE_DRY")) c.dry_reuse = std::string(dr) != "off";
  if (const char* sp = std::getenv("PRESTIGE_SORT")) {
    if (std::string(sp) == "radix") c.sort_policy = Ctx::kSortRadix;
    else if (std::string(sp) == "incremental") c.sort_policy = Ctx::kSortIncremental;
  }
  c.lb.alloc(((size_t)c.grid.n_cells + 1) * 4);
  c.flags.alloc(FLAG_COUNT * 4);
  c.tile_q.alloc(kQueueWords * 4);
  if (c.desc.model == PRESTIGE_MODEL_WCSPH) {
    c.movcell.alloc(((size_t)c.grid.n_cells + 1) * 4);
    PR_HIP(hipMemsetAsync(c.movcell.p, 0, ((size_t)c.grid.n_cells + 1) * 4, c.stream));
    c.dry_pass.alloc(nn * 4);
    PR_HIP(hipMemsetAsync(c.dry_pass.p, 0, nn * 4, c.stream));
  }
  PR_HIP(hipMemsetAsync(c.tile_q.p, 0, kQueueWords * 4, c.stream));
  PR_HIP(hipDeviceGetAttribute(&c.n_cus, hipDeviceAttributeMultiprocessorCount, c.device));
  PR_HIP(hipMemsetAsync(c.flags.p, 0, FLAG_COUNT * 4, c.stream));
  const size_t es = f32 || (mixed && c.desc.model == PRESTIGE_MODEL_WCSPH) ? 4 : 8;
  c.rp.alloc(nn * 2 * es);
  c.pres.alloc(nn * es);
  c.acc.alloc(nn * e4);
  PR_HIP(hipMemsetAsync(c.rp.p, 0, nn * 2 * es, c.stream));
  PR_HIP(hipMemsetAsync(c.pres.p, 0, nn * es, c.stream));
  if (c.desc.flags & PRESTIGE_FLAG_DENSITY_F64) {
    c.rp64.alloc(nn * 16);
    c.pres64.alloc(nn * 8);
    PR_HIP(hipMemsetAsync(c.rp64.p, 0, nn * 16, c.stream));
    PR_HIP(hipMemsetAsync(c.pres64.p, 0, nn * 8, c.stream));
  }
  if (c.desc.flags & PRESTIGE_FLAG_DEM_EVENTS) {
    c.events.alloc(nn * 8);
    PR_HIP(hipMemsetAsync(c.events.p, 0, nn * 8, c.stream));
  }
  if (c.desc.model == PRESTIGE_MODEL_WCSPH && c.desc.precision != PRESTIGE_F64) {
    c.rp_id.alloc(nn * 2 * es);   // outputs of dry wall particles by id
    c.pres_id.alloc(nn * es);
  }
  PR_HIP(hipMemsetAsync(c.acc.p, 0, nn * e4, c.stream));
  if (mixed) c.off32.alloc(nn * sizeof(float4));
  if (mixed) {
    c.vel32.alloc(nn * sizeof(float4));
    if (dem) c.om32.alloc(nn * sizeof(float4));
  }
  if (c.desc.model == PRESTIGE_MODEL_WCSPH) {
    c.nmeta.alloc(nn * 9 * sizeof(uint2));
    c.nword.alloc(nn * 36 * sizeof(uint32_t));
  }
  c.scratch.alloc(nn * 8);
  if (dem) {
    const size_t mc = c.desc.max_contacts;
    const size_t hr = mixed ? 4 : 8;
    for (int b = 0; b < 2; ++b) {
      c.hcnt[b].alloc(nn * 4);
      c.hid[b].alloc(nn * mc * 4);
      c.hxi[b].alloc(nn * mc * 3 * hr);
      c.wxi[b].alloc(nn * 18 * hr);
      PR_HIP(hipMemsetAsync(c.hcnt[b].p, 0, nn * 4, c.stream));
      PR_HIP(hipMemsetAsync(c.wxi[b].p, 0, nn * 18 * hr, c.stream));
    }
  }
  if (n) {
    k_iota_ids<<<blocks_for(n), kBlock, 0, c.stream>>>((uint32_t)n, c.ids[0].as<uint32_t>());
    PR_HIP(hipGetLastError());
  }
  c.cur_ids = 0;
  c.cur = 0;
  PR_HIP(hipStreamSynchronize(c.stream));
}

// Field table: where a named SoA slice lives.
struct FieldRef {
  void* buf;
  int kind;    // 0 double, 1 float, 2 uint32 keys
  int comp;    // -1 scalar, 0..3 component of a 4-vector, 100+k component of a 2-vector
  bool id_order;
  bool writable;
};

static FieldRef find_field(Ctx& c, const std::string& name) {
  const bool f32 = c.desc.precision == PRESTIGE_F32;
  const int sk = f32 ? 1 : 0;
  const bool dem = c.desc.model == PRESTIGE_MODEL_DEM;
  const bool sph_small = f32 || (c.desc.precision == PRESTIGE_MIXED && c.desc.model == PRESTIGE_MODEL_WCSPH);
  const int rk = sph_small ? 1 : 0;
  if (name == "x") return {c.pos[c.cur].p, sk, 0, false, true};
  if (name == "y") return {c.pos[c.cur].p, sk, 1, false, true};
  if (name == "z") return {c.pos[c.cur].p, sk, 2, false, true};
  if (name == "vx") return {c.vel[c.cur].p, sk, 0, false, true};
  if (name == "vy") return {c.vel[c.cur].p, sk, 1, false, true};
  if (name == "vz") return {c.vel[c.cur].p, sk, 2, false, true};
  if (dem && name == "wx") return {c.omega[c.cur].p, sk, 0, false, true};
  if (dem && name == "wy") return {c.omega[c.cur].p, sk, 1, false, true};
  if (dem && name == "wz") return {c.omega[c.cur].p, sk, 2, false, true};
  if (dem && name == "radius") return {c.pos[c.cur].p, sk, 3, false, true};
  if (c.desc.flags & PRESTIGE_FLAG_DENSITY_F64) {
    if (name == "rho") return {c.rp64.p, 0, 100, false, false};
    if (name == "p") return {c.pres64.p, 0, -1, false, false};
  }
  if (name == "rho") return {c.rp.p, rk, 100, false, false};
  if (name == "p") return {c.pres.p, rk, -1, false, false};
  if (name == "ax") return {c.acc.p, sk, 0, false, false};
  if (name == "ay") return {c.acc.p, sk, 1, false, false};
  if (name == "az") return {c.acc.p, sk, 2, false, false};
  if (name == "key") return {c.keys[c.cur_key].p, 2, -1, false, false};
  if (name == "force") return {c.force.p, 0, -1, true, true};
  if (name == "mass") return {c.mass.p, 0, -1, true, true};
  throw Error(PRESTIGE_E_INVALID, "unknown field '" + name + "'");
}

// Names of fields that generated equations own: any identifier that is not a
// field of the context's model (state, eq1's force / mass, the key; WCSPH
// outputs in WCSPH contexts).
bool generic_field_name(const Ctx& c, const std::string& name) {
  if (name.empty() || !(std::isalpha((unsigned char)name[0]) || name[0] == '_')) return false;
  for (char ch : name)
    if (!(std::isalnum((unsigned char)ch) || ch == '_')) return false;
  static const char* model[] = {"x", "y", "z", "vx", "vy", "vz", "wx", "wy", "wz", "radius", "force", "mass", "key"};
  for (const char* m : model)
    if (name == m) return false;
  if (c.desc.model != PRESTIGE_MODEL_PAIRS)
    for (const char* m : {"rho", "p", "ax", "ay", "az"})
      if (name == m) return false;
  return true;
}

// FP64 copy, in particle-id order, of the model field `name` (state or step
// output) into the device array d_out; false when the context has no such field.
bool model_field_ids(Ctx& c, const std::string& name, double* d_out) {
  if (name == "key" || name == "force" || name == "mass") return false;
  FieldRef f;
  try {
    f = find_field(c, name);
  } catch (const Error&) {
    return false;
  }
  if (c.n) scatter_to_ids(c, f.buf, f.kind, f.comp, d_out);
  return true;
}

static void upload(Ctx& c, const std::string& name, const double* host, uint64_t n) {
  PR_CHECK(n == c.n, PRESTIGE_E_INVALID, "field length must equal n_particles");
  PR_CHECK(host != nullptr || n == 0, PRESTIGE_E_INVALID, "null data");
  if (generic_field_name(c, name)) {
    std::unique_ptr<DevBuf>& b = c.gfields[name];
    if (!b) b.reset(new DevBuf());
    b->alloc((n ? n : 1) * 8);
    if (n) PR_HIP(hipMemcpyAsync(b->p, host, n * 8, hipMemcpyHostToDevice, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  if (name == "force" || name == "mass") {
    DevBuf& b = name == "force" ? c.force : c.mass;
    b.alloc((n ? n : 1) * 8);
    if (n) PR_HIP(hipMemcpyAsync(b.p, host, n * 8, hipMemcpyHostToDevice, c.stream));
    if (name == "force") c.have_force = true; else c.have_mass = true;
    if (name == "mass" && c.desc.model == PRESTIGE_MODEL_DEM) {
      for (uint64_t i = 0; i < n; ++i)
        PR_CHECK(host[i] > 0.0 && std::isfinite(host[i]), PRESTIGE_E_INVALID, "DEM masses must be finite and > 0");
      gather_from_ids(c, c.mass.as<double>(), c.vel[c.cur].p, c.desc.precision == PRESTIGE_F32 ? 1 : 0, 3);
    }
    PR_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  FieldRef f = find_field(c, name);
  PR_CHECK(f.writable, PRESTIGE_E_INVALID, "field '" + name + "' is an output");
  for (uint64_t i = 0; i < n; ++i)
    PR_CHECK(std::isfinite(host[i]), PRESTIGE_E_NONFINITE, "non-finite value in field '" + name + "'");
  if (name == "radius") {
    double rmax = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      PR_CHECK(host[i] > 0.0, PRESTIGE_E_INVALID, "radii must be > 0");
      rmax = host[i] > rmax ? host[i] : rmax;
    }
    PR_CHECK(2.0 * rmax <= c.grid.cs, PRESTIGE_E_INVALID, "cell_size must be >= 2 * max radius");
  }
  if (n) {
    PR_HIP(hipMemcpyAsync(c.scratch.p, host, n * 8, hipMemcpyHostToDevice, c.stream));
    gather_from_ids(c, c.scratch.as<double>(), f.buf, f.kind, f.comp);
    PR_HIP(hipStreamSynchronize(c.stream));
  }
  if (name == "x" || name == "y" || name == "z") {
    c.keys_valid = false;
    c.sorted = false;
    c.inc_ready = false;
    if (c.events.p) {
      PR_HIP(hipMemsetAsync(c.events.p, 0, c.n * 8, c.stream));
      c.ev_steps = 0;
    }
    if (c.desc.model == PRESTIGE_MODEL_DEM) {
      // new positions invalidate the tangential contact history
      const size_t hr = c.desc.precision == PRESTIGE_MIXED ? 4 : 8;
      PR_HIP(hipMemsetAsync(c.hcnt[c.cur_hist].p, 0, c.n * 4, c.stream));
      PR_HIP(hipMemsetAsync(c.wxi[c.cur_hist].p, 0, c.n * 18 * hr, c.stream));
      PR_HIP(hipStreamSynchronize(c.stream));
    }
  }
  c.tape.clear();
  c.forget_dry();
}

static void download(Ctx& c, const std::string& name, double* host, uint64_t n) {
  PR_CHECK(n == c.n, PRESTIGE_E_INVALID, "field length must equal n_particles");
  PR_CHECK(host != nullptr || n == 0, PRESTIGE_E_INVALID, "null output");
  if (c.events.p && (name == "contact_events" || name == "slip_events")) {
    std::vector<uint32_t> u(n ? n : 1);
    const size_t off = name == "contact_events" ? 0 : c.n;
    if (n) PR_HIP(hipMemcpyAsync(u.data(), c.events.as<uint32_t>() + off, n * 4, hipMemcpyDeviceToHost, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    for (uint64_t i = 0; i < n; ++i) host[i] = (double)u[i];
    return;
  }
  if (generic_field_name(c, name)) {
    auto it = c.gfields.find(name);
    PR_CHECK(it != c.gfields.end(), PRESTIGE_E_STATE, "field '" + name + "' was never uploaded or written");
    if (n) PR_HIP(hipMemcpyAsync(host, it->second->p, n * 8, hipMemcpyDeviceToHost, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  FieldRef f = find_field(c, name);
  if (f.id_order) {
    PR_CHECK(name == "force" ? c.have_force : c.have_mass, PRESTIGE_E_STATE, "field '" + name + "' was never uploaded");
    if (n) PR_HIP(hipMemcpyAsync(host, f.buf, n * 8, hipMemcpyDeviceToHost, c.stream));
    PR_HIP(hipStreamSynchronize(c.stream));
    return;
  }
  if (name == "key" && !c.keys_valid) keys_from_positions(c);
  if (name == "key") f.buf = c.keys[c.cur_key].p;
  if (n) {
    scatter_to_ids(c, f.buf, f.kind, f.comp, c.scratch.as<double>());
    PR_HIP(hipMemcpyAsync(host, c.scratch.p, n * 8, hipMemcpyDeviceToHost, c.stream));
  }
  PR_HIP(hipStreamSynchronize(c.stream));
}

// Device flags are read every kFlagEvery steps (one stream synchronisation),
// so a step that raises one ends the call within that many steps, and the
// error names the stretch in which it was raised.
constexpr uint32_t kFlagEvery = 16;

static void step(Ctx& c, uint32_t nsteps) {
  const int m = c.desc.model, p = c.desc.precision;
  if (m == PRESTIGE_MODEL_WCSPH && nsteps)
    PR_CHECK(c.sph.cfl_v2 > 0.0, PRESTIGE_E_CFL, "dt c0 > h / 2: the time step breaks the CFL bound at rest");
  uint32_t first = 0;
  for (uint32_t s = 0; s < nsteps; ++s) {
    if (m == PRESTIGE_MODEL_WCSPH) {
      if (p == PRESTIGE_F64) sph_step_f64(c);
      else if (p == PRESTIGE_F32) sph_step_f32(c);
      else sph_step_mixed(c);
    } else if (m == PRESTIGE_MODEL_DEM) {
      if (c.events.p) c.ev_weight = ++c.ev_steps;
      if (p == PRESTIGE_F64) dem_step_f64(c);
      else if (p == PRESTIGE_MIXED) dem_step_mixed(c);
      else throw Error(PRESTIGE_E_UNSUPPORTED, "DEM runs in F64 or MIXED precision");
      c.ev_weight = 0;
    } else {
      throw Error(PRESTIGE_E_UNSUPPORTED, "model PAIRS has no step; use prestige_eval_equation");
    }
    if ((s + 1) % kFlagEvery == 0 || s + 1 == nsteps) {
      c.check_flags("raised in steps " + std::to_string(first + 1) + "-" + std::to_string(s + 1) + " of " +
                    std::to_string(nsteps) + " of this call");
      first = s + 1;
    }
  }
  if (!nsteps) c.check_flags();
}

// Tape entry points live in adjoint code when built; declared weak-free here.
void tape_forward(Ctx& c, uint32_t nsteps, uint32_t every);
void adjoint_run(Ctx& c, const double* seed_x, const double* seed_v, double* grad_params,
                 uint32_t n_params, double* grad_v0);

}  // namespace prestige

using namespace prestige;

struct prestige_ctx : Ctx {};

namespace {
thread_local std::string g_create_error;
}  // namespace

namespace prestige {
void set_context_free_error(const std::string& msg) { g_create_error = msg; }
}  // namespace prestige

namespace {

template <class Fn>
prestige_status guard(prestige_ctx* ctx, Fn&& fn) {
  if (!ctx) {
    g_create_error = "null context";
    return PRESTIGE_E_INVALID;
  }
  try {
    ctx->err.clear();
    fn();
    return PRESTIGE_OK;
  } catch (const Error& e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    ctx->err = "host allocation failed";
    return PRESTIGE_E_OOM;
  } catch (const std::exception& e) {
    ctx->err = e.what();
    return PRESTIGE_E_INVALID;
  } catch (...) {
    ctx->err = "unknown C++ exception";
    return PRESTIGE_E_INVALID;
  }
}
}  // namespace

extern "C" {

void prestige_desc_init(prestige_desc* d, int32_t model) {
  if (!d) return;
  std::memset(d, 0, sizeof(*d));
  d->abi_version = PRESTIGE_ABI_VERSION;
  d->model = model;
  d->precision = PRESTIGE_F64;
  d->max_contacts = 16;
  d->gravity[2] = -9.81;
  d->rho0 = 1000.0;
  d->alpha = 0.1;
  d->kn = 1.0e5;
  d->kt = 2.0 / 7.0 * 1.0e5;
  d->restitution = 0.3;
  d->mu = 0.5;
}

prestige_status prestige_create(const prestige_desc* desc, prestige_ctx** out) {
  if (!desc || !out) {
    g_create_error = "null descriptor or output pointer";
    return PRESTIGE_E_INVALID;
  }
  *out = nullptr;
  prestige_ctx* c = nullptr;
  try {
    if (desc->abi_version != PRESTIGE_ABI_VERSION) throw Error(PRESTIGE_E_INVALID, "abi_version mismatch");
    if (desc->model < 0 || desc->model > PRESTIGE_MODEL_DEM) throw Error(PRESTIGE_E_INVALID, "unknown model");
    if (desc->precision < 0 || desc->precision > PRESTIGE_MIXED)
      throw Error(PRESTIGE_E_INVALID, "unknown precision");
    if (desc->n_particles >= 0xFFFFFFFFull) throw Error(PRESTIGE_E_INVALID, "n_particles must be < 2^32 - 1");
    // the mixed WCSPH density walk tags slots with their run in bit 31
    if (desc->model == PRESTIGE_MODEL_WCSPH && desc->precision == PRESTIGE_MIXED && desc->n_particles >= 0x80000000ull)
      throw Error(PRESTIGE_E_INVALID, "n_particles must be < 2^31 in the MIXED WCSPH form");
    if (desc->flags & ~(PRESTIGE_FLAG_DENSITY_F64 | PRESTIGE_FLAG_DEM_EVENTS))
      throw Error(PRESTIGE_E_INVALID, "unknown bits in flags");
    if ((desc->flags & PRESTIGE_FLAG_DEM_EVENTS) && desc->model != PRESTIGE_MODEL_DEM)
      throw Error(PRESTIGE_E_INVALID, "PRESTIGE_FLAG_DEM_EVENTS applies to DEM contexts only");
    if ((desc->flags & PRESTIGE_FLAG_DENSITY_F64) &&
        !(desc->model == PRESTIGE_MODEL_WCSPH && desc->precision == PRESTIGE_MIXED))
      throw Error(PRESTIGE_E_INVALID, "PRESTIGE_FLAG_DENSITY_F64 applies to the MIXED WCSPH form only");
    if (desc->n_moving > desc->n_particles) throw Error(PRESTIGE_E_INVALID, "n_moving > n_particles");
    if (!(desc->cell_size > 0.0) || !std::isfinite(desc->cell_size))
      throw Error(PRESTIGE_E_INVALID, "cell_size must be finite and > 0");
    if (desc->model != PRESTIGE_MODEL_PAIRS && !(desc->dt > 0.0)) throw Error(PRESTIGE_E_INVALID, "dt must be > 0");
    c = new prestige_ctx();
    c->desc = *desc;
    c->n = desc->n_particles;
    c->n_moving = desc->n_moving;
    derive_constants(*c);
    PR_HIP(hipGetDevice(&c->device));
    PR_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    allocate(*c);
  } catch (const Error& e) {
    g_create_error = e.what();
    if (c) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
    }
    return e.code;
  } catch (const std::exception& e) {
    g_create_error = e.what();
    if (c) {
      if (c->stream) (void)hipStreamDestroy(c->stream);
      delete c;
    }
    return PRESTIGE_E_INVALID;
  }
  g_create_error.clear();
  *out = c;
  return PRESTIGE_OK;
}

void prestige_destroy(prestige_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  hipStream_t s = ctx->stream;
  delete ctx;   // buffers and events released by their destructors
  if (s) (void)hipStreamDestroy(s);
}

prestige_status prestige_field_upload(prestige_ctx* ctx, const char* name, const double* data, uint64_t n) {
  return guard(ctx, [&] {
    PR_CHECK(name != nullptr, PRESTIGE_E_INVALID, "null field name");
    upload(*ctx, name, data, n);
  });
}

prestige_status prestige_field_download(prestige_ctx* ctx, const char* name, double* out, uint64_t n) {
  return guard(ctx, [&] {
    PR_CHECK(name != nullptr, PRESTIGE_E_INVALID, "null field name");
    download(*ctx, name, out, n);
  });
}

prestige_status prestige_write_particles(prestige_ctx* ctx, const char* path, int32_t format,
                                         const char* fields, double time) {
  return guard(ctx, [&] {
    PR_CHECK(path != nullptr, PRESTIGE_E_INVALID, "null path");
    PR_CHECK(format == PRESTIGE_OUT_VTK || format == PRESTIGE_OUT_CSV, PRESTIGE_E_INVALID, "unknown format");
    std::vector<std::string> names;
    if (fields) {
      std::string cur;
      for (const char* q = fields;; ++q) {
        if (*q == ',' || *q == '\0') {
          if (!cur.empty()) names.push_back(cur);
          cur.clear();
          if (*q == '\0') break;
        } else if (*q != ' ') {
          cur += *q;
        }
      }
    }
    const uint64_t n = ctx->n;
    std::vector<double> xyz[3];
    for (int a = 0; a < 3; ++a) {
      xyz[a].resize(n);
      download(*ctx, a == 0 ? "x" : (a == 1 ? "y" : "z"), xyz[a].data(), n);
    }
    std::vector<std::vector<double>> data(names.size());
    for (size_t k = 0; k < names.size(); ++k) {
      data[k].resize(n);
      download(*ctx, names[k], data[k].data(), n);
    }
    FILE* f = std::fopen(path, format == PRESTIGE_OUT_VTK ? "wb" : "w");
    PR_CHECK(f != nullptr, PRESTIGE_E_INVALID, std::string("cannot open '") + path + "' for writing");
    bool ok = true;
    if (format == PRESTIGE_OUT_CSV) {
      ok &= std::fprintf(f, "# time %.17g\nid,x,y,z", time) > 0;
      for (const auto& nm : names) ok &= std::fprintf(f, ",%s", nm.c_str()) > 0;
      ok &= std::fputc('\n', f) != EOF;
      for (uint64_t i = 0; i < n && ok; ++i) {
        ok &= std::fprintf(f, "%llu,%.17g,%.17g,%.17g", (unsigned long long)i, xyz[0][i], xyz[1][i], xyz[2][i]) > 0;
        for (const auto& col : data) ok &= std::fprintf(f, ",%.17g", col[i]) > 0;
        ok &= std::fputc('\n', f) != EOF;
      }
    } else {
      // legacy VTK binary data are big-endian
      auto be64 = [](double v) {
        uint64_t u;
        std::memcpy(&u, &v, 8);
        return __builtin_bswap64(u);
      };
      auto put = [&](const void* p, size_t b) { ok &= std::fwrite(p, 1, b, f) == b; };
      std::vector<uint64_t> buf;
      ok &= std::fprintf(f, "# vtk DataFile Version 3.0\nprestige particles, time %.17g\nBINARY\n"
                            "DATASET POLYDATA\nFIELD FieldData 1\nTIME 1 1 double\n", time) > 0;
      const uint64_t tb = be64(time);
      put(&tb, 8);
      ok &= std::fprintf(f, "\nPOINTS %llu double\n", (unsigned long long)n) > 0;
      buf.resize(3 * n);
      for (uint64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) buf[3 * i + a] = be64(xyz[a][i]);
      put(buf.data(), buf.size() * 8);
      ok &= std::fprintf(f, "\nVERTICES %llu %llu\n", (unsigned long long)n, (unsigned long long)(2 * n)) > 0;
      std::vector<uint32_t> cells(2 * n);
      for (uint64_t i = 0; i < n; ++i) {
        cells[2 * i] = __builtin_bswap32(1u);
        cells[2 * i + 1] = __builtin_bswap32((uint32_t)i);
      }
      put(cells.data(), cells.size() * 4);
      ok &= std::fprintf(f, "\nPOINT_DATA %llu\n", (unsigned long long)n) > 0;
      buf.resize(n);
      for (size_t k = 0; k < names.size(); ++k) {
        ok &= std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", names[k].c_str()) > 0;
        for (uint64_t i = 0; i < n; ++i) buf[i] = be64(data[k][i]);
        put(buf.data(), n * 8);
        ok &= std::fputc('\n', f) != EOF;
      }
    }
    ok &= std::fclose(f) == 0;
    PR_CHECK(ok, PRESTIGE_E_INVALID, std::string("write to '") + path + "' failed");
  });
}

prestige_status prestige_write_series(const char* path, const char* const* files, const double* times,
                                      uint32_t n) {
  if (!path || (n && (!files || !times))) {
    g_create_error = "null path or list";
    return PRESTIGE_E_INVALID;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    g_create_error = std::string("cannot open '") + path + "' for writing";
    return PRESTIGE_E_INVALID;
  }
  bool ok = std::fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"Collection\" version=\"0.1\">\n"
                            "  <Collection>\n") > 0;
  for (uint32_t k = 0; k < n; ++k) {
    if (!files[k]) {
      ok = false;
      break;
    }
    ok &= std::fprintf(f, "    <DataSet timestep=\"%.17g\" part=\"0\" file=\"%s\"/>\n", times[k], files[k]) > 0;
  }
  ok &= std::fprintf(f, "  </Collection>\n</VTKFile>\n") > 0;
  ok &= std::fclose(f) == 0;
  if (!ok) {
    g_create_error = std::string("write to '") + path + "' failed";
    return PRESTIGE_E_INVALID;
  }
  return PRESTIGE_OK;
}

prestige_status prestige_neighbors_build(prestige_ctx* ctx) {
  return guard(ctx, [&] {
    neighbors_build(*ctx);
    ctx->check_flags();
  });
}

prestige_status prestige_neighbors_download(prestige_ctx* ctx, uint32_t* keys, uint32_t* perm, uint32_t* ids,
                                            uint32_t* cell_lb) {
  return guard(ctx, [&] {
    PR_CHECK(ctx->built, PRESTIGE_E_STATE, "no neighbour build yet");
    PR_CHECK(ctx->sorted, PRESTIGE_E_STATE, "state moved since the last build; call prestige_neighbors_build");
    const size_t n = ctx->n;
    if (n && keys) PR_HIP(hipMemcpyAsync(keys, ctx->keys[ctx->cur_key].p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (n && perm) PR_HIP(hipMemcpyAsync(perm, ctx->vals[ctx->perm_buf].p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (n && ids) PR_HIP(hipMemcpyAsync(ids, ctx->ids[ctx->cur_ids].p, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (cell_lb)
      PR_HIP(hipMemcpyAsync(cell_lb, ctx->lb.p, ((size_t)ctx->grid.n_cells + 1) * 4, hipMemcpyDeviceToHost,
                            ctx->stream));
    PR_HIP(hipStreamSynchronize(ctx->stream));
  });
}

prestige_status prestige_grid_dims(prestige_ctx* ctx, uint32_t dims[3]) {
  return guard(ctx, [&] {
    PR_CHECK(dims != nullptr, PRESTIGE_E_INVALID, "null dims");
    for (int a = 0; a < 3; ++a) dims[a] = (uint32_t)ctx->grid.dims[a];
  });
}

prestige_status prestige_eval_equation(prestige_ctx* ctx, int32_t eq_id, int32_t loop, double cutoff) {
  return guard(ctx, [&] {
    PR_CHECK(eq_id == PRESTIGE_EQ1, PRESTIGE_E_UNSUPPORTED,
             "only eq1 is evaluated standalone; the WCSPH/DEM equations run inside prestige_step");
    eq1_eval(*ctx, loop, cutoff);
    ctx->check_flags();
  });
}

prestige_status prestige_run_equations(prestige_ctx* ctx, const int32_t* eq_ids, uint32_t n, int32_t loop,
                                      double cutoff) {
  return guard(ctx, [&] { run_equations(*ctx, eq_ids, n, loop, cutoff); });
}

prestige_status prestige_step(prestige_ctx* ctx, uint32_t nsteps) {
  return guard(ctx, [&] { step(*ctx, nsteps); });
}

prestige_status prestige_forward_tape(prestige_ctx* ctx, uint32_t nsteps, uint32_t checkpoint_every) {
  return guard(ctx, [&] { tape_forward(*ctx, nsteps, checkpoint_every); });
}

prestige_status prestige_adjoint(prestige_ctx* ctx, const double* seed_x, const double* seed_v,
                                 double* grad_params, uint32_t n_params, double* grad_v0) {
  return guard(ctx, [&] { adjoint_run(*ctx, seed_x, seed_v, grad_params, n_params, grad_v0); });
}

prestige_status prestige_timing_enable(prestige_ctx* ctx, int32_t enable) {
  return guard(ctx, [&] {
    PR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.flush();
    ctx->timing.clear();
    ctx->timing.enabled = enable != 0;
  });
}

uint32_t prestige_timing_count(prestige_ctx* ctx) {
  if (!ctx) return 0;
  uint32_t k = 0;
  prestige_status s = guard(ctx, [&] {
    PR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.flush();
    k = (uint32_t)ctx->timing.stats.size();
  });
  return s == PRESTIGE_OK ? k : 0;
}

prestige_status prestige_timing_get(prestige_ctx* ctx, uint32_t idx, const char** name, double* total_ms,
                                    uint64_t* launches, double* bytes_per_launch) {
  return guard(ctx, [&] {
    PR_HIP(hipStreamSynchronize(ctx->stream));
    ctx->timing.flush();
    PR_CHECK(idx < ctx->timing.stats.size(), PRESTIGE_E_INVALID, "timing index out of range");
    const KernelStat& s = ctx->timing.stats[idx];
    if (name) *name = s.name;
    if (total_ms) *total_ms = s.total_ms;
    if (launches) *launches = s.launches;
    if (bytes_per_launch) *bytes_per_launch = s.bytes_per_launch;
  });
}

prestige_status prestige_sync(prestige_ctx* ctx) {
  return guard(ctx, [&] { PR_HIP(hipStreamSynchronize(ctx->stream)); });
}

const char* prestige_last_error(const prestige_ctx* ctx) {
  if (!ctx) return g_create_error.c_str();
  return ctx->err.c_str();
}

}  // extern "C"
