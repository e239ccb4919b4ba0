// scene.cpp — deterministic particle layouts of the benchmark workloads.
//
// Draws come from SplitMix64 (Steele, Lea & Flood 2014) mapped to [0, 1) as
// (state >> 11) * 2^-53, so the same seed gives the same particles on every
// host.  Geometry choices (count vs block size, tank size, pour lattice) are
// frozen here; the reference defines no scene.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/prestige_hip.h"

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

constexpr double kPi = 3.14159265358979323846;
constexpr double kG = 9.81;

// ---- dam break: fluid cube of side s*dx in the corner of a 3-layer tank ----
struct DamDims {
  uint64_t side, nix, niy, niz, n_fluid, n_wall;
};

DamDims dam_dims(const prestige_scene_params& sp) {
  DamDims d{};
  uint64_t target = sp.n_target ? sp.n_target : 1000000ull;
  uint64_t s = (uint64_t)std::cbrt((double)target);
  while (s * s * s < target) ++s;
  while (s > 1 && (s - 1) * (s - 1) * (s - 1) >= target) --s;
  d.side = s;
  // tank inner size 3.2 x 1.34 x 2.4 block sides (1.6 x 0.67 x 1.2 m for a 0.5 m block)
  d.nix = (uint64_t)std::llround(3.2 * (double)s);
  d.niy = (uint64_t)std::llround(1.34 * (double)s);
  d.niz = (uint64_t)std::llround(2.4 * (double)s);
  d.n_fluid = s * s * s;
  uint64_t ox = d.nix + 6, oy = d.niy + 6, oz = d.niz + 3;
  d.n_wall = ox * oy * oz - d.nix * d.niy * d.niz;
  return d;
}

// ---- DEM lattice: 0.5 m box, r = 0.01, spacing 0.021 ----
constexpr double kLatR = 0.01, kLatBox = 0.5, kLatSpacing = 0.021, kSolidRho = 2500.0;

uint64_t lattice_sites_per_side() {
  return (uint64_t)std::floor((kLatBox - 2.0 * kLatR) / kLatSpacing + 1e-9) + 1;
}

// ---- DEM pour: 0.3 m bin, r ~ U(0.008, 0.012), lattice spacing 0.025 ----
constexpr double kPourBin = 0.3, kPourRmin = 0.008, kPourRmax = 0.012, kPourSpacing = 0.025;

uint64_t pour_per_row() { return (uint64_t)std::floor(kPourBin / kPourSpacing + 1e-9); }

void dem_common(prestige_desc* d) {
  prestige_desc_init(d, PRESTIGE_MODEL_DEM);
  d->kn = 1.0e5;
  d->kt = 2.0 / 7.0 * d->kn;
  d->restitution = 0.3;
  d->mu = 0.5;
  d->dt = 1.0e-5;
  d->max_contacts = 16;
}

}  // namespace

extern "C" {

prestige_status prestige_scene_count(int32_t scene, const prestige_scene_params* sp, uint64_t* n_total,
                                     uint64_t* n_moving) {
  if (!sp || !n_total || !n_moving) return PRESTIGE_E_INVALID;
  switch (scene) {
    case PRESTIGE_SCENE_DEM_LATTICE: {
      uint64_t n = sp->n_target ? sp->n_target : 10000ull;
      uint64_t ns = lattice_sites_per_side();
      if (n > ns * ns * ns) return PRESTIGE_E_INVALID;
      *n_total = n;
      *n_moving = n;
      return PRESTIGE_OK;
    }
    case PRESTIGE_SCENE_DAM_BREAK: {
      if (!(sp->dx > 0.0)) return PRESTIGE_E_INVALID;
      DamDims d = dam_dims(*sp);
      *n_total = d.n_fluid + d.n_wall;
      *n_moving = d.n_fluid;
      return PRESTIGE_OK;
    }
    case PRESTIGE_SCENE_DEM_POUR: {
      uint64_t n = sp->n_target ? sp->n_target : 100000ull;
      *n_total = n;
      *n_moving = n;
      return PRESTIGE_OK;
    }
    default:
      return PRESTIGE_E_INVALID;
  }
}

prestige_status prestige_scene_build(int32_t scene, const prestige_scene_params* sp, prestige_desc* desc,
                                     double* x, double* y, double* z, double* vx, double* vy, double* vz,
                                     double* radius, double* mass) {
  uint64_t n = 0, nm = 0;
  prestige_status st = prestige_scene_count(scene, sp, &n, &nm);
  if (st != PRESTIGE_OK) return st;
  if (!desc || !x || !y || !z || !vx || !vy || !vz) return PRESTIGE_E_INVALID;
  SplitMix64 rng(sp->seed);
  if (scene == PRESTIGE_SCENE_DAM_BREAK) {
    const double dx = sp->dx;
    DamDims d = dam_dims(*sp);
    prestige_desc_init(desc, PRESTIGE_MODEL_WCSPH);
    desc->n_particles = n;
    desc->n_moving = nm;
    desc->h = 1.2 * dx;
    desc->cell_size = 2.0 * desc->h;
    desc->rho0 = 1000.0;
    desc->mass = desc->rho0 * dx * dx * dx;
    const double H = (double)d.side * dx;   // initial water height
    desc->c0 = 10.0 * std::sqrt(2.0 * kG * H);
    desc->alpha = 0.1;
    desc->dt = 0.25 * desc->h / desc->c0;
    desc->gravity[0] = 0.0; desc->gravity[1] = 0.0; desc->gravity[2] = -kG;
    desc->grid_lo[0] = -3.0 * dx; desc->grid_lo[1] = -3.0 * dx; desc->grid_lo[2] = -3.0 * dx;
    desc->grid_hi[0] = ((double)d.nix + 3.0) * dx;
    desc->grid_hi[1] = ((double)d.niy + 3.0) * dx;
    desc->grid_hi[2] = ((double)d.niz + 3.0) * dx;
    uint64_t k = 0;
    for (uint64_t c = 0; c < d.side; ++c)
      for (uint64_t b = 0; b < d.side; ++b)
        for (uint64_t a = 0; a < d.side; ++a, ++k) {
          x[k] = ((double)a + 0.5) * dx;
          y[k] = ((double)b + 0.5) * dx;
          z[k] = ((double)c + 0.5) * dx;
          vx[k] = vy[k] = vz[k] = 0.0;
        }
    const int64_t nix = (int64_t)d.nix, niy = (int64_t)d.niy, niz = (int64_t)d.niz;
    for (int64_t c = -3; c < niz; ++c)
      for (int64_t b = -3; b < niy + 3; ++b)
        for (int64_t a = -3; a < nix + 3; ++a) {
          bool inside = a >= 0 && a < nix && b >= 0 && b < niy && c >= 0;
          if (inside) continue;
          x[k] = ((double)a + 0.5) * dx;
          y[k] = ((double)b + 0.5) * dx;
          z[k] = ((double)c + 0.5) * dx;
          vx[k] = vy[k] = vz[k] = 0.0;
          ++k;
        }
    if (radius) for (uint64_t i = 0; i < n; ++i) radius[i] = 0.0;
    if (mass) for (uint64_t i = 0; i < n; ++i) mass[i] = desc->mass;
    return k == n ? PRESTIGE_OK : PRESTIGE_E_INVALID;
  }
  if (!radius || !mass) return PRESTIGE_E_INVALID;
  if (scene == PRESTIGE_SCENE_DEM_LATTICE) {
    dem_common(desc);
    desc->n_particles = n;
    desc->n_moving = n;
    desc->cell_size = 2.0 * kLatR;
    for (int a = 0; a < 3; ++a) {
      desc->grid_lo[a] = 0.0;
      desc->grid_hi[a] = kLatBox;
      desc->wall_lo[a] = 0.0;
      desc->wall_hi[a] = kLatBox;
    }
    desc->wall_mask = 0x3F;
    const uint64_t ns = lattice_sites_per_side();
    const double first = 0.5 * (kLatBox - (double)(ns - 1) * kLatSpacing);
    const double m = kSolidRho * (4.0 / 3.0) * kPi * kLatR * kLatR * kLatR;
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t a = i % ns, b = (i / ns) % ns, c = i / (ns * ns);
      double jx = rng.uniform(-5e-4, 5e-4), jy = rng.uniform(-5e-4, 5e-4), jz = rng.uniform(-5e-4, 5e-4);
      x[i] = first + (double)a * kLatSpacing + jx;
      y[i] = first + (double)b * kLatSpacing + jy;
      z[i] = first + (double)c * kLatSpacing + jz;
      vx[i] = rng.uniform(-0.1, 0.1);
      vy[i] = rng.uniform(-0.1, 0.1);
      vz[i] = rng.uniform(-0.1, 0.1);
      radius[i] = kLatR;
      mass[i] = m;
    }
    return PRESTIGE_OK;
  }
  if (scene == PRESTIGE_SCENE_DEM_POUR) {
    dem_common(desc);
    desc->n_particles = n;
    desc->n_moving = n;
    desc->cell_size = 2.0 * kPourRmax;
    const uint64_t row = pour_per_row();
    const uint64_t layers = (n + row * row - 1) / (row * row);
    const double top = ((double)layers + 0.5) * kPourSpacing + 0.05;
    for (int a = 0; a < 3; ++a) {
      desc->grid_lo[a] = 0.0;
      desc->wall_lo[a] = 0.0;
    }
    desc->grid_hi[0] = kPourBin; desc->grid_hi[1] = kPourBin; desc->grid_hi[2] = top;
    desc->wall_hi[0] = kPourBin; desc->wall_hi[1] = kPourBin; desc->wall_hi[2] = top;
    desc->wall_mask = 0x1F;   // four sides and the floor; open top
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t a = i % row, b = (i / row) % row, c = i / (row * row);
      double r = rng.uniform(kPourRmin, kPourRmax);
      double jx = rng.uniform(-5e-4, 5e-4), jy = rng.uniform(-5e-4, 5e-4), jz = rng.uniform(-5e-4, 5e-4);
      x[i] = ((double)a + 0.5) * kPourSpacing + jx;
      y[i] = ((double)b + 0.5) * kPourSpacing + jy;
      z[i] = ((double)c + 0.5) * kPourSpacing + jz;
      // a stirred stream: +-0.5 m/s closes ~2.5 mm of gap in 500 steps, so
      // the contacts the calibration differentiates through form inside the run
      vx[i] = rng.uniform(-0.5, 0.5);
      vy[i] = rng.uniform(-0.5, 0.5);
      vz[i] = rng.uniform(-0.5, 0.5);
      radius[i] = r;
      mass[i] = kSolidRho * (4.0 / 3.0) * kPi * r * r * r;
    }
    return PRESTIGE_OK;
  }
  return PRESTIGE_E_INVALID;
}

}  // extern "C"
