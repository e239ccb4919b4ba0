/*
 * oracle.c — CPU FP64 oracle (test infrastructure; see oracle.h).
 *
 * The hot-path semantics come from prestige/src/codegen/simple_cpu.rs:7-19
 * (loop order) and prestige/src/lib.rs:7-12 (eq1); the physics bodies are
 * restated from the textbook equations named in oracle.h.  The same physics
 * source (physics_impl.h) is instantiated for double and for double complex.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* building blocks                                                         */
/* ---------------------------------------------------------------------- */

int64_t orc_cell_key(const orc_params* P, double x, double y, double z) {
  const double inv = 1.0 / P->cs;
  double fx = floor((x - P->lo[0]) * inv);
  double fy = floor((y - P->lo[1]) * inv);
  double fz = floor((z - P->lo[2]) * inv);
  if (!(fx >= 0.0 && fy >= 0.0 && fz >= 0.0)) return -1;
  if (!(fx < (double)P->dims[0] && fy < (double)P->dims[1] && fz < (double)P->dims[2])) return -1;
  int64_t cx = (int64_t)fx, cy = (int64_t)fy, cz = (int64_t)fz;
  return (cz * P->dims[1] + cy) * P->dims[0] + cx;
}

void orc_stable_sort(uint64_t n, const uint32_t* keys, uint32_t n_keys, uint32_t* perm) {
  uint64_t* count = (uint64_t*)calloc((size_t)n_keys + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) count[keys[i] + 1]++;
  for (uint32_t c = 0; c < n_keys; ++c) count[c + 1] += count[c];
  for (uint64_t i = 0; i < n; ++i) perm[count[keys[i]]++] = (uint32_t)i;
  free(count);
}

void orc_cell_lb(uint64_t n, const uint32_t* sorted_keys, uint32_t n_cells, uint32_t* lb) {
  uint64_t k = 0;
  for (uint32_t c = 0; c <= n_cells; ++c) {
    while (k < n && sorted_keys[k] < c) ++k;
    lb[c] = (uint32_t)k;
  }
}

/* the WCSPH passes' fused forms of the cubic spline (physics_impl.h) */
double orc_spline_f(double q) {
  if (q < 1.0) return fma(q * q, fma(0.75, q, -1.5), 1.0);
  if (q < 2.0) {
    double t = 2.0 - q;
    return 0.25 * t * t * t;
  }
  return 0.0;
}

double orc_spline_df(double q) {
  if (q < 1.0) return q * fma(2.25, q, -3.0);
  if (q < 2.0) {
    double t = 2.0 - q;
    return -0.75 * t * t;
  }
  return 0.0;
}

/* prestige/src/codegen/simple_cpu.rs:7-19 with the body of prestige/src/lib.rs:10 */
void orc_eq1_allpairs(uint64_t n, double* force, const double* mass) {
  for (uint64_t i = 0; i < n; ++i) {
    for (uint64_t j = 0; j < n; ++j) {
      force[i] += mass[j];
    }
  }
}

/* ---------------------------------------------------------------------- */
/* FP64 instantiation                                                      */
/* ---------------------------------------------------------------------- */
#define REAL double
#define NAME(f) f##_d
#define STATE orc_state
#define RE(v) (v)
#define SQRT(v) sqrt(v)
#define LOG(v) log(v)
#define FMA(a, b, c) fma((a), (b), (c))
#include "physics_impl.h"
#undef REAL
#undef NAME
#undef STATE
#undef RE
#undef SQRT
#undef LOG
#undef FMA

/* ---------------------------------------------------------------------- */
/* complex-step instantiation                                              */
/* ---------------------------------------------------------------------- */
#define REAL double complex
#define NAME(f) f##_cs
#define STATE orc_state_cs
#define RE(v) creal(v)
#define SQRT(v) csqrt(v)
#define LOG(v) clog(v)
/* complex step: the tangent of a fused multiply-add is that of a b + c */
#define FMA(a, b, c) ((a) * (b) + (c))
#include "physics_impl.h"
#undef REAL
#undef NAME
#undef STATE
#undef RE
#undef SQRT
#undef LOG
#undef FMA

/* ---------------------------------------------------------------------- */
/* public FP64 API                                                         */
/* ---------------------------------------------------------------------- */

orc_state* orc_create(const orc_params* P, const double* x, const double* y,
                      const double* z, const double* vx, const double* vy,
                      const double* vz, const double* radius, const double* mass) {
  if (!P || P->max_contacts > ORC_MAX_CONTACTS) return NULL;
  return state_new_d(P, x, y, z, vx, vy, vz, radius, mass);
}

void orc_destroy(orc_state* s) { state_free_d(s); }

int orc_build(orc_state* s) { return build_d(s); }

void orc_neighbors(const orc_state* s, uint32_t* keys, uint32_t* perm, uint32_t* ids,
                   uint32_t* lb) {
  if (keys) memcpy(keys, s->key, s->n * sizeof(uint32_t));
  if (perm) memcpy(perm, s->perm, s->n * sizeof(uint32_t));
  if (ids) memcpy(ids, s->id, s->n * sizeof(uint32_t));
  if (lb) memcpy(lb, s->lb, ((size_t)s->n_cells + 1) * sizeof(uint32_t));
}

void orc_eq1_cells(const orc_state* s, double cutoff, double* force, const double* mass) {
  const double c2 = cutoff * cutoff;
  int64_t n = (int64_t)s->n;
  /* gather into [i]: bodies see only the sorted slots, results go back by id */
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t rlo[9], rhi[9];
    int nr = runs_d(s, (uint64_t)i, rlo, rhi);
    double acc = force[s->id[i]];
    for (int r = 0; r < nr; ++r) {
      for (uint32_t j = rlo[r]; j < rhi[r]; ++j) {
        double dx = s->x[i] - s->x[j], dy = s->y[i] - s->y[j], dz = s->z[i] - s->z[j];
        double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < c2) acc += mass[s->id[j]];
      }
    }
    force[s->id[i]] = acc;
  }
}

int orc_density(orc_state* s) {
  int st = build_d(s);
  if (st) return st;
  density_d(s);
  return ORC_OK;
}

int orc_step(orc_state* s, uint32_t nsteps) { return step_d(s, nsteps); }

int orc_set_threads(int n) {
#ifdef _OPENMP
  int prev = omp_get_max_threads();
  if (n >= 1) omp_set_num_threads(n);
  return prev;
#else
  (void)n;
  return 1;
#endif
}

static const double* pick_field_d(const orc_state* s, const char* name) {
  struct { const char* n; const double* p; } tab[] = {
      {"x", s->x}, {"y", s->y}, {"z", s->z}, {"vx", s->vx}, {"vy", s->vy},
      {"vz", s->vz}, {"wx", s->wx}, {"wy", s->wy}, {"wz", s->wz}, {"rho", s->rho},
      {"p", s->p}, {"ax", s->ax}, {"ay", s->ay}, {"az", s->az}, {"radius", s->rad},
      {"mass", s->mass}, {"fx", s->fx}, {"fy", s->fy}, {"fz", s->fz}};
  for (size_t k = 0; k < sizeof(tab) / sizeof(tab[0]); ++k)
    if (strcmp(tab[k].n, name) == 0) return tab[k].p;
  return NULL;
}

int orc_get(const orc_state* s, const char* name, double* out) {
  if (strcmp(name, "key") == 0) {
    for (uint64_t k = 0; k < s->n; ++k) out[s->id[k]] = (double)s->key[k];
    return ORC_OK;
  }
  const double* src = pick_field_d(s, name);
  if (!src) return ORC_E_INVALID;
  for (uint64_t k = 0; k < s->n; ++k) out[s->id[k]] = src[k];
  return ORC_OK;
}

uint64_t orc_contact_count(const orc_state* s) {
  uint64_t c = 0;
  if (!s->hcnt) return 0;
  for (uint64_t k = 0; k < s->n; ++k) c += s->hcnt[k];
  return c;
}

/* ---------------------------------------------------------------------- */
/* complex-step tangent                                                    */
/* ---------------------------------------------------------------------- */

int orc_cs_tangent(const orc_params* P, const double* x, const double* y,
                   const double* z, const double* vx, const double* vy,
                   const double* vz, const double* radius, const double* mass,
                   const double* dparam, const double* dv0, uint32_t nsteps,
                   double* x_out, double* v_out, double* tx, double* tv) {
  const double eps = 1e-30;
  if (!P || P->max_contacts > ORC_MAX_CONTACTS) return ORC_E_INVALID;
  orc_state_cs* s = state_new_cs(P, x, y, z, vx, vy, vz, radius, mass);
  uint64_t n = P->n;
  double d0 = dparam ? dparam[0] : 0.0, d1 = dparam ? dparam[1] : 0.0,
         d2 = (dparam && P->model == ORC_MODEL_DEM) ? dparam[2] : 0.0;
  if (P->model == ORC_MODEL_WCSPH) {
    set_params_cs(s, P->alpha + I * eps * d0, P->c0 + I * eps * d1, P->kn, P->e, P->mu);
  } else {
    set_params_cs(s, P->alpha, P->c0, P->kn + I * eps * d0, P->e + I * eps * d1,
                  P->mu + I * eps * d2);
  }
  if (dv0) {
    for (uint64_t i = 0; i < n; ++i) {
      s->vx[i] += I * eps * dv0[i];
      s->vy[i] += I * eps * dv0[n + i];
      s->vz[i] += I * eps * dv0[2 * n + i];
    }
  }
  int st = step_cs(s, nsteps);
  for (uint64_t k = 0; k < n; ++k) {
    uint64_t i = s->id[k];
    if (x_out) {
      x_out[i] = creal(s->x[k]);
      x_out[n + i] = creal(s->y[k]);
      x_out[2 * n + i] = creal(s->z[k]);
    }
    if (v_out) {
      v_out[i] = creal(s->vx[k]);
      v_out[n + i] = creal(s->vy[k]);
      v_out[2 * n + i] = creal(s->vz[k]);
    }
    if (tx) {
      tx[i] = cimag(s->x[k]) / eps;
      tx[n + i] = cimag(s->y[k]) / eps;
      tx[2 * n + i] = cimag(s->z[k]) / eps;
    }
    if (tv) {
      tv[i] = cimag(s->vx[k]) / eps;
      tv[n + i] = cimag(s->vy[k]) / eps;
      tv[2 * n + i] = cimag(s->vz[k]) / eps;
    }
  }
  state_free_cs(s);
  return st;
}
