/*
 * oracle.h — CPU FP64 restatement of the particle hot path.  TEST
 * INFRASTRUCTURE ONLY: tests/, the smoke check and bench.py's CPU-baseline
 * leg load liboracle.so as the checker; nothing in prestige_amd/ links,
 * loads or calls it.
 *
 * What it restates (paths relative to the reference tree):
 *   - the generated loop of prestige/src/codegen/simple_cpu.rs:7-19: i outer,
 *     j inner, j == i included, fused bodies applied in order per (i, j),
 *     gather form (writes only through [i]);
 *   - eq1, force[i] += mass[j] (prestige/src/lib.rs:7-12), both over all
 *     pairs (the literal loop) and over the 27-cell neighbour iterator;
 *   - the neighbour iterator, WCSPH and DEM bodies and the integrator that
 *     the reference lacks, written from the textbook equations (cubic spline
 *     of Monaghan & Lattanzio 1985; Tait EOS and artificial viscosity of
 *     Monaghan 1992/1994; linear spring-dashpot + Coulomb of Cundall & Strack
 *     1979; symplectic Euler).
 * The same source is compiled twice: REAL = double (the FP64 oracle) and
 * REAL = double complex (complex-step forward-mode derivatives, Squire &
 * Trapp 1998: f(x + i*eps*d) = f(x) + i*eps*J*d + O(eps^2), read off as
 * Im/eps with eps = 1e-30, so the tangent is exact to FP64 rounding).
 *
 * Build: cc -O3 -march=x86-64-v3 -ffp-contract=off -fopenmp (oracle/Makefile).
 * -ffp-contract=off keeps every a*b+c as two roundings, the same as the
 * device FP64 kernels, so the FP64 oracle and the FP64 device path can be
 * compared bit for bit.
 */
#ifndef PRESTIGE_ORACLE_H
#define PRESTIGE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_E_INVALID = 1, ORC_E_CAPACITY = 4, ORC_E_DOMAIN = 6 };
enum { ORC_MODEL_PAIRS = 0, ORC_MODEL_WCSPH = 1, ORC_MODEL_DEM = 2 };
#define ORC_MAX_CONTACTS 32

typedef struct orc_params {
  int32_t model;
  uint32_t max_contacts;
  uint64_t n, n_moving;
  double lo[3];
  double cs;
  int32_t dims[3];
  int32_t pad0;
  double g[3];
  double dt;
  double h, rho0, c0, alpha, mass;
  double kn, kt, e, mu;
  double wall_lo[3], wall_hi[3];
  uint32_t wall_mask;
  uint32_t pad1;
} orc_params;

typedef struct orc_state orc_state;       /* FP64 */
typedef struct orc_state_cs orc_state_cs; /* complex step */

/* ---- building blocks (pure functions) ---- */
/* Cell key of one point: key = (cz*ny + cy)*nx + cx with c = floor((x-lo)/cs)
 * evaluated as floor((x - lo) * (1/cs)); returns -1 if outside the grid. */
int64_t orc_cell_key(const orc_params* P, double x, double y, double z);
/* Stable counting sort of keys (all < n_keys): perm[k] = source slot. */
void orc_stable_sort(uint64_t n, const uint32_t* keys, uint32_t n_keys, uint32_t* perm);
/* lb[c] = first k with sorted_keys[k] >= c, c in [0, n_cells]. */
void orc_cell_lb(uint64_t n, const uint32_t* sorted_keys, uint32_t n_cells, uint32_t* lb);
/* Cubic spline shape f(q) (W = sigma f) and df/dq. */
double orc_spline_f(double q);
double orc_spline_df(double q);

/* eq1 over all pairs, literally for i { for j { force[i] += mass[j] } }. */
void orc_eq1_allpairs(uint64_t n, double* force, const double* mass);

/* ---- FP64 state ---- */
orc_state* orc_create(const orc_params* P, const double* x, const double* y,
                      const double* z, const double* vx, const double* vy,
                      const double* vz, const double* radius, const double* mass);
void orc_destroy(orc_state* s);
/* Neighbour build (keys, stable sort, reorder, lb); returns ORC_* status. */
int orc_build(orc_state* s);
/* Copies of the last build: sorted keys, perm, ids (n each), lb (n_cells+1). */
void orc_neighbors(const orc_state* s, uint32_t* keys, uint32_t* perm,
                   uint32_t* ids, uint32_t* lb);
/* eq1 through the cell iterator over the last build, |x_i-x_j| < cutoff:
 * force[i] += mass[j] for every neighbour j (j == i included).  force and mass
 * are in id order. */
void orc_eq1_cells(const orc_state* s, double cutoff, double* force, const double* mass);
/* One rebuild + density/EOS (all particles) only, no integration. */
int orc_density(orc_state* s);
/* nsteps full steps; OpenMP threads = omp_get_max_threads(). */
int orc_step(orc_state* s, uint32_t nsteps);
/* Set the OpenMP thread count of later calls (n >= 1); returns the previous
 * omp_get_max_threads(). */
int orc_set_threads(int n);
/* Field in id order: x y z vx vy vz wx wy wz rho p ax ay az radius mass key. */
int orc_get(const orc_state* s, const char* name, double* out);
/* Number of active DEM contacts summed over particles (pairs counted twice). */
uint64_t orc_contact_count(const orc_state* s);

/* ---- complex-step tangent ----
 * Runs nsteps from the given initial state with the parameters perturbed by
 * i*eps*dparam and the initial velocities by i*eps*dv0 (3*n, id order, x
 * block then y then z; NULL = 0).  dparam: WCSPH {alpha, c0}, DEM {kn, e, mu}.
 * Writes the FP64 final state (x_out, v_out: 3*n id order, may be NULL) and
 * its tangent (tx, tv: 3*n id order). */
int orc_cs_tangent(const orc_params* P, const double* x, const double* y,
                   const double* z, const double* vx, const double* vy,
                   const double* vz, const double* radius, const double* mass,
                   const double* dparam, const double* dv0, uint32_t nsteps,
                   double* x_out, double* v_out, double* tx, double* tv);

#ifdef __cplusplus
}
#endif
#endif
