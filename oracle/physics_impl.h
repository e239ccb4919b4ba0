/*
 * physics_impl.h — the oracle's state and step, written once for a scalar
 * type REAL.  oracle.c includes it twice:
 *   REAL = double          NAME(x) = x##_d    RE(v) = v          (FP64 oracle)
 *   REAL = double complex  NAME(x) = x##_cs   RE(v) = creal(v)   (complex step)
 * Branches (cutoff tests, Coulomb cap, p >= 0 clamp, cell keys) take the real
 * part, so the complex instantiation differentiates the same piecewise
 * function the FP64 one evaluates.
 *
 * Loop semantics follow prestige/src/codegen/simple_cpu.rs:7-19 (i outer, j
 * inner, j == i visited, bodies in order, writes only through [i]) with the
 * j range replaced by the 27 cells around i: for dz in -1..1, dy in -1..1 the
 * three cells cx-1..cx+1 of that row are one contiguous run of the key-sorted
 * order, visited in ascending slot order.  DEM visits the runs row by row
 * (dz outer, dy inner); the WCSPH passes visit opposite rows one after the
 * other (sph_runs), the order in which the device walks them.
 */

typedef struct STATE {
  orc_params P;
  uint64_t n;
  uint32_t n_cells;
  /* sorted-order arrays */
  REAL *x, *y, *z, *vx, *vy, *vz, *wx, *wy, *wz, *rad, *mass;
  REAL *rho, *p, *pr, *ax, *ay, *az;       /* pr = p / rho^2 */
  REAL *fx, *fy, *fz, *tqx, *tqy, *tqz;    /* DEM force / torque */
  uint32_t *id, *key, *perm, *lb;
  /* DEM tangential history, keyed by partner id, moved with the particle */
  uint32_t *hcnt, *hid, *hcnt2, *hid2;
  REAL *hxi, *hxi2;                        /* [n][maxc][3] */
  REAL *wxi, *wxi2;                        /* [n][6][3] walls */
  /* scratch */
  REAL* tmp;
  uint32_t* utmp;
  /* derived constants */
  REAL alpha_c0, B;
  double inv_h, H2, eps2, inv_rho0, sigma, msig, msh, inv_cs;
  REAL two_zeta, kn, kt, mu, neg_inv_kt;
  int built;
} STATE;

static void NAME(derive)(STATE* s) {
  const orc_params* P = &s->P;
  s->inv_cs = 1.0 / P->cs;
  if (P->model == ORC_MODEL_WCSPH) {
    double h = P->h;
    s->inv_h = 1.0 / h;
    s->H2 = (2.0 * h) * (2.0 * h);
    s->eps2 = (0.01 * h) * h;
    s->inv_rho0 = 1.0 / P->rho0;
    s->sigma = 1.0 / (((M_PI * h) * h) * h);
    s->msig = P->mass * s->sigma;
    s->msh = (P->mass * s->sigma) * s->inv_h;
  }
}

static void NAME(set_params)(STATE* s, REAL alpha, REAL c0, REAL kn, REAL e, REAL mu) {
  const orc_params* P = &s->P;
  s->alpha_c0 = alpha * c0;
  s->B = ((P->rho0 * c0) * c0) / 7.0;
  /* linear spring-dashpot: zeta = -ln e / sqrt(pi^2 + ln^2 e), cn = 2 zeta sqrt(meff kn) */
  if (RE(e) > 0.0) {
    REAL le = LOG(e);
    s->two_zeta = 2.0 * (-le / SQRT(M_PI * M_PI + le * le));
  } else {
    s->two_zeta = 0.0;
  }
  s->kn = kn;
  /* kt tracks kn at the configured ratio (kt = 2/7 kn in the benchmark) */
  s->kt = P->kn > 0.0 ? P->kt * (kn / P->kn) : P->kt;
  s->mu = mu;
  s->neg_inv_kt = RE(s->kt) > 0.0 ? -1.0 / s->kt : 0.0;
}

#define ALLOC_R(n) ((REAL*)calloc((size_t)(n), sizeof(REAL)))
#define ALLOC_U(n) ((uint32_t*)calloc((size_t)(n), sizeof(uint32_t)))

static STATE* NAME(state_new)(const orc_params* P, const double* x, const double* y,
                              const double* z, const double* vx, const double* vy,
                              const double* vz, const double* radius,
                              const double* mass) {
  STATE* s = (STATE*)calloc(1, sizeof(STATE));
  s->P = *P;
  uint64_t n = P->n;
  s->n = n;
  s->n_cells = (uint32_t)P->dims[0] * (uint32_t)P->dims[1] * (uint32_t)P->dims[2];
  REAL** fields[] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->wx, &s->wy,
                     &s->wz, &s->rad, &s->mass, &s->rho, &s->p, &s->pr, &s->ax,
                     &s->ay, &s->az, &s->fx, &s->fy, &s->fz, &s->tqx, &s->tqy,
                     &s->tqz, &s->tmp};
  for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); ++f) *fields[f] = ALLOC_R(n ? n : 1);
  s->id = ALLOC_U(n ? n : 1);
  s->key = ALLOC_U(n ? n : 1);
  s->perm = ALLOC_U(n ? n : 1);
  s->utmp = ALLOC_U(n ? n : 1);
  s->lb = ALLOC_U((uint64_t)s->n_cells + 1);
  uint32_t mc = P->max_contacts ? P->max_contacts : 1;
  if (P->model == ORC_MODEL_DEM) {
    s->hcnt = ALLOC_U(n); s->hcnt2 = ALLOC_U(n);
    s->hid = ALLOC_U(n * mc); s->hid2 = ALLOC_U(n * mc);
    s->hxi = ALLOC_R(n * mc * 3); s->hxi2 = ALLOC_R(n * mc * 3);
    s->wxi = ALLOC_R(n * 18); s->wxi2 = ALLOC_R(n * 18);
  }
  for (uint64_t i = 0; i < n; ++i) {
    s->x[i] = x[i]; s->y[i] = y[i]; s->z[i] = z[i];
    s->vx[i] = vx[i]; s->vy[i] = vy[i]; s->vz[i] = vz[i];
    s->rad[i] = radius ? radius[i] : 0.0;
    s->mass[i] = mass ? mass[i] : P->mass;
    s->id[i] = (uint32_t)i;
  }
  NAME(derive)(s);
  NAME(set_params)(s, P->alpha, P->c0, P->kn, P->e, P->mu);
  return s;
}

static void NAME(state_free)(STATE* s) {
  if (!s) return;
  REAL* rf[] = {s->x, s->y, s->z, s->vx, s->vy, s->vz, s->wx, s->wy, s->wz, s->rad,
                s->mass, s->rho, s->p, s->pr, s->ax, s->ay, s->az, s->fx, s->fy,
                s->fz, s->tqx, s->tqy, s->tqz, s->tmp, s->hxi, s->hxi2, s->wxi, s->wxi2};
  for (size_t f = 0; f < sizeof(rf) / sizeof(rf[0]); ++f) free(rf[f]);
  uint32_t* uf[] = {s->id, s->key, s->perm, s->utmp, s->lb, s->hcnt, s->hcnt2, s->hid, s->hid2};
  for (size_t f = 0; f < sizeof(uf) / sizeof(uf[0]); ++f) free(uf[f]);
  free(s);
}

static void NAME(permute)(STATE* s, REAL** a) {
  REAL* src = *a;
  REAL* dst = s->tmp;
  const uint32_t* perm = s->perm;
  int64_t n = (int64_t)s->n;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < n; ++k) dst[k] = src[perm[k]];
  s->tmp = src;
  *a = dst;
}

/* Keys of the current positions, stable sort, reorder of every field, lb. */
static int NAME(build)(STATE* s) {
  const orc_params* P = &s->P;
  int64_t n = (int64_t)s->n;
  int bad = 0;
  for (int64_t i = 0; i < n; ++i) {
    int64_t k = orc_cell_key(P, RE(s->x[i]), RE(s->y[i]), RE(s->z[i]));
    if (k < 0) { bad = 1; k = 0; }
    s->key[i] = (uint32_t)k;
  }
  if (bad) return ORC_E_DOMAIN;
  orc_stable_sort(s->n, s->key, s->n_cells, s->perm);
  /* keys and ids follow the permutation */
  for (int64_t k = 0; k < n; ++k) s->utmp[k] = s->key[s->perm[k]];
  { uint32_t* t = s->key; s->key = s->utmp; s->utmp = t; }
  for (int64_t k = 0; k < n; ++k) s->utmp[k] = s->id[s->perm[k]];
  { uint32_t* t = s->id; s->id = s->utmp; s->utmp = t; }
  REAL** fields[] = {&s->x, &s->y, &s->z, &s->vx, &s->vy, &s->vz, &s->wx, &s->wy,
                     &s->wz, &s->rad, &s->mass};
  for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); ++f) NAME(permute)(s, fields[f]);
  if (P->model == ORC_MODEL_DEM) {
    uint32_t mc = P->max_contacts;
    for (int64_t k = 0; k < n; ++k) {
      uint32_t src = s->perm[k];
      s->hcnt2[k] = s->hcnt[src];
      for (uint32_t c = 0; c < mc; ++c) {
        s->hid2[(uint64_t)k * mc + c] = s->hid[(uint64_t)src * mc + c];
        for (int d = 0; d < 3; ++d)
          s->hxi2[((uint64_t)k * mc + c) * 3 + d] = s->hxi[((uint64_t)src * mc + c) * 3 + d];
      }
      for (int c = 0; c < 18; ++c) s->wxi2[(uint64_t)k * 18 + c] = s->wxi[(uint64_t)src * 18 + c];
    }
    { uint32_t* t = s->hcnt; s->hcnt = s->hcnt2; s->hcnt2 = t; }
    { uint32_t* t = s->hid; s->hid = s->hid2; s->hid2 = t; }
    { REAL* t = s->hxi; s->hxi = s->hxi2; s->hxi2 = t; }
    { REAL* t = s->wxi; s->wxi = s->wxi2; s->wxi2 = t; }
  }
  orc_cell_lb(s->n, s->key, s->n_cells, s->lb);
  s->built = 1;
  return ORC_OK;
}

/* The nine contiguous candidate runs of slot i: run r covers [lo[r], hi[r]). */
static int NAME(runs)(const STATE* s, uint64_t i, uint32_t* rlo, uint32_t* rhi) {
  const int32_t* D = s->P.dims;
  uint32_t k = s->key[i];
  int32_t cx = (int32_t)(k % (uint32_t)D[0]);
  int32_t cy = (int32_t)((k / (uint32_t)D[0]) % (uint32_t)D[1]);
  int32_t cz = (int32_t)(k / ((uint32_t)D[0] * (uint32_t)D[1]));
  int32_t x0 = cx > 0 ? cx - 1 : 0;
  int32_t x1 = cx < D[0] - 1 ? cx + 1 : D[0] - 1;
  int nr = 0;
  for (int32_t dz = -1; dz <= 1; ++dz) {
    int32_t zz = cz + dz;
    if (zz < 0 || zz >= D[2]) continue;
    for (int32_t dy = -1; dy <= 1; ++dy) {
      int32_t yy = cy + dy;
      if (yy < 0 || yy >= D[1]) continue;
      uint32_t row = ((uint32_t)zz * (uint32_t)D[1] + (uint32_t)yy) * (uint32_t)D[0];
      rlo[nr] = s->lb[row + (uint32_t)x0];
      rhi[nr] = s->lb[row + (uint32_t)x1 + 1];
      ++nr;
    }
  }
  return nr;
}

/* Summation density over all particles, then Tait EOS (gamma = 7), p >= 0. */
#ifndef ORC_SPH_ORDER
#define ORC_SPH_ORDER
/* WCSPH run order: run r = 3 (dz + 1) + (dy + 1) is followed by the
 * opposite row 8 - r, the centre row last. */
static const int orc_sph_order[9] = {0, 8, 1, 7, 2, 6, 3, 5, 4};
#endif

/* The nine runs of slot i in the WCSPH order (rows outside the grid left out). */
static int NAME(sph_runs)(const STATE* s, uint64_t i, uint32_t* rlo, uint32_t* rhi) {
  const int32_t* D = s->P.dims;
  uint32_t k = s->key[i];
  int32_t cx = (int32_t)(k % (uint32_t)D[0]);
  int32_t cy = (int32_t)((k / (uint32_t)D[0]) % (uint32_t)D[1]);
  int32_t cz = (int32_t)(k / ((uint32_t)D[0] * (uint32_t)D[1]));
  int32_t x0 = cx > 0 ? cx - 1 : 0;
  int32_t x1 = cx < D[0] - 1 ? cx + 1 : D[0] - 1;
  int nr = 0;
  for (int t = 0; t < 9; ++t) {
    const int r = orc_sph_order[t];
    int32_t zz = cz + r / 3 - 1, yy = cy + r % 3 - 1;
    if (zz < 0 || zz >= D[2] || yy < 0 || yy >= D[1]) continue;
    uint32_t row = ((uint32_t)zz * (uint32_t)D[1] + (uint32_t)yy) * (uint32_t)D[0];
    rlo[nr] = s->lb[row + (uint32_t)x0];
    rhi[nr] = s->lb[row + (uint32_t)x1 + 1];
    ++nr;
  }
  return nr;
}

static void NAME(density)(STATE* s) {
  int64_t n = (int64_t)s->n;
  const double H2 = s->H2, inv_h = s->inv_h;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    uint32_t rlo[9], rhi[9];
    int nr = NAME(sph_runs)(s, (uint64_t)i, rlo, rhi);
    REAL xi = s->x[i], yi = s->y[i], zi = s->z[i];
    REAL acc = 0.0;
    for (int r = 0; r < nr; ++r) {
      for (uint32_t j = rlo[r]; j < rhi[r]; ++j) {
        REAL dx = xi - s->x[j], dy = yi - s->y[j], dz = zi - s->z[j];
        /* fused forms (fma: one rounding), the device's operation order */
        REAL r2 = FMA(dx, dx, FMA(dy, dy, dz * dz));
        if (RE(r2) < H2) {
          REAL q = SQRT(r2) * inv_h;
          REAL w;
          if (RE(q) < 1.0) {
            /* 1 - 1.5 q^2 + 0.75 q^3 = 1 + q^2 (0.75 q - 1.5) */
            w = FMA(q * q, FMA(0.75, q, -1.5), 1.0);
          } else {
            REAL t = 2.0 - q;
            w = 0.25 * t * t * t;
          }
          acc = acc + w;
        }
      }
    }
    REAL rho = s->msig * acc;
    REAL ratio = rho * s->inv_rho0;
    REAL r2_ = ratio * ratio;
    REAL r4 = r2_ * r2_;
    REAL r7 = r4 * r2_ * ratio;
    REAL p = s->B * (r7 - 1.0);
    if (RE(p) < 0.0) p = 0.0;
    s->rho[i] = rho;
    s->p[i] = p;
    s->pr[i] = p / (rho * rho);
  }
}

/* Momentum with artificial viscosity for moving particles, then symplectic Euler. */
static void NAME(sph_force_integrate)(STATE* s) {
  const orc_params* P = &s->P;
  int64_t n = (int64_t)s->n;
  const double H2 = s->H2, inv_h = s->inv_h, eps2 = s->eps2, h = P->h, msh = s->msh;
  const double dt = P->dt;
  const REAL ach = s->alpha_c0 * h;
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t i = 0; i < n; ++i) {
    if (s->id[i] >= P->n_moving) {
      s->ax[i] = 0.0; s->ay[i] = 0.0; s->az[i] = 0.0;
      continue;
    }
    uint32_t rlo[9], rhi[9];
    int nr = NAME(sph_runs)(s, (uint64_t)i, rlo, rhi);
    REAL xi = s->x[i], yi = s->y[i], zi = s->z[i];
    REAL vxi = s->vx[i], vyi = s->vy[i], vzi = s->vz[i];
    REAL pri = s->pr[i], rhoi = s->rho[i];
    REAL sx = 0.0, sy = 0.0, sz = 0.0;
    for (int r = 0; r < nr; ++r) {
      for (uint32_t j = rlo[r]; j < rhi[r]; ++j) {
        REAL dx = xi - s->x[j], dy = yi - s->y[j], dz = zi - s->z[j];
        REAL r2 = FMA(dx, dx, FMA(dy, dy, dz * dz));
        if (RE(r2) < H2 && RE(r2) > 0.0) {
          REAL rr = SQRT(r2);
          REAL q = rr * inv_h;
          REAL dw;
          if (RE(q) < 1.0) {
            /* -3 q + 2.25 q^2 = q (2.25 q - 3) */
            dw = q * FMA(2.25, q, -3.0);
          } else {
            REAL t = 2.0 - q;
            dw = -0.75 * t * t;
          }
          REAL dvx = vxi - s->vx[j], dvy = vyi - s->vy[j], dvz = vzi - s->vz[j];
          REAL vdotr = FMA(dvx, dx, FMA(dvy, dy, dvz * dz));
          /* sc = (pr_i + pr_j + Pi_ij) f'(q) / r with Pi_ij = -alpha c0 h (v.r) /
           * ((r^2 + eps2) rhobar) for approaching pairs, put over one common
           * denominator so that every pair costs a single division */
          REAL num, den;
          if (RE(vdotr) < 0.0) {
            REAL drb = (r2 + eps2) * (0.5 * (rhoi + s->rho[j]));
            num = FMA(pri + s->pr[j], drb, -(ach * vdotr)) * dw;
            den = drb * rr;
          } else {
            num = (pri + s->pr[j]) * dw;
            den = rr;
          }
          REAL sc = num / den;
          sx = FMA(sc, dx, sx);
          sy = FMA(sc, dy, sy);
          sz = FMA(sc, dz, sz);
        }
      }
    }
    s->ax[i] = P->g[0] - msh * sx;
    s->ay[i] = P->g[1] - msh * sy;
    s->az[i] = P->g[2] - msh * sz;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (s->id[i] >= P->n_moving) continue;
    s->vx[i] = s->vx[i] + dt * s->ax[i];
    s->vy[i] = s->vy[i] + dt * s->ay[i];
    s->vz[i] = s->vz[i] + dt * s->az[i];
    s->x[i] = s->x[i] + dt * s->vx[i];
    s->y[i] = s->y[i] + dt * s->vy[i];
    s->z[i] = s->z[i] + dt * s->vz[i];
  }
}

/* Tangential spring update shared by pair and wall contacts: project the old
 * spring onto the tangent plane, add vt*dt, Coulomb-cap |Ft| <= mu |Fn|. */
static inline void NAME(tangential)(const STATE* s, REAL fn, REAL nx, REAL ny, REAL nz,
                                    REAL vtx, REAL vty, REAL vtz, REAL xox, REAL xoy,
                                    REAL xoz, REAL* ft, REAL* xn) {
  const double dt = s->P.dt;
  REAL xdn = xox * nx + xoy * ny + xoz * nz;
  REAL xpx = xox - xdn * nx, xpy = xoy - xdn * ny, xpz = xoz - xdn * nz;
  REAL xnx = xpx + vtx * dt, xny = xpy + vty * dt, xnz = xpz + vtz * dt;
  REAL ftx = -s->kt * xnx, fty = -s->kt * xny, ftz = -s->kt * xnz;
  REAL ftn = SQRT(ftx * ftx + fty * fty + ftz * ftz);
  REAL fcap = s->mu * (RE(fn) < 0.0 ? -fn : fn);
  if (RE(ftn) > RE(fcap)) {
    REAL sc = fcap / ftn;
    ftx = ftx * sc; fty = fty * sc; ftz = ftz * sc;
    xnx = ftx * s->neg_inv_kt; xny = fty * s->neg_inv_kt; xnz = ftz * s->neg_inv_kt;
  }
  ft[0] = ftx; ft[1] = fty; ft[2] = ftz;
  xn[0] = xnx; xn[1] = xny; xn[2] = xnz;
}

/* DEM contact forces (pairs then walls) with history, then symplectic Euler
 * for v, x and omega. */
static int NAME(dem_force_integrate)(STATE* s) {
  const orc_params* P = &s->P;
  int64_t n = (int64_t)s->n;
  const uint32_t mc = P->max_contacts;
  const double dt = P->dt;
  int overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
  for (int64_t i = 0; i < n; ++i) {
    s->hcnt2[i] = 0;
    for (int c = 0; c < 18; ++c) s->wxi2[(uint64_t)i * 18 + c] = 0.0;
    if (s->id[i] >= P->n_moving) {
      s->fx[i] = 0.0; s->fy[i] = 0.0; s->fz[i] = 0.0;
      s->tqx[i] = 0.0; s->tqy[i] = 0.0; s->tqz[i] = 0.0;
      continue;
    }
    uint32_t rlo[9], rhi[9];
    int nr = NAME(runs)(s, (uint64_t)i, rlo, rhi);
    REAL xi = s->x[i], yi = s->y[i], zi = s->z[i];
    REAL vxi = s->vx[i], vyi = s->vy[i], vzi = s->vz[i];
    REAL wxi = s->wx[i], wyi = s->wy[i], wzi = s->wz[i];
    REAL ri = s->rad[i], mi = s->mass[i];
    REAL Fx = 0.0, Fy = 0.0, Fz = 0.0, Tx = 0.0, Ty = 0.0, Tz = 0.0;
    uint32_t nc = 0;
    const uint32_t ocnt = s->hcnt[i];
    for (int r = 0; r < nr; ++r) {
      for (uint32_t j = rlo[r]; j < rhi[r]; ++j) {
        if (j == (uint32_t)i) continue;
        REAL dx = xi - s->x[j], dy = yi - s->y[j], dz = zi - s->z[j];
        REAL r2 = dx * dx + dy * dy + dz * dz;
        REAL R = ri + s->rad[j];
        if (!(RE(r2) < RE(R * R) && RE(r2) > 0.0)) continue;
        REAL d = SQRT(r2);
        REAL invd = 1.0 / d;
        REAL nx = dx * invd, ny = dy * invd, nz = dz * invd;
        REAL delta = R - d;
        REAL wsx = ri * wxi + s->rad[j] * s->wx[j];
        REAL wsy = ri * wyi + s->rad[j] * s->wy[j];
        REAL wsz = ri * wzi + s->rad[j] * s->wz[j];
        REAL cx = wsy * nz - wsz * ny, cy = wsz * nx - wsx * nz, cz = wsx * ny - wsy * nx;
        REAL vrx = (vxi - s->vx[j]) - cx, vry = (vyi - s->vy[j]) - cy, vrz = (vzi - s->vz[j]) - cz;
        REAL vn = vrx * nx + vry * ny + vrz * nz;
        REAL vtx = vrx - vn * nx, vty = vry - vn * ny, vtz = vrz - vn * nz;
        REAL mj = s->mass[j];
        REAL meff = mi * mj / (mi + mj);
        REAL cn = s->two_zeta * SQRT(meff * s->kn);
        REAL fn = s->kn * delta - cn * vn;
        uint32_t pid = s->id[j];
        REAL xox = 0.0, xoy = 0.0, xoz = 0.0;
        for (uint32_t c = 0; c < ocnt; ++c) {
          if (s->hid[(uint64_t)i * mc + c] == pid) {
            const REAL* xo = &s->hxi[((uint64_t)i * mc + c) * 3];
            xox = xo[0]; xoy = xo[1]; xoz = xo[2];
            break;
          }
        }
        REAL ft[3], xn[3];
        NAME(tangential)(s, fn, nx, ny, nz, vtx, vty, vtz, xox, xoy, xoz, ft, xn);
        Fx = Fx + (fn * nx + ft[0]);
        Fy = Fy + (fn * ny + ft[1]);
        Fz = Fz + (fn * nz + ft[2]);
        REAL tx = ny * ft[2] - nz * ft[1], ty = nz * ft[0] - nx * ft[2], tz = nx * ft[1] - ny * ft[0];
        Tx = Tx - ri * tx; Ty = Ty - ri * ty; Tz = Tz - ri * tz;
        if (nc < mc) {
          s->hid2[(uint64_t)i * mc + nc] = pid;
          REAL* dst = &s->hxi2[((uint64_t)i * mc + nc) * 3];
          dst[0] = xn[0]; dst[1] = xn[1]; dst[2] = xn[2];
        } else {
          overflow = 1;
        }
        ++nc;
      }
    }
    /* walls: bit 2a = lower wall of axis a (normal +e_a), bit 2a+1 = upper (normal -e_a) */
    for (int w = 0; w < 6; ++w) {
      if (!(P->wall_mask & (1u << w))) continue;
      int a = w >> 1;
      REAL pa = a == 0 ? xi : (a == 1 ? yi : zi);
      REAL dist = (w & 1) ? (P->wall_hi[a] - pa) : (pa - P->wall_lo[a]);
      REAL delta = ri - dist;
      if (!(RE(delta) > 0.0)) continue;
      double sg = (w & 1) ? -1.0 : 1.0;
      REAL nx = a == 0 ? sg : 0.0, ny = a == 1 ? sg : 0.0, nz = a == 2 ? sg : 0.0;
      REAL wsx = ri * wxi, wsy = ri * wyi, wsz = ri * wzi;
      REAL cx = wsy * nz - wsz * ny, cy = wsz * nx - wsx * nz, cz = wsx * ny - wsy * nx;
      REAL vrx = vxi - cx, vry = vyi - cy, vrz = vzi - cz;
      REAL vn = vrx * nx + vry * ny + vrz * nz;
      REAL vtx = vrx - vn * nx, vty = vry - vn * ny, vtz = vrz - vn * nz;
      REAL cn = s->two_zeta * SQRT(mi * s->kn);
      REAL fn = s->kn * delta - cn * vn;
      const REAL* xo = &s->wxi[(uint64_t)i * 18 + (uint64_t)w * 3];
      REAL ft[3], xn[3];
      NAME(tangential)(s, fn, nx, ny, nz, vtx, vty, vtz, xo[0], xo[1], xo[2], ft, xn);
      Fx = Fx + (fn * nx + ft[0]);
      Fy = Fy + (fn * ny + ft[1]);
      Fz = Fz + (fn * nz + ft[2]);
      REAL tx = ny * ft[2] - nz * ft[1], ty = nz * ft[0] - nx * ft[2], tz = nx * ft[1] - ny * ft[0];
      Tx = Tx - ri * tx; Ty = Ty - ri * ty; Tz = Tz - ri * tz;
      REAL* dst = &s->wxi2[(uint64_t)i * 18 + (uint64_t)w * 3];
      dst[0] = xn[0]; dst[1] = xn[1]; dst[2] = xn[2];
    }
    s->hcnt2[i] = nc < mc ? nc : mc;
    s->fx[i] = Fx; s->fy[i] = Fy; s->fz[i] = Fz;
    s->tqx[i] = Tx; s->tqy[i] = Ty; s->tqz[i] = Tz;
  }
  { uint32_t* t = s->hcnt; s->hcnt = s->hcnt2; s->hcnt2 = t; }
  { uint32_t* t = s->hid; s->hid = s->hid2; s->hid2 = t; }
  { REAL* t = s->hxi; s->hxi = s->hxi2; s->hxi2 = t; }
  { REAL* t = s->wxi; s->wxi = s->wxi2; s->wxi2 = t; }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (s->id[i] >= P->n_moving) continue;
    REAL mi = s->mass[i], ri = s->rad[i];
    REAL axx = s->fx[i] / mi + P->g[0];
    REAL ayy = s->fy[i] / mi + P->g[1];
    REAL azz = s->fz[i] / mi + P->g[2];
    s->ax[i] = axx; s->ay[i] = ayy; s->az[i] = azz;
    s->vx[i] = s->vx[i] + dt * axx;
    s->vy[i] = s->vy[i] + dt * ayy;
    s->vz[i] = s->vz[i] + dt * azz;
    s->x[i] = s->x[i] + dt * s->vx[i];
    s->y[i] = s->y[i] + dt * s->vy[i];
    s->z[i] = s->z[i] + dt * s->vz[i];
    REAL inertia = 0.4 * mi * ri * ri;
    s->wx[i] = s->wx[i] + dt * (s->tqx[i] / inertia);
    s->wy[i] = s->wy[i] + dt * (s->tqy[i] / inertia);
    s->wz[i] = s->wz[i] + dt * (s->tqz[i] / inertia);
  }
  return overflow ? ORC_E_CAPACITY : ORC_OK;
}

static int NAME(step)(STATE* s, uint32_t nsteps) {
  for (uint32_t it = 0; it < nsteps; ++it) {
    int st = NAME(build)(s);
    if (st) return st;
    if (s->P.model == ORC_MODEL_WCSPH) {
      NAME(density)(s);
      NAME(sph_force_integrate)(s);
    } else if (s->P.model == ORC_MODEL_DEM) {
      st = NAME(dem_force_integrate)(s);
      if (st) return st;
    } else {
      return ORC_E_INVALID;
    }
  }
  return ORC_OK;
}
