/*
 * prestige_hip.h — C ABI of libprestige_hip.so, the MI355X (gfx950) particle
 * hot path: cell-linked-list neighbour search, the per-particle WCSPH / DEM
 * gather loops, the FP64 symplectic-Euler integrator and their adjoint.
 *
 * Reference interfaces this boundary replaces (paths relative to the
 * reference tree):
 *   - particle data: one slice per named field, indexed by particle, `i` the
 *     destination and `j` the source of every pairwise body
 *     (prestige/src/lib.rs:8);
 *   - the equation record {name, reads, writes, body}
 *     (prestige/src/equations/ir.rs:3-14) produced by #[equation]
 *     (prestige_macros/src/lib.rs:8-58);
 *   - fuse(Vec<EquationIR>) -> FusedEquations (prestige/src/equations/fuse.rs:14-40);
 *   - generate_simple_cpu(&FusedEquations) -> String, the i-outer / j-inner
 *     all-pairs loop (prestige/src/codegen/simple_cpu.rs:3-22), which this
 *     library executes instead of printing, with `for j in 0..n` replaced by
 *     the 27-cell neighbour iterator.
 * The reference has no C ABI and no runtime; every entry point below names the
 * reference item whose behaviour it carries.
 *
 * Conventions: every function returns prestige_status (0 = OK) unless stated;
 * no C++ exception crosses the ABI; host pointers are borrowed for the call
 * only; one context is externally synchronised (the `&mut` of the reference),
 * separate contexts may be used from separate threads.  All per-particle host
 * arrays are in ORIGINAL particle-id order (the order of the first upload),
 * whatever order the device currently keeps them in.
 */
#ifndef PRESTIGE_HIP_H
#define PRESTIGE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRESTIGE_ABI_VERSION 1u

typedef enum prestige_status {
  PRESTIGE_OK = 0,
  PRESTIGE_E_INVALID = 1,     /* bad argument, unknown field name, wrong length */
  PRESTIGE_E_HIP = 2,         /* HIP runtime error (message in prestige_last_error) */
  PRESTIGE_E_OOM = 3,         /* device allocation failed */
  PRESTIGE_E_CAPACITY = 4,    /* more than max_contacts DEM partners on one particle, or (WCSPH)
                                 more than 128 particles in one row of three cells */
  PRESTIGE_E_NONFINITE = 5,   /* NaN or Inf found in a state field */
  PRESTIGE_E_DOMAIN = 6,      /* particle outside the cell grid */
  PRESTIGE_E_UNSUPPORTED = 7, /* valid call that this model/precision does not provide */
  PRESTIGE_E_STATE = 8,       /* call out of order (e.g. adjoint before a taped forward) */
  PRESTIGE_E_CFL = 9          /* a particle outran the time step: WCSPH dt (c0 + |v|) > h / 2
                                 (or dt c0 > h / 2 before any step), DEM dt |v| > r / 4 */
} prestige_status;

typedef enum prestige_model {
  PRESTIGE_MODEL_PAIRS = 0,   /* neighbour search + equation probes only */
  PRESTIGE_MODEL_WCSPH = 1,   /* weakly compressible SPH (density, Tait EOS, momentum) */
  PRESTIGE_MODEL_DEM = 2      /* spring-dashpot + Coulomb spheres */
} prestige_model;

typedef enum prestige_precision {
  PRESTIGE_F64 = 0,           /* everything FP64 */
  PRESTIGE_F32 = 1,           /* everything FP32, positions absolute */
  PRESTIGE_MIXED = 2          /* FP32 gathers on FP32 offsets from FP64 cell origins,
                                 Kahan-compensated sums, FP64 integrator */
} prestige_precision;

typedef enum prestige_loop {
  PRESTIGE_LOOP_ALLPAIRS = 0, /* for j in 0..n (prestige/src/codegen/simple_cpu.rs:8) */
  PRESTIGE_LOOP_CELLS = 1     /* j over the 27 cells around i's cell, |x_i-x_j| < cell_size */
} prestige_loop;

/* Particle output formats (prestige_write_particles). */
typedef enum prestige_output_format {
  PRESTIGE_OUT_VTK = 0,   /* legacy VTK POLYDATA, binary (big-endian FP64) */
  PRESTIGE_OUT_CSV = 1    /* text, one row per particle, 17 significant digits */
} prestige_output_format;

/* Built-in equations (prestige/src/lib.rs:7-12 holds the reference's one, eq1). */
typedef enum prestige_equation_id {
  PRESTIGE_EQ1 = 0,           /* force[i] += mass[j]  (prestige/src/lib.rs:10) */
  PRESTIGE_EQ_SPH_DENSITY = 1,
  PRESTIGE_EQ_SPH_EOS = 2,
  PRESTIGE_EQ_SPH_MOMENTUM = 3,
  PRESTIGE_EQ_INTEGRATE = 4,
  PRESTIGE_EQ_DEM_CONTACT = 5,
  PRESTIGE_EQ_COUNT = 6
} prestige_equation_id;

/* Role bits of a field inside one equation body. */
#define PRESTIGE_ROLE_I 1u    /* indexed by the destination particle i */
#define PRESTIGE_ROLE_J 2u    /* indexed by the source particle j */

typedef struct prestige_desc {
  uint32_t abi_version;       /* must be PRESTIGE_ABI_VERSION */
  int32_t model;              /* prestige_model */
  int32_t precision;          /* prestige_precision */
  uint32_t max_contacts;      /* DEM history slots per particle (<= 32) */
  uint64_t n_particles;       /* total particles (< 2^32 - 1; < 2^31 in the MIXED WCSPH form); 0 is
                                 valid: steps and gathers then do nothing and fields are empty */
  uint64_t n_moving;          /* ids [0, n_moving) are integrated; [n_moving, n) stay fixed */
  double grid_lo[3];          /* cell-grid origin; particles must stay in [lo, hi) */
  double grid_hi[3];
  double cell_size;           /* >= interaction cutoff (2h for WCSPH, 2 r_max for DEM) */
  double gravity[3];
  double dt;
  /* WCSPH */
  double h;                   /* smoothing length; support 2h, cubic spline */
  double rho0;
  double c0;
  double alpha;               /* artificial viscosity */
  double mass;                /* particle mass (uniform) */
  /* DEM */
  double kn;                  /* normal stiffness, N/m */
  double kt;                  /* tangential stiffness, N/m */
  double restitution;         /* normal restitution e in (0, 1] */
  double mu;                  /* Coulomb friction */
  double wall_lo[3];          /* half-space walls x>=lo, x<=hi per axis */
  double wall_hi[3];
  uint32_t wall_mask;         /* bit 2a: wall at lo of axis a; bit 2a+1: wall at hi */
  uint32_t flags;             /* PRESTIGE_FLAG_* bits, 0 for none */
} prestige_desc;

/* prestige_desc.flags.  PRESTIGE_FLAG_DENSITY_F64: with PRESTIGE_MIXED WCSPH,
 * the density pass runs the F64 form's pair arithmetic and sums (its rho and p
 * are the F64 form's bit for bit from the same positions) and only the force
 * pass is FP32 with Kahan sums; "rho" and "p" download the FP64 values.  Not
 * available for prestige_forward_tape (E_UNSUPPORTED). */
#define PRESTIGE_FLAG_DENSITY_F64 1u
/* PRESTIGE_FLAG_DEM_EVENTS: DEM contexts keep two event fingerprints per
 * particle over the prestige_step calls since the last position upload:
 * "contact_events" = sum over steps s (1, 2, ... counted from that upload) of
 * s times the particle's contact count (pairs and walls) at step s, and
 * "slip_events" = the same sum over its contacts at the Coulomb cap, both
 * modulo 2^32, downloaded as doubles.  Two runs whose fingerprints agree
 * opened, closed and slipped the same contacts at the same steps (the check
 * a finite-difference gradient of the piecewise-smooth DEM loss needs). */
#define PRESTIGE_FLAG_DEM_EVENTS 2u

/* Opaque context: owns every device buffer and one HIP stream. */
typedef struct prestige_ctx prestige_ctx;

/* Fill *d with defaults for `model` (abi_version, gravity -9.81 z, gamma 7, ...). */
void prestige_desc_init(prestige_desc* d, int32_t model);

/* Create / destroy a context on the current HIP device. */
prestige_status prestige_create(const prestige_desc* desc, prestige_ctx** out);
void prestige_destroy(prestige_ctx* ctx);

/* Copy one SoA field in id order.  Names follow the reference's slice naming
 * (prestige/src/lib.rs:8): "x","y","z" positions; "vx","vy","vz" velocities;
 * "wx","wy","wz" angular velocities (DEM); "radius", "mass" (DEM, per
 * particle); "rho","p" (WCSPH, after a step); "ax","ay","az" accelerations of
 * the last step; "force" (eq1 output, PAIRS/any model); "key" cell key of the
 * current sort (download only, as double).  n must equal n_particles.
 * Uploading a position component of a DEM context clears the tangential
 * contact history.  Download synchronises the context's stream. */
prestige_status prestige_field_upload(prestige_ctx* ctx, const char* name,
                                      const double* data, uint64_t n);
prestige_status prestige_field_download(prestige_ctx* ctx, const char* name,
                                        double* out, uint64_t n);

/* Cell keys of the current positions, stable LSD radix sort, reorder of every
 * field into key order, cell lower-bound table.  The neighbour relation the
 * gathers use afterwards is |x_i - x_j| < cutoff over the 27 surrounding
 * cells. */
prestige_status prestige_neighbors_build(prestige_ctx* ctx);

/* Parity view of the last build: sorted keys (n), perm (n; sorted slot k holds
 * the particle that was in slot perm[k] of the previous order), ids (n; the
 * original id in slot k), cell_lb (n_cells + 1; first sorted slot with key >=
 * c).  Any pointer may be NULL. */
prestige_status prestige_neighbors_download(prestige_ctx* ctx, uint32_t* keys,
                                            uint32_t* perm, uint32_t* ids,
                                            uint32_t* cell_lb);
/* n_cells = nx*ny*nz of the grid. */
prestige_status prestige_grid_dims(prestige_ctx* ctx, uint32_t dims[3]);

/* Run one built-in equation over the current build (prestige/src/codegen/simple_cpu.rs:7-19
 * semantics: i outer, j inner, j == i included, gather into [i]).  Only
 * PRESTIGE_EQ1 is evaluated standalone: force[i] = force[i] + sum_j mass[j],
 * with j over all n (ALLPAIRS) or over neighbours within `cutoff` (CELLS;
 * cutoff <= cell_size).  "force" and "mass" must have been uploaded. */
prestige_status prestige_eval_equation(prestige_ctx* ctx, int32_t eq_id,
                                       int32_t loop, double cutoff);

/* Advance nsteps full steps (neighbour rebuild, gathers, integrator) on the
 * device.  The kernels' error flags (E_NONFINITE, E_CFL, E_CAPACITY, E_DOMAIN,
 * reported in that order of precedence) are read every 16 steps and after
 * the last: a bad step ends the call within 16 steps, the state left is the
 * state after the step that was last read, and prestige_last_error names the
 * stretch of steps. */
prestige_status prestige_step(prestige_ctx* ctx, uint32_t nsteps);

/* Reverse mode.  prestige_forward_tape runs nsteps like prestige_step from
 * the current state and keeps FP64 checkpoints of the state every
 * `checkpoint_every` steps.  prestige_adjoint then takes dL/dx and dL/dv of
 * the final state (3*n each, id order, x-block then y then z; NULL = 0) and
 * returns dL/dparams and dL/dv0 (3*n, id order, NULL to skip), recomputing
 * each checkpoint segment forward with its neighbour lists frozen.
 * params: WCSPH {alpha, c0}; DEM {kn, restitution, mu}. */
prestige_status prestige_forward_tape(prestige_ctx* ctx, uint32_t nsteps,
                                      uint32_t checkpoint_every);
prestige_status prestige_adjoint(prestige_ctx* ctx, const double* seed_x,
                                 const double* seed_v, double* grad_params,
                                 uint32_t n_params, double* grad_v0);

/* Per-kernel HIP-event timing on the context's stream.  enable=1 starts
 * recording (and clears the totals). */
prestige_status prestige_timing_enable(prestige_ctx* ctx, int32_t enable);
/* Number of distinct kernels timed so far. */
uint32_t prestige_timing_count(prestige_ctx* ctx);
/* Kernel idx: name (static storage), summed milliseconds, launches, and the
 * algorithmic bytes of one launch (every array element the kernel must read
 * or write counted once). */
prestige_status prestige_timing_get(prestige_ctx* ctx, uint32_t idx,
                                    const char** name, double* total_ms,
                                    uint64_t* launches, double* bytes_per_launch);

/* Block until all queued work on the context's stream is done. */
prestige_status prestige_sync(prestige_ctx* ctx);

/* Last error message of this context ("" if none; static storage valid until
 * the next call on ctx).  ctx may be NULL for the errors of calls without a
 * context (creation, equation definition and code generation). */
const char* prestige_last_error(const prestige_ctx* ctx);

/* eq1 through the cell iterator on caller-owned device memory, allocating
 * nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same result, bit for bit, as
 * prestige_eval_equation(ctx, PRESTIGE_EQ1, PRESTIGE_LOOP_CELLS, cutoff) on a
 * PRESTIGE_MODEL_PAIRS context with the same grid, positions, force and mass:
 * force[i] += mass[j] over every j with |x_i - x_j| < cutoff (j == i included).
 * x, y, z, mass (read) and force (read and written) are DEVICE pointers to n
 * doubles in particle-id order; grid_lo / grid_hi / cell_size as in
 * prestige_desc; stream is a hipStream_t (NULL: the null stream).  The work
 * runs on that stream and the call returns once it has finished (it reads
 * back the domain flag).  Intermediate arrays live in a fixed device array
 * of about 64.5 MiB per GPU, held until the process ends: n <= 1,048,576
 * particles and at most 4,194,304 cells, else PRESTIGE_E_CAPACITY.  Calls on
 * one GPU share that array, so calls that overlap on one GPU race.
 * PRESTIGE_E_DOMAIN when a particle lies outside the grid.  Errors are read
 * with prestige_last_error(NULL). */
prestige_status prestige_eq1_cells_noalloc(const double* x, const double* y, const double* z,
                                           const double* mass, double* force, uint64_t n,
                                           const double grid_lo[3], const double grid_hi[3],
                                           double cell_size, double cutoff, void* stream);

/* ---- equation interface (prestige/src/equations/ir.rs:3-14, fuse.rs:14-40,
 *      codegen/simple_cpu.rs:3-22) -------------------------------------- */

/* Describe equation eq_id (built-in or defined): name, read and write field names
 * (comma-separated, sorted; every indexed name is a read, LHS included, as in
 * prestige_macros/src/lib.rs:71-133), the i/j role of every read (one
 * PRESTIGE_ROLE_* mask per read, same order) and the Rust body text.  Buffers
 * are NUL-terminated and truncated to their capacity. */
prestige_status prestige_equation_describe(int32_t eq_id, char* name, size_t name_cap,
                                           char* reads, size_t reads_cap,
                                           char* writes, size_t writes_cap,
                                           uint8_t* read_roles, size_t roles_cap,
                                           char* body, size_t body_cap);

/* Fuse equations eq_ids[0..n) in order (fuse.rs:20-31) and split them into
 * gather passes wherever a body reads at [j] a field an earlier body of the
 * same pass writes, or a body with [j] accesses reads a field an earlier
 * i-only body of the pass writes (i-only bodies run once per i after the j
 * loop).  pass_of[k] receives the pass index of eq_ids[k]; *n_passes the
 * count. */
prestige_status prestige_fuse_schedule(const int32_t* eq_ids, uint32_t n,
                                       uint32_t* pass_of, uint32_t* n_passes);

/* Text of the reference's CPU loop for the fused bodies
 * (prestige/src/codegen/simple_cpu.rs:5-21).  *len receives the full length
 * (without NUL); the text is truncated to cap-1 bytes. */
prestige_status prestige_generate_simple_cpu(const int32_t* eq_ids, uint32_t n,
                                             char* out, size_t cap, size_t* len);

/* Register an equation (no counterpart in the reference, where #[equation]
 * runs at compile time; prestige_macros/src/lib.rs:8-58): name is an
 * identifier, body the block text the macro captures (quote!(#body), e.g.
 * "{ force [i] += mass [j] ; }", any spacing).  Reads and writes follow the
 * macro's rule (prestige_macros/src/lib.rs:71-133); each field also gets the
 * i/j role of its accesses (the index inside the brackets).  *eq_id receives
 * the new id (>= PRESTIGE_EQ_COUNT).  The registry is process-wide. */
prestige_status prestige_equation_define(const char* name, const char* body, int32_t* eq_id);

/* HIP source of the fused equations' gather passes: the GPU sibling of
 * prestige_generate_simple_cpu (prestige/src/codegen/simple_cpu.rs:3-22).
 * One extern "C" kernel per pass of prestige_fuse_schedule, in loop form
 * PRESTIGE_LOOP_ALLPAIRS (thread = id i, j over ids 0..n) or
 * PRESTIGE_LOOP_CELLS (thread = slot i of the last build, j over the nine row
 * runs of its cell with |x_i - x_j|^2 < cutoff^2).  Bodies may use fields at
 * [i] and [j], the positions x y z (read only), the descriptor constants h dt
 * rho0 c0 alpha kn kt mu gx gy gz, the helpers kernel_w(dx, dy, dz, h) (cubic
 * spline) and tait(rho, rho0, c0), let bindings, if statements and f64
 * functions in method (a.sqrt(), a.powi(2)) or path (f64::exp(a)) form.
 * Other constructs (tuples, if expressions, writes through [j] or to x y z)
 * return PRESTIGE_E_UNSUPPORTED.  Text as prestige_generate_simple_cpu. */
prestige_status prestige_generate_hip(const int32_t* eq_ids, uint32_t n, int32_t loop,
                                      char* out, size_t cap, size_t* len);

/* Compile the passes of prestige_generate_hip for this device (hiprtc, kept
 * per context) and run them in order on the context's particles.  Fields are
 * FP64 arrays in particle-id order: eq1's force and mass, and named fields
 * (prestige_field_upload / _download with any identifier that is not a field
 * of the context's model).  A field the equations write is created
 * zero-filled when absent; one they only read must have been uploaded
 * (E_STATE).  cutoff is used by PRESTIGE_LOOP_CELLS, 0 < cutoff <= cell_size.
 * i-only bodies run once per i after the j loop. */
prestige_status prestige_run_equations(prestige_ctx* ctx, const int32_t* eq_ids, uint32_t n,
                                       int32_t loop, double cutoff);

/* ---- output (the VTK / CSV particle and time-series files the reference's
 *      .gitignore:4,6 anticipates; no writer exists there) ----------------- */

/* Write the particles to path in particle-id order: the positions (x, y, z)
 * and the comma-separated fields (any name prestige_field_download accepts;
 * NULL or "" for none), with the simulation time.  VTK: points, one vertex
 * cell per particle, the fields as POINT_DATA scalars and the time as field
 * data TIME.  CSV: a "# time" line, a header "id,x,y,z,<fields>" and one row
 * per particle. */
prestige_status prestige_write_particles(prestige_ctx* ctx, const char* path, int32_t format,
                                         const char* fields, double time);

/* Write a ParaView collection file (.pvd) that lists files[k] (paths as
 * written into the file) at times[k], k < n: the time series of particle
 * outputs.  Host only. */
prestige_status prestige_write_series(const char* path, const char* const* files, const double* times,
                                      uint32_t n);

/* ---- boundary geometry from Gmsh meshes (the *.msh files .gitignore:5
 *      anticipates; no reader exists there).  Host only. ----------------- */

/* Boundary particles of the surface mesh in the ASCII MSH 2.x / 4.x file
 * path, at spacing dx: the triangles and quadrangles' coplanar patches are
 * filled with a square dx lattice (points at least dx/2 inside), the feature
 * lines between patches and along open rims are split into equal steps of at
 * most dx, and their corners are kept; a box whose sides are multiples of dx
 * gives exactly the lattice points of its surface.  layers >= 1 surfaces are
 * written, layer l offset by l*dx against (side = -1) or along (side = +1)
 * the face normals (b - a) x (c - a): mesh a tank with normals into the fluid
 * and pass side = -1 for walls behind it.  With x, y and z all NULL only *n
 * is set; otherwise the cap-sized arrays receive the *n points (layer 0
 * first).  Append them as fixed particles (ids >= n_moving). */
prestige_status prestige_mesh_boundary(const char* path, double dx, uint32_t layers, int32_t side,
                                       double* x, double* y, double* z, uint64_t cap, uint64_t* n);

/* ---- scenes (host only; no device needed) -------------------------------
 * Deterministic particle layouts of the benchmark configurations, drawn with
 * SplitMix64.  Arrays are caller-owned, sized from prestige_scene_count. */
typedef enum prestige_scene_id {
  PRESTIGE_SCENE_DEM_LATTICE = 0,   /* 10,000 spheres r=0.01 jittered lattice in a 0.5 m box */
  PRESTIGE_SCENE_DAM_BREAK = 1,     /* fluid cube + 3-layer tank */
  PRESTIGE_SCENE_DEM_POUR = 2       /* polydisperse spheres above a 0.3 m bin */
} prestige_scene_id;

typedef struct prestige_scene_params {
  double dx;                  /* DAM_BREAK lattice spacing */
  uint64_t n_target;          /* particles requested (fluid for DAM_BREAK) */
  uint64_t seed;
} prestige_scene_params;

/* n_total and n_moving for a scene. */
prestige_status prestige_scene_count(int32_t scene, const prestige_scene_params* sp,
                                     uint64_t* n_total, uint64_t* n_moving);
/* Fill a descriptor (grid, physics constants) and the particle arrays
 * (x,y,z,vx,vy,vz each n_total; radius and mass may be NULL for WCSPH). */
prestige_status prestige_scene_build(int32_t scene, const prestige_scene_params* sp,
                                     prestige_desc* desc, double* x, double* y,
                                     double* z, double* vx, double* vy, double* vz,
                                     double* radius, double* mass);

#ifdef __cplusplus
}
#endif
#endif /* PRESTIGE_HIP_H */
