// common.hpp — internal types of libprestige_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/prestige_hip.h"

namespace prestige {

// Error carrying a prestige_status across internal C++ code; converted to a
// status code at the C ABI.
struct Error : std::runtime_error {
  prestige_status code;
  Error(prestige_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define PR_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::prestige::Error(e_ == hipErrorOutOfMemory ? PRESTIGE_E_OOM : PRESTIGE_E_HIP, \
                              std::string(#call) + ": " + hipGetErrorString(e_));    \
  } while (0)

#define PR_CHECK(cond, code, msg)                                   \
  do {                                                              \
    if (!(cond)) throw ::prestige::Error((code), (msg));            \
  } while (0)

// Device flags written by kernels on error paths only.
// FLAG_CFL: a moving particle outran the step (WCSPH: dt (c0 + |v|) > h / 2;
// DEM: dt |v| > r / 4).
enum FlagSlot : int { FLAG_DOMAIN = 0, FLAG_CAPACITY = 1, FLAG_NONFINITE = 2, FLAG_CFL = 3, FLAG_COUNT = 4 };

// Cell grid: key = (cz*ny + cy)*nx + cx, c = floor((x - lo) * inv_cs).
struct Grid {
  double lo[3];
  double cs;
  double inv_cs;
  double off_scale;   // MIXED offsets are (x - origin) * off_scale (1/h for WCSPH, 1 for DEM)
  int32_t dims[3];
  uint32_t n_cells;
};

// Grid of the box [lo, hi) in cells of size cs: dims = ceil((hi - lo) / cs)
// per axis (every context and the allocation-free entry point build it
// here, so their keys agree bit for bit).
inline Grid make_grid(const double lo[3], const double hi[3], double cs) {
  Grid g{};
  PR_CHECK(cs > 0.0 && cs < 1.0e300, PRESTIGE_E_INVALID, "cell_size must be finite and > 0");
  for (int a = 0; a < 3; ++a) {
    g.lo[a] = lo[a];
    const double span = hi[a] - lo[a];
    PR_CHECK(span > 0.0, PRESTIGE_E_INVALID, "grid_hi must exceed grid_lo on every axis");
    const double cells = __builtin_ceil(span / cs);
    PR_CHECK(cells >= 1.0 && cells < 2.0e9, PRESTIGE_E_INVALID, "grid has too many cells along an axis");
    g.dims[a] = (int32_t)cells;
  }
  g.cs = cs;
  g.inv_cs = 1.0 / cs;
  const double nc = (double)g.dims[0] * (double)g.dims[1] * (double)g.dims[2];
  PR_CHECK(nc < 4.0e9, PRESTIGE_E_INVALID, "grid has more than 4e9 cells");
  g.n_cells = (uint32_t)nc;
  g.off_scale = 1.0;
  return g;
}

// Constants of the WCSPH bodies, derived once on the host from the
// descriptor.  Every device form reads the same FP64 values (FP32 forms
// round them once).
struct SphConst {
  double h, inv_h, H2, eps2, inv_rho0, B, alpha_c0, msig, msh, dt;
  double g[3];
  double cfl_v2;   // largest |v|^2 with dt (c0 + |v|) <= h / 2 (negative: dt alone breaks it)
};

struct DemConst {
  double kn, kt, mu, two_zeta, neg_inv_kt, dt;
  double g[3];
  double wall_lo[3], wall_hi[3];
  uint32_t wall_mask;
  uint32_t max_contacts;
};

// Owned device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    if (b == bytes && p) return;
    release();
    if (b == 0) return;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
      p = nullptr;
      throw Error(PRESTIGE_E_OOM, "hipMalloc(" + std::to_string(b) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = b;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Per-kernel HIP-event timing on the context stream.
struct KernelStat {
  const char* name;
  double total_ms = 0.0;
  uint64_t launches = 0;
  double bytes_per_launch = 0.0;
};

struct Timing {
  bool enabled = false;
  std::vector<KernelStat> stats;
  struct Pending { hipEvent_t a, b; size_t stat; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  hipEvent_t get_event();
  size_t stat_index(const char* name, double bytes);
  void flush();   // collect elapsed times of completed pairs (synchronises on each)
  void clear();
  ~Timing();
};

struct Ctx;

// Records `name` around one launch when timing is on.
// Moved-slot mask of the next incremental sort, written by the step kernels:
// word w has bit l set when slot 64 w + l gets a different cell key from the
// step, and tile_cnt[t] counts the moved slots of sort tile t (kSortTile
// slots; cleared by the sort that consumes it).  Fixed particles never move,
// so the ballot of the live lanes is the whole word: the first live lane of
// the wavefront stores it (live_mask is the wavefront's live lanes, taken
// while all its lanes with a slot below n were active), and a wavefront
// without live lanes stores 0 from its fixed path.
constexpr uint32_t kSortTile = 2048;
struct MovedOut {
  uint64_t* mask;
  uint32_t* tile_cnt;
};
__device__ inline void moved_word_live(const MovedOut& mo, uint32_t i, bool moved, uint64_t live_mask) {
  const uint64_t b = __ballot(moved);
  if (mo.mask && __lane_id() == (uint32_t)__builtin_ctzll(live_mask)) {
    mo.mask[i >> 6] = b;
    if (b) atomicAdd(&mo.tile_cnt[i / kSortTile], (uint32_t)__popcll(b));
  }
}
__device__ inline void moved_word_fixed(const MovedOut& mo, uint32_t i, uint64_t live_mask) {
  if (mo.mask && live_mask == 0ull && __lane_id() == 0u) mo.mask[i >> 6] = 0ull;
}

// Hand a buffer a step just wrote to an adjoint record without copying: the
// next step only writes into the context side of the swap before reading it.
inline void take_buffer(DevBuf& ctx_buf, DevBuf& rec_buf) {
  if (rec_buf.bytes != ctx_buf.bytes) rec_buf.alloc(ctx_buf.bytes);
  std::swap(ctx_buf.p, rec_buf.p);
}

struct LaunchScope {
  Ctx* ctx;
  size_t stat;
  hipEvent_t a = nullptr, b = nullptr;
  LaunchScope(Ctx* c, const char* name, double bytes);
  ~LaunchScope() noexcept(false);
};

// Device copy of the full integration state (used as an adjoint checkpoint).
struct Snapshot {
  DevBuf pos, vel, omega, ids, keys;
  DevBuf hcnt, hid, hxi, wxi;   // DEM contact history
  int cur = 0, cur_ids = 0, cur_key = 0, cur_hist = 0;
  bool sorted = false, keys_valid = false, built = false;
};

// What the reverse sweep needs from one forward step: the sorted pre-step
// pair geometry, the gather outputs, and the frozen neighbour structure.
struct StepRecord {
  DevBuf gpos, gvel, gom, rp, ids, perm, nmeta, nword;
  DevBuf hcnt, hid, hxi, wxi;   // DEM history read by the step
  DevBuf hcnt_out, hid_out;     // DEM contacts the step found (its frozen contact lists)
  DevBuf state_pos, state_vel, state_om;   // DEM FP64 pre-step state (sorted)
};

// Adjoint tape: checkpoints every `every` steps of an nsteps forward run.
struct Tape {
  bool valid = false;
  bool last_recorded = false;   // rec holds the last segment of the taped forward
  uint32_t nsteps = 0, every = 0;
  std::vector<std::unique_ptr<Snapshot>> ckpt;   // state before step s * every
  size_t n_ckpt = 0;                             // checkpoints in use (ckpt keeps its buffers)
  std::unique_ptr<Snapshot> final_state;
  std::vector<std::unique_ptr<StepRecord>> rec;  // reused across segments
  DevBuf xb, vb, xb2, vb2, cvec, xbf, rhob, partial, acc, seeds, wb, wb2;
  DevBuf hb[2], wallb[2], slot_of;   // DEM history adjoints (ping-pong), id -> slot
  // Start a new tape; checkpoint buffers are kept and reused.
  void clear() {
    n_ckpt = 0;
    valid = false;
    last_recorded = false;
  }
  Snapshot& next_ckpt() {
    if (n_ckpt == ckpt.size()) ckpt.emplace_back(new Snapshot());
    return *ckpt[n_ckpt++];
  }
};

struct Ctx {
  Ctx() = default;
  Ctx(const Ctx&) = delete;
  Ctx& operator=(const Ctx&) = delete;
  ~Ctx() {
    if (inc_host) (void)hipHostFree(inc_host);
  }
  prestige_desc desc{};
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t n = 0, n_moving = 0;
  Grid grid{};
  SphConst sph{};
  DemConst dem{};
  std::string err;

  // slot bookkeeping: cur = buffer (0/1) holding the newest positions and
  // velocities; `sorted` = cur is in key order of the last build;
  // `keys_valid` = keys[cur_key] hold the cell keys of buffer cur.
  int cur = 0;
  bool sorted = false;
  bool keys_valid = false;
  bool built = false;
  int cur_key = 0;

  // state, ping-pong pairs (element type depends on precision)
  DevBuf pos[2], vel[2];         // double4 (F64, MIXED) or float4 (F32); DEM: pos.w = radius, vel.w = mass
  DevBuf omega[2];               // DEM: double4 / float4
  DevBuf ids[2];                 // uint32 original id per slot
  int cur_ids = 0;
  DevBuf keys[2], vals[2];       // radix-sort ping-pong
  int perm_buf = 0;              // which vals[] holds the last permutation
  DevBuf sort_hist, sort_tot;    // per-tile digit counts, per-digit totals
  // incremental sort: moved-slot mask (bit per slot), per-tile moved counts
  // and offsets, moved count, moved entries (key << 32 | slot), the same
  // sorted, and the moved slots before each mask word
  DevBuf inc_mask, inc_tiles, inc_count, inc_mov, inc_tmp, inc_woff;
  uint32_t* inc_host = nullptr;  // pinned, mapped: moved count of the last incremental sort
  uint32_t* inc_host_dev = nullptr;  // the same word as the device writes it
  bool inc_ready = false;        // keys[cur_key] are the next keys of the last build's slots
  uint32_t inc_skipped = 0;      // radix builds since the moved count was last measured
  enum SortPolicy { kSortAuto = 0, kSortRadix = 1, kSortIncremental = 2 };
  int sort_policy = kSortAuto;   // PRESTIGE_SORT=radix|incremental overrides the choice
  DevBuf lb;                     // n_cells + 1
  DevBuf tile_q;                 // tile-queue counters of the WCSPH gathers (kQueueWords)
  // WCSPH wall particles far from the fluid (DryWalls): cell stamps (n_cells),
  // last dry pass per particle id, pass counter, the pass the stamps describe,
  // and whether steps must leave every particle usable masks (adjoint records)
  DevBuf movcell, dry_pass, rp_id, pres_id;
  uint32_t dens_pass = 16, stamp_pass = 0;
  bool record_all = false;
  bool dry_reuse = true;         // PRESTIGE_DRY=off turns the reuse off (comparisons and tests)
  void forget_dry() { dens_pass += 2; }   // state changed outside a step
  int n_cus = 0;                 // compute units of the device
  DevBuf flags;                  // FLAG_COUNT uint32
  // gather outputs
  DevBuf rp;                     // (rho, p/rho^2) per slot: double2 or float2
  DevBuf pres;                   // p per slot (double or float)
  DevBuf rp64, pres64;           // MIXED with PRESTIGE_FLAG_DENSITY_F64: the FP64 density outputs
  // PRESTIGE_FLAG_DEM_EVENTS: per particle id, step-weighted sums of contact
  // and slip counts (2 n uint32); steps counted since the last position
  // upload; weight of the step being run (0: not recording)
  DevBuf events;
  uint32_t ev_steps = 0, ev_weight = 0;
  DevBuf acc;                    // a per slot (double4 / float4)
  // MIXED form gather copies
  DevBuf off32, vel32, om32;     // float4: offset from cell origin (cell x in .w); velocity
                                 // (mass in .w); DEM angular velocity (radius in .w)
  // row-run masks of the last density pass: run r of slot i has (first
  // candidate slot, candidate count) at nmeta[r n + i] (uint2) and its hit
  // words at nword[(4 r + w) n + i] (uint32, written for 32 w < count)
  DevBuf nmeta, nword;
  // eq1 probe
  DevBuf force, mass, scratch;
  bool have_force = false, have_mass = false;
  // DEM history (indexed by original id, double-buffered)
  DevBuf hcnt[2], hid[2], hxi[2], wxi[2];
  int cur_hist = 0;
  Timing timing;
  Tape tape;
  bool keep_acc = true;
  // named FP64 fields of generated equations (particle-id order), and the
  // modules compiled for them, keyed by source text
  std::map<std::string, std::unique_ptr<DevBuf>> gfields;
  DevBuf posid;   // positions in id order for the all-pairs form of generated passes
  // read-only model fields bound by generated passes (id-order FP64 copies,
  // refreshed per call), the slot-order copies of [j] fields and the separate
  // outputs of fields a pass reads at [j] and writes, kept between calls
  std::map<std::string, std::unique_ptr<DevBuf>> gmodel, gslots, gsplit;
  struct JitCache {
    std::map<std::string, hipModule_t> mods;
    ~JitCache() {
      for (auto& kv : mods) (void)hipModuleUnload(kv.second);
    }
  } jit;

  // Syncs the stream; throws on a raised flag.  `where` names the stretch of
  // the call the flags cover (e.g. "steps 17-32 of 100").
  void check_flags(const std::string& where = std::string());
};

// Launch helpers -------------------------------------------------------------
constexpr int kBlock = 256;

inline uint32_t blocks_for(uint64_t n, int block = kBlock) {
  return (uint32_t)((n + (uint64_t)block - 1) / (uint64_t)block);
}

// Gather kernels run many blocks over spatially sorted slots; the hardware
// deals workgroup b to XCD b % 8, so consecutive tiles would land on eight
// different L2s.  Launch a multiple of 8 blocks and map physical block b to
// logical tile (b % 8) * per + b / 8: each XCD then walks one contiguous
// range of tiles and its L2 keeps the neighbour cells those tiles share.
inline uint32_t xcd_grid(uint32_t nb) { return ((nb + 7u) / 8u) * 8u; }

__device__ inline uint32_t xcd_tile(uint32_t b, uint32_t grid) {
  uint32_t per = grid / 8u;
  return (b % 8u) * per + b / 8u;
}

// Dynamic form of the same split for the WCSPH gathers, whose tiles cost very
// different amounts (the top and bottom of a tank are wall particles only):
// XCD q owns tiles [q per, (q+1) per) and its wavefronts take them in order
// from a counter; a wavefront whose range is used up takes tiles from the
// other ranges in turn, so no XCD idles while another still has work.  A tile
// is claimed only when the wavefront is free to start it (a claim held during
// the current tile would keep it from idle wavefronts at the end).  The grid
// is at most the resident wavefronts, each looping over tiles; the last
// wavefront to finish clears the counters for the next launch (they start at
// zero when allocated).  Launches sharing the counters must be stream-ordered.
constexpr uint32_t kQueueStride = 32;   // one 128-byte line per counter
constexpr uint32_t kQueueWords = 9 * kQueueStride;
struct TileQueue {
  uint32_t* ctr;      // range heads at ctr[q * kQueueStride], finished wavefronts at ctr[8 * kQueueStride]
  uint32_t n_tiles;   // tiles of one wavefront's slots
  uint32_t per;       // tiles per XCD range
};

// Blocks of `block` threads of `kernel` one compute unit holds at once.
template <class K>
int blocks_per_cu(K kernel, int block) {
  int n = 0;
  PR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, 0));
  return n;
}

// Grid of a tile-queue launch: the wavefronts the device holds at once
// (blocks_per_cu from the occupancy of the kernel), a multiple of 8, never
// more than the tiles need.
inline uint32_t queue_grid(int blocks_per_cu, int n_cus, uint32_t n_tiles) {
  const uint32_t resident = (uint32_t)(blocks_per_cu > 0 ? blocks_per_cu : 1) * (uint32_t)(n_cus > 0 ? n_cus : 1);
  const uint32_t cap = resident >= 8u ? resident / 8u * 8u : 8u;
  const uint32_t need = xcd_grid(n_tiles);
  return need < cap ? need : cap;
}

// Minimum / maximum over the wavefront, returned as a wave-uniform value.
__device__ inline uint32_t wave_min(uint32_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const uint32_t t = __shfl_xor(v, d, 64);
    v = t < v ? t : v;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

__device__ inline uint32_t wave_max(uint32_t v) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const uint32_t t = __shfl_xor(v, d, 64);
    v = t > v ? t : v;
  }
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Maximum of a non-negative float over the wavefront (their bit patterns
// order like the values).
__device__ inline float wave_max_pos(float v) { return __uint_as_float(wave_max(__float_as_uint(v))); }

// Threshold of the FP32 cell scan in units of the cutoff: 1 plus the margin
// that keeps the scan a superset of the exact FP64 test, for offsets up to the
// grid's extent E (in cutoffs) from a wavefront's reference point: 2^-10
// (the test's own rounding) + 16 E 2^-24 (the offsets' rounding).
inline float cell_scan_threshold(const Grid& g, double cutoff) {
  double E = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double lo = g.lo[a], hi = lo + g.dims[a] * g.cs;
    const double e = (std::fabs(lo) > std::fabs(hi) ? std::fabs(lo) : std::fabs(hi)) + (hi - lo);
    E = e > E ? e : E;
  }
  E /= cutoff;
  return (float)(1.0 + 1.0 / 1024.0 + 16.0 * E * 0x1.0p-24);
}

// body(tile) for every tile, each exactly once over the grid; one wavefront
// per block, every lane calls.
template <class Body>
__device__ inline void for_each_tile(const TileQueue& tq, Body&& body) {
  uint32_t q = blockIdx.x & 7u;
  for (uint32_t tries = 0; tries < 8u;) {
    uint32_t t = 0;
    if (__lane_id() == 0u) t = atomicAdd(&tq.ctr[q * kQueueStride], 1u);
    t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
    if (t >= tq.per) {   // this range is used up: try the next one
      q = (q + 1u) & 7u;
      ++tries;
      continue;
    }
    const uint32_t tile = q * tq.per + t;
    if (tile < tq.n_tiles) body(tile);
  }
  if (__lane_id() == 0u && atomicAdd(&tq.ctr[8 * kQueueStride], 1u) == gridDim.x - 1u)
    for (uint32_t k = 0; k <= 8u; ++k) __hip_atomic_store(&tq.ctr[k * kQueueStride], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The FP32 probe's .w carries the particle's cell coordinates modulo 4 (2
// bits per axis: x in bits 6-7, y in 8-9, z in 10-11, bits 0-5 clear).
// Neighbours sit at a cell difference in {-1, 0, 1}, which the residues
// recover exactly: ((c_j - c_i + 1) mod 4) - 1.  With bits 0-5 clear, byte 0
// of w_j - w_i + 64 is ((cx_j - cx_i + 1) mod 4) * 64, which the scan reads
// with a single byte-to-float conversion.
constexpr uint32_t kCodeShift = 6;

__device__ inline uint32_t cell_code(uint32_t cx, uint32_t cy, uint32_t cz) {
  return ((cx & 3u) | ((cy & 3u) << 2) | ((cz & 3u) << 4)) << kCodeShift;
}

__device__ inline int cell_shift(uint32_t wi, uint32_t wj, int axis) {
  const uint32_t s = kCodeShift + 2u * (uint32_t)axis;
  uint32_t t = ((wj >> s) - (wi >> s) + 1u) & 3u;
  return (int)t - 1;
}

// Error message of calls without a context (prestige_last_error(NULL)).
void set_context_free_error(const std::string& msg);

// Per-form kernel entry points (defined in the .hip files).
void neighbors_build(Ctx& c);
uint64_t sort_tiles(uint64_t n);   // radix-sort tiles of n keys (histogram rows per digit)
void keys_from_positions(Ctx& c);
void eq1_eval(Ctx& c, int loop, double cutoff);
void sph_step_f64(Ctx& c);
void sph_step_f32(Ctx& c);
void sph_step_mixed(Ctx& c);
void sph_density_f64_for_mixed(Ctx& c);   // density of PRESTIGE_FLAG_DENSITY_F64 (FP64 translation unit)
void dem_step_f64(Ctx& c);
void dem_step_mixed(Ctx& c);
void run_equations(Ctx& c, const int32_t* eq_ids, uint32_t n, int32_t loop, double cutoff);
bool generic_field_name(const Ctx& c, const std::string& name);
bool model_field_ids(Ctx& c, const std::string& name, double* d_out);
void scatter_to_ids(Ctx& c, const void* src, int src_kind, int comp, double* d_out);
void gather_from_ids(Ctx& c, const double* d_in, void* dst, int dst_kind, int comp);

}  // namespace prestige
