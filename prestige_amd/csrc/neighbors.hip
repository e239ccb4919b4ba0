// neighbors.hip — cell keys, stable LSD radix sort, incremental stable sort
// between steps, reorder into key order, cell lower-bound table, id-order
// scatter/gather, and the eq1 probe.
//
// The sorted layout is what every gather kernel iterates: for a particle in
// cell (cx, cy, cz) the three cells cx-1..cx+1 of each of the nine (dy, dz)
// rows are one contiguous run [lb[row + cx-1], lb[row + cx+2]) of slots.
#include <mutex>

#include "common.hpp"

namespace prestige {

// ---------------------------------------------------------------------------
// cell keys
// ---------------------------------------------------------------------------

// floor((x - lo) * inv_cs) per axis in FP64, exactly as the host oracle does;
// out-of-grid particles raise FLAG_DOMAIN and get key 0 so the sort stays in
// bounds.
__device__ inline uint32_t cell_key_of(double x, double y, double z, const Grid& g,
                                       uint32_t* flags) {
  double fx = floor((x - g.lo[0]) * g.inv_cs);
  double fy = floor((y - g.lo[1]) * g.inv_cs);
  double fz = floor((z - g.lo[2]) * g.inv_cs);
  bool ok = fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)g.dims[0] &&
            fy < (double)g.dims[1] && fz < (double)g.dims[2];
  if (!ok) {
    atomicOr(&flags[FLAG_DOMAIN], 1u);
    return 0u;
  }
  uint32_t cx = (uint32_t)fx, cy = (uint32_t)fy, cz = (uint32_t)fz;
  return (cz * (uint32_t)g.dims[1] + cy) * (uint32_t)g.dims[0] + cx;
}

template <class T4>
__global__ __launch_bounds__(kBlock) void k_cell_keys(uint32_t n, const T4* __restrict__ pos,
                                                      Grid g, uint32_t* __restrict__ keys,
                                                      uint32_t* flags) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T4 p = pos[i];
  keys[i] = cell_key_of((double)p.x, (double)p.y, (double)p.z, g, flags);
}

void keys_from_positions(Ctx& c) {
  c.inc_ready = false;
  if (c.n == 0) { c.keys_valid = true; return; }
  uint32_t n = (uint32_t)c.n;
  LaunchScope ls(&c, "cell_keys", 32.0 * n);
  if (c.desc.precision == PRESTIGE_F32)
    k_cell_keys<float4><<<blocks_for(n), kBlock, 0, c.stream>>>(
        n, c.pos[c.cur].as<float4>(), c.grid, c.keys[c.cur_key].as<uint32_t>(), c.flags.as<uint32_t>());
  else
    k_cell_keys<double4><<<blocks_for(n), kBlock, 0, c.stream>>>(
        n, c.pos[c.cur].as<double4>(), c.grid, c.keys[c.cur_key].as<uint32_t>(), c.flags.as<uint32_t>());
  PR_HIP(hipGetLastError());
  c.keys_valid = true;
}

// ---------------------------------------------------------------------------
// stable LSD radix sort of (key, slot) pairs, 8-bit digits
// ---------------------------------------------------------------------------
constexpr int kRadixBits = 8;
constexpr int kBuckets = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;   // keys per thread and tile
constexpr int kTile = kSortThreads * kSortItems;
static_assert(kTile == (int)kSortTile, "the step kernels count moved slots per sort tile");
static_assert(kSortThreads == kBuckets, "one thread per bucket in the scan phases");

// Lanes of this wavefront whose digit equals mine (and whose item is valid):
// one ballot per digit bit, no shared memory.
__device__ inline uint64_t match_digit(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kRadixBits; ++b) {
    uint64_t bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__device__ inline uint64_t lanes_below() {
  return (1ull << __lane_id()) - 1ull;
}

// Per-tile digit histogram: hist[d * n_tiles + tile].
__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                            uint32_t n, int shift,
                                                            uint32_t n_tiles,
                                                            uint32_t* __restrict__ hist,
                                                            uint32_t* __restrict__ tile_cnt) {
  __shared__ uint32_t h[kBuckets];
  const uint32_t t = threadIdx.x;
  h[t] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * (uint32_t)kTile;
  for (int it = 0; it < kSortItems; ++it) {
    uint32_t idx = base + (uint32_t)it * kSortThreads + t;
    bool valid = idx < n;
    uint32_t d = valid ? (keys[idx] >> shift) & (kBuckets - 1) : 0u;
    uint64_t peers = match_digit(d, valid);
    if (valid && (peers & lanes_below()) == 0ull) atomicAdd(&h[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[t * n_tiles + blockIdx.x] = h[t];
  // the moved-slot counts the last step kernel left for an incremental sort
  // this build does not use start again from zero
  if (t == 0 && tile_cnt) tile_cnt[blockIdx.x] = 0;
}

// Inclusive scan within a wavefront of 64 lanes.
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
  const int lane = (int)__lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(x, (unsigned)o, 64);
    if (lane >= o) x += y;
  }
  return x;
}

// Exclusive scan over a 256-thread block; *total receives the sum.
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* s4, uint32_t* total) {
  const uint32_t w = threadIdx.x >> 6;
  uint32_t x = wave_incl_scan(v);
  if (__lane_id() == 63) s4[w] = x;
  __syncthreads();
  uint32_t base = 0;
  for (uint32_t k = 0; k < w; ++k) base += s4[k];
  *total = s4[0] + s4[1] + s4[2] + s4[3];
  __syncthreads();
  return base + x - v;
}

// One block per digit: exclusive scan of hist[d][*] across tiles, tot[d] = count.
__global__ __launch_bounds__(kSortThreads) void k_sort_scan(uint32_t* __restrict__ hist,
                                                            uint32_t n_tiles,
                                                            uint32_t* __restrict__ tot) {
  __shared__ uint32_t s4[4];
  uint32_t* row = hist + (size_t)blockIdx.x * n_tiles;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < n_tiles; b += kSortThreads) {
    uint32_t idx = b + threadIdx.x;
    uint32_t v = idx < n_tiles ? row[idx] : 0u;
    uint32_t total;
    uint32_t ex = block_excl_scan(v, s4, &total);
    if (idx < n_tiles) row[idx] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = carry;
}

// Stable scatter of one tile.  Items are taken in rounds of 256 consecutive
// slots; inside a round the rank of an item among equal digits is its
// wavefront rank (ballot match) plus the counts of earlier wavefronts, and
// rounds advance a per-digit running offset, so equal digits keep their
// input order.
__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, int shift, uint32_t n_tiles,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot) {
  __shared__ uint32_t s_run[kBuckets];
  __shared__ uint32_t s_wave[4][kBuckets];
  __shared__ uint32_t s4[4];
  const uint32_t t = threadIdx.x;
  const uint32_t w = t >> 6;
  {
    uint32_t total;
    uint32_t digit_base = block_excl_scan(tot[t], s4, &total);
    s_run[t] = digit_base + hist[t * n_tiles + blockIdx.x];
  }
  const uint32_t base = blockIdx.x * (uint32_t)kTile;
  for (int it = 0; it < kSortItems; ++it) {
    uint32_t idx = base + (uint32_t)it * kSortThreads + t;
    bool valid = idx < n;
    uint32_t k = valid ? kin[idx] : 0u;
    uint32_t v = valid ? (vin ? vin[idx] : idx) : 0u;
    uint32_t d = (k >> shift) & (kBuckets - 1);
    s_wave[0][t] = 0; s_wave[1][t] = 0; s_wave[2][t] = 0; s_wave[3][t] = 0;
    __syncthreads();
    uint64_t peers = match_digit(d, valid);
    uint32_t rank = (uint32_t)__popcll(peers & lanes_below());
    if (valid && rank == 0) s_wave[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    {
      uint32_t c0 = s_wave[0][t], c1 = s_wave[1][t], c2 = s_wave[2][t], c3 = s_wave[3][t];
      uint32_t r = s_run[t];
      s_wave[0][t] = r;
      s_wave[1][t] = r + c0;
      s_wave[2][t] = r + c0 + c1;
      s_wave[3][t] = r + c0 + c1 + c2;
      s_run[t] = r + c0 + c1 + c2 + c3;
    }
    __syncthreads();
    if (valid) {
      uint32_t dst = s_wave[w][d] + rank;
      kout[dst] = k;
      vout[dst] = v;
    }
    __syncthreads();
  }
}

__global__ void k_iota(uint32_t n, uint32_t* v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

uint64_t sort_tiles(uint64_t n) { return (n + (uint64_t)kTile - 1) / (uint64_t)kTile; }

static int key_bits(uint32_t n_cells) {
  int b = 0;
  while (b < 32 && ((uint64_t)1 << b) < (uint64_t)n_cells) ++b;
  return b;
}

// Buffers of one radix sort: key and value ping-pong pairs, the per-tile digit
// histogram (kBuckets x tiles), the digit totals, and optionally the moved-slot
// tile counts the sort clears for the next incremental sort.
struct SortBufs {
  uint32_t* keys[2];
  uint32_t* vals[2];
  uint32_t* hist;
  uint32_t* tot;
  uint32_t* tile_cnt;
};

// Sorts the n keys (< n_cells) in keys[src]; *key_out and *val_out receive
// the buffers holding the sorted keys and the permutation (sorted slot k <-
// old slot perm[k]).  c (may be null) records the launches when timing is on.
static void radix_sort_bufs(Ctx* c, const SortBufs& b, int src, uint32_t n, uint32_t n_cells, hipStream_t st,
                            int* key_out, int* val_out) {
  const int bits = key_bits(n_cells);
  const int passes = (bits + kRadixBits - 1) / kRadixBits;
  const uint32_t n_tiles = (uint32_t)sort_tiles(n);
  *key_out = src;
  *val_out = 0;
  if (passes == 0) {
    k_iota<<<blocks_for(n), kBlock, 0, st>>>(n, b.vals[0]);
    PR_HIP(hipGetLastError());
    return;
  }
  int vsrc = -1;   // first pass generates slot indices itself
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kRadixBits;
    const int dst = src ^ 1;
    const int vdst = (vsrc < 0) ? 0 : (vsrc ^ 1);
    {
      LaunchScope ls(c, "sort_hist", 4.0 * n);
      k_sort_hist<<<n_tiles, kSortThreads, 0, st>>>(b.keys[src], n, shift, n_tiles, b.hist,
                                                    p == 0 ? b.tile_cnt : nullptr);
      PR_HIP(hipGetLastError());
    }
    {
      LaunchScope ls(c, "sort_scan", 8.0 * kBuckets * n_tiles);
      k_sort_scan<<<kBuckets, kSortThreads, 0, st>>>(b.hist, n_tiles, b.tot);
      PR_HIP(hipGetLastError());
    }
    {
      LaunchScope ls(c, "sort_scatter", (vsrc < 0 ? 12.0 : 16.0) * n);
      k_sort_scatter<<<n_tiles, kSortThreads, 0, st>>>(b.keys[src], vsrc < 0 ? nullptr : b.vals[vsrc], b.keys[dst],
                                                       b.vals[vdst], n, shift, n_tiles, b.hist, b.tot);
      PR_HIP(hipGetLastError());
    }
    src = dst;
    vsrc = vdst;
  }
  *key_out = src;
  *val_out = vsrc;
}

// Sorts keys[cur_key]; afterwards keys[cur_key] holds sorted keys and
// vals[perm_buf] the permutation (sorted slot k <- old slot perm[k]).
static void radix_sort(Ctx& c) {
  const SortBufs b{{c.keys[0].as<uint32_t>(), c.keys[1].as<uint32_t>()},
                   {c.vals[0].as<uint32_t>(), c.vals[1].as<uint32_t>()},
                   c.sort_hist.as<uint32_t>(), c.sort_tot.as<uint32_t>(), c.inc_tiles.as<uint32_t>()};
  radix_sort_bufs(&c, b, c.cur_key, (uint32_t)c.n, c.grid.n_cells, c.stream, &c.cur_key, &c.perm_buf);
}

// ---------------------------------------------------------------------------
// reorder into key order.  PROBE also writes the FP32 probe (offset from the
// FP64 cell origin times off_scale, cell-coordinate residues mod 4 in .w);
// MIXED also writes FP32 copies of the velocity (mass in .w) and, with OMEGA,
// of the angular velocity (radius in .w).
// ---------------------------------------------------------------------------

// Gather-side copies of one particle moved from slot s to slot k (cell key
// of slot k: key): state, id, and for the mixed form the FP32 probe (offset
// from the FP64 origin of the cell, cell code in .w) and velocity copies.
template <class T4>
struct ReorderIO {
  const T4 *pos_in, *vel_in, *om_in;
  const uint32_t* id_in;
  T4 *pos_out, *vel_out, *om_out;
  uint32_t* id_out;
  Grid g;
  float4 *off32, *vel32, *om32;
};

template <class T4, bool OMEGA, bool PROBE, bool MIXED>
__device__ inline void move_slot(const ReorderIO<T4>& io, uint32_t s, uint32_t k, uint32_t key) {
  const T4 p = io.pos_in[s];
  const T4 v = io.vel_in[s];
  io.pos_out[k] = p;
  io.vel_out[k] = v;
  io.id_out[k] = io.id_in[s];
  if constexpr (OMEGA) {
    const T4 w = io.om_in[s];
    io.om_out[k] = w;
    if constexpr (MIXED) io.om32[k] = make_float4((float)w.x, (float)w.y, (float)w.z, (float)p.w);
  }
  if constexpr (PROBE) {
    const Grid& g = io.g;
    const uint32_t cx = key % (uint32_t)g.dims[0];
    const uint32_t rest = key / (uint32_t)g.dims[0];
    const uint32_t cy = rest % (uint32_t)g.dims[1];
    const uint32_t cz = rest / (uint32_t)g.dims[1];
    const float ox = (float)(((double)p.x - (g.lo[0] + (double)cx * g.cs)) * g.off_scale);
    const float oy = (float)(((double)p.y - (g.lo[1] + (double)cy * g.cs)) * g.off_scale);
    const float oz = (float)(((double)p.z - (g.lo[2] + (double)cz * g.cs)) * g.off_scale);
    io.off32[k] = make_float4(ox, oy, oz, __uint_as_float(cell_code(cx, cy, cz)));
  }
  if constexpr (MIXED) io.vel32[k] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
}

template <class T4, bool OMEGA, bool PROBE, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_reorder(uint32_t n, const uint32_t* __restrict__ perm,
                                                    const uint32_t* __restrict__ skeys, ReorderIO<T4> io) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) move_slot<T4, OMEGA, PROBE, MIXED>(io, perm[k], k, skeys[k]);
}

// The copies of one build from buffer src to dst, as the context's model and
// precision need them.
template <class T4, bool OMEGA, bool MIXED>
ReorderIO<T4> reorder_io(Ctx& c, int src, int dst) {
  return ReorderIO<T4>{c.pos[src].as<T4>(), c.vel[src].as<T4>(), OMEGA ? c.omega[src].as<T4>() : nullptr,
                       c.ids[c.cur_ids].as<uint32_t>(), c.pos[dst].as<T4>(), c.vel[dst].as<T4>(),
                       OMEGA ? c.omega[dst].as<T4>() : nullptr, c.ids[c.cur_ids ^ 1].as<uint32_t>(), c.grid,
                       MIXED ? c.off32.as<float4>() : nullptr, MIXED ? c.vel32.as<float4>() : nullptr,
                       (MIXED && OMEGA) ? c.om32.as<float4>() : nullptr};
}

// The copy variant as a type: fn(ReorderForm<...>{}, io) picks the kernel.
template <class T4, bool OMEGA, bool PROBE, bool MIXED>
struct ReorderForm {
  using T = T4;
  static constexpr bool omega = OMEGA, probe = PROBE, mixed = MIXED;
};

// Calls fn(form, io) for the context's form (the mixed forms build the FP32
// probe; DEM moves omega).
template <class Fn>
void with_reorder_form(Ctx& c, int src, int dst, Fn&& fn) {
  const bool dem = c.desc.model == PRESTIGE_MODEL_DEM;
  const bool sph = c.desc.model == PRESTIGE_MODEL_WCSPH;
  const bool mixed = c.desc.precision == PRESTIGE_MIXED;
  const bool f32 = c.desc.precision == PRESTIGE_F32;
  if (dem) {
    if (mixed) fn(ReorderForm<double4, true, true, true>{}, reorder_io<double4, true, true>(c, src, dst));
    else fn(ReorderForm<double4, true, false, false>{}, reorder_io<double4, true, false>(c, src, dst));
  } else if (f32) {
    fn(ReorderForm<float4, false, false, false>{}, reorder_io<float4, false, false>(c, src, dst));
  } else if (sph && mixed) {
    fn(ReorderForm<double4, false, true, true>{}, reorder_io<double4, false, true>(c, src, dst));
  } else {
    fn(ReorderForm<double4, false, false, false>{}, reorder_io<double4, false, false>(c, src, dst));
  }
}

// Bytes one build moves per particle in the reorder.
double reorder_bytes(const Ctx& c) {
  const bool dem = c.desc.model == PRESTIGE_MODEL_DEM;
  const bool mixed = c.desc.precision == PRESTIGE_MIXED;
  const bool f32 = c.desc.precision == PRESTIGE_F32;
  const double elem = f32 ? 12.0 : 24.0;   // three coordinates per vector
  return 4.0 + 2.0 * (2.0 * elem + 4.0) + (dem ? 2.0 * elem + 16.0 : 0.0) + (mixed ? 4.0 + 16.0 + 16.0 : 0.0) +
         (mixed && dem ? 16.0 : 0.0);
}

// ---------------------------------------------------------------------------
// incremental stable sort.  Between two builds only the slots whose particle
// changed cell ("moved", key_new != key_old) break the order: slots are in
// key_old order, so the slots that kept their key are still sorted and keep
// their relative order, and the stable sort of key_new merges them with the
// few moved ones sorted by (key_new, slot).  With P(t) = moved slots before
// slot t and M the moved entries sorted by (key, slot):
//   kept slot i (key c):  position i - P(i) + #{M entries < (c, i)}
//   moved entry b = (c, i): position b + t - P(t), t = lb_old[c] if i <
//                           lb_old[c], else lb_old[c + 1]
// (the kept slots before (c, i) are the slots before t that kept their key;
// i never lies inside cell c's old range, whose slots all had key c).  The
// step kernel that computed the keys also wrote the moved-slot mask and the
// per-tile moved counts; three launches (gather, sort of M, placement, which
// also moves the particle state) replace the radix sort's nine and the
// reorder, and the permutation and keys are exactly the stable sort's, as
// the radix sort would give them.
// ---------------------------------------------------------------------------
// Up to kRankMax moved entries are ordered by counting their ranks, one
// block per 64 entries; more take a one-block radix sort (the host then
// prefers the full radix sort for the next builds).
constexpr uint32_t kRankMax = 8192;
constexpr int kIncSortThreads = 1024;

// The gather walks a sort tile with kIncThreads threads, kIncItems slots
// each.
constexpr int kIncThreads = 1024;
constexpr int kIncItems = kTile / kIncThreads;
constexpr uint32_t kIncWaves = kIncThreads / 64;
static_assert(kIncItems * kIncThreads == kTile, "whole tiles");

__device__ inline uint32_t wave_sum(uint32_t x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Block b of n_tiles: sums the per-tile moved counts (from the step kernel
// that wrote the keys) of the tiles before b, writes that to tile_off[b] and
// the count before each mask word to word_off, and lists its moved slots at
// their rank, as (key << 32 | slot) in slot order.  Block 0 sums every tile
// and writes the total to tile_off[n_tiles], *count and the host-mapped word
// *host_count, which the host reads without waiting when it picks the next
// build's sort.  The tile's mask words are read before the counts are summed,
// so one barrier serves both.  The placement clears the tile counts.
__global__ __launch_bounds__(kIncThreads) void k_inc_gather(
    uint32_t n, uint32_t n_tiles, const uint32_t* __restrict__ tile_cnt, const uint32_t* __restrict__ knew,
    const uint64_t* __restrict__ mask, uint32_t* __restrict__ tile_off, uint32_t* __restrict__ word_off,
    uint32_t* __restrict__ count, uint32_t* __restrict__ host_count, uint64_t* __restrict__ mov) {
  __shared__ uint32_t s_red[kIncWaves];
  __shared__ uint32_t s_w[kIncItems * kIncWaves];
  const uint32_t w = threadIdx.x >> 6, lane = __lane_id();
  const uint32_t base = blockIdx.x * (uint32_t)kTile;
  uint64_t word[kIncItems];
#pragma unroll
  for (int it = 0; it < kIncItems; ++it) {
    const uint32_t i0 = base + (uint32_t)it * kIncThreads + 64u * w;   // the wavefront's first slot
    word[it] = i0 < n ? mask[i0 >> 6] : 0ull;
  }
  const uint32_t lim = blockIdx.x == 0 ? n_tiles : blockIdx.x;
  uint32_t part = 0;
  for (uint32_t t = threadIdx.x; t < lim; t += kIncThreads) part += tile_cnt[t];
  part = wave_sum(part);
  if (lane == 0) {
    s_red[w] = part;
#pragma unroll
    for (int it = 0; it < kIncItems; ++it) s_w[it * kIncWaves + w] = (uint32_t)__popcll(word[it]);
  }
  __syncthreads();
  uint32_t sum = 0;
  for (uint32_t k = 0; k < kIncWaves; ++k) sum += s_red[k];
  const uint32_t pre = blockIdx.x == 0 ? 0u : sum;
  if (threadIdx.x == 0) {
    tile_off[blockIdx.x] = pre;
    if (blockIdx.x == 0) {
      tile_off[n_tiles] = sum;
      count[0] = sum;
      *host_count = sum;
      __threadfence_system();
    }
  }
  uint32_t run = pre;
#pragma unroll
  for (int it = 0; it < kIncItems; ++it) {
    uint32_t before = run, total = 0;
    for (uint32_t k = 0; k < kIncWaves; ++k) {
      const uint32_t cnt = s_w[it * kIncWaves + k];
      before += k < w ? cnt : 0u;
      total += cnt;
    }
    const uint32_t i = base + (uint32_t)it * kIncThreads + 64u * w + lane;
    before += (uint32_t)__popcll(word[it] & lanes_below());
    if (i < n && (i & 63u) == 0u) word_off[i >> 6] = before;
    if (i < n && ((word[it] >> lane) & 1ull) != 0ull) mov[before] = ((uint64_t)knew[i] << 32) | i;
    run += total;
  }
}

// Writes the moved entries sorted by (key, slot) to out.  Up to kRankMax
// entries, block b places the 64-entry chunk b of them at its rank
// #{y : (key_y, y) < (key_x, x)}, counted by its 16 wavefronts over
// interleaved 64-entry chunks of y and summed in LDS.  The entries are in
// slot order, so (key_y, y) < (key_x, x) is key_y < key_x + [y < x]: one
// 32-bit compare per pair, the y keys read by scalar loads.  Above kRankMax
// entries block 0 runs a stable LSD radix sort on the key over mov and out
// (entries arrive in slot order, so key order alone keeps (key, slot) order).
__global__ __launch_bounds__(kIncSortThreads) void k_inc_sort(const uint32_t* __restrict__ count, int key_bits,
                                                              uint64_t* __restrict__ mov,
                                                              uint64_t* __restrict__ out) {
  constexpr uint32_t kW = kIncSortThreads / 64;
  const uint32_t m = count[0];
  const uint32_t t = threadIdx.x;
  if (m <= kRankMax) {
    const uint32_t nch = (m + 63u) >> 6;
    const uint32_t xc = blockIdx.x;
    if (xc >= nch) return;
    __shared__ uint32_t s_cnt[kW][64];
    const uint32_t lane = __lane_id();
    const uint32_t w = (uint32_t)__builtin_amdgcn_readfirstlane((int)(t >> 6));
    const uint32_t* keys = reinterpret_cast<const uint32_t*>(mov) + 1;   // high halves, stride 2
    const uint32_t x = xc * 64u + lane;
    const uint32_t kx = x < m ? keys[2u * x] : 0u;
    uint32_t cnt = 0;
    for (uint32_t c = w; c < nch; c += kW) {
      const uint32_t y0 = c * 64u;
      if (c != xc && y0 + 64u <= m) {
        const uint32_t thr = c < xc ? kx + 1u : kx;   // all y before x, or all after
#pragma unroll 16
        for (uint32_t y = y0; y < y0 + 64u; ++y) cnt += keys[2u * y] < thr ? 1u : 0u;
      } else {
        const uint32_t y1 = min(y0 + 64u, m);
        for (uint32_t y = y0; y < y1; ++y) cnt += keys[2u * y] < (y < x ? kx + 1u : kx) ? 1u : 0u;
      }
    }
    s_cnt[w][lane] = cnt;
    __syncthreads();
    if (w == 0 && x < m) {
      uint32_t r = 0;
#pragma unroll
      for (uint32_t k = 0; k < kW; ++k) r += s_cnt[k][lane];
      if (r < m) out[r] = mov[x];   // always
    }
    return;
  }
  if (blockIdx.x != 0) return;
  // digit histogram, then a stable scatter in chunks of kIncSortThreads
  // entries (wavefront rank by ballot match + counts of the earlier
  // wavefronts of the chunk)
  __shared__ uint32_t hist[kBuckets];
  __shared__ uint32_t wcnt[kW * kBuckets];
  const uint32_t w = t >> 6;
  uint64_t* src = mov;
  uint64_t* dst = out;
  const int passes = (key_bits + kRadixBits - 1) / kRadixBits;
  for (int p = 0; p < passes; ++p) {
    const int shift = 32 + p * kRadixBits;
    for (uint32_t k = t; k < kBuckets; k += kIncSortThreads) hist[k] = 0;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < m; c0 += kIncSortThreads) {
      const uint32_t k = c0 + t;
      const bool valid = k < m;
      const uint32_t d = valid ? (uint32_t)(src[k] >> shift) & (kBuckets - 1) : 0u;
      const uint64_t peers = match_digit(d, valid);
      if (valid && (peers & lanes_below()) == 0ull) atomicAdd(&hist[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    if (t < 64) {   // exclusive scan of the 256 counts by one wavefront, 4 per lane
      uint32_t v[4], sum = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        v[q] = hist[4 * t + q];
        sum += v[q];
      }
      const uint32_t incl = wave_incl_scan(sum);
      uint32_t run = incl - sum;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        hist[4 * t + q] = run;
        run += v[q];
      }
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < m; c0 += kIncSortThreads) {
      const uint32_t k = c0 + t;
      const bool valid = k < m;
      const uint64_t e = valid ? src[k] : 0ull;
      const uint32_t d = (uint32_t)(e >> shift) & (kBuckets - 1);
      for (uint32_t q = t; q < kW * kBuckets; q += kIncSortThreads) wcnt[q] = 0;
      __syncthreads();
      const uint64_t peers = match_digit(d, valid);
      const uint32_t r = (uint32_t)__popcll(peers & lanes_below());
      if (valid && r == 0) wcnt[w * kBuckets + d] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (t < kBuckets) {   // per digit: offsets of the wavefronts of this chunk
        uint32_t run = hist[t];
        for (uint32_t q = 0; q < kW; ++q) {
          const uint32_t c = wcnt[q * kBuckets + t];
          wcnt[q * kBuckets + t] = run;
          run += c;
        }
        hist[t] = run;
      }
      __syncthreads();
      if (valid) dst[wcnt[w * kBuckets + d] + r] = e;
      __syncthreads();
    }
    uint64_t* x = src;
    src = dst;
    dst = x;
  }
  if (src != out)
    for (uint32_t k = t; k < m; k += kIncSortThreads) out[k] = src[k];
}

// Moved slots before slot t (t <= n; m = all moved slots).
__device__ inline uint32_t inc_moved_before(uint32_t t, uint32_t n, uint32_t m, const uint64_t* __restrict__ mask,
                                            const uint32_t* __restrict__ word_off) {
  if (t >= n) return m;
  const uint32_t r = t & 63u;
  return word_off[t >> 6] + (r ? (uint32_t)__popcll(mask[t >> 6] & ((1ull << r) - 1ull)) : 0u);
}

// Lower bounds in the sorted moved entries mov[0 .. m) of ea (returned in .x)
// and eb >= ea (in .y), one per half-wavefront: each step probes 32 evenly
// spaced entries of the remaining range and keeps the gap that holds the
// bound, so m = 32k entries take three dependent loads.  Every lane of the
// wavefront calls.
__device__ inline uint2 wave_lower_bounds(const uint64_t* __restrict__ mov, uint32_t m, uint64_t ea, uint64_t eb) {
  const uint32_t lane = __lane_id();
  const bool upper = lane >= 32u;
  const uint64_t e = upper ? eb : ea;
  const uint32_t j = lane & 31u;
  uint32_t lo = 0, hi = m;   // the bound lies in [lo, hi]
  while (__ballot(lo < hi)) {
    const uint32_t chunk = (hi - lo + 31u) >> 5;   // 0 when this half is done
    const uint32_t idx = lo + j * chunk;
    const bool below = chunk != 0u && idx < hi && mov[idx] < e;
    const uint64_t bal = __ballot(below);
    const uint32_t c = (uint32_t)__popcll(upper ? (bal >> 32) : (bal & 0xffffffffull));
    if (chunk != 0u) {
      if (c == 0u) {
        hi = lo;   // mov[lo] >= e
      } else {
        const uint32_t nlo = lo + (c - 1u) * chunk + 1u;
        hi = min(lo + c * chunk, hi);
        lo = nlo;
      }
    }
  }
  return make_uint2((uint32_t)__builtin_amdgcn_readlane((int)lo, 0),
                    (uint32_t)__builtin_amdgcn_readlane((int)lo, 32));
}

// One thread per slot (a wavefront covers one mask word, so the moved slots
// before slot i are word_off[i / 64] plus a popcount), then the moved
// entries spread over the blocks.
template <class T4, bool OMEGA, bool PROBE, bool MIXED>
__global__ __launch_bounds__(kBlock) void k_inc_place(
    uint32_t n, uint32_t n_tiles, const uint32_t* __restrict__ knew, const uint64_t* __restrict__ mask,
    const uint32_t* __restrict__ word_off, const uint64_t* __restrict__ mov, const uint32_t* __restrict__ count,
    const uint32_t* __restrict__ lb_old, uint32_t* __restrict__ tile_cnt, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ perm, ReorderIO<T4> io) {
  static_assert(kBlock % 64 == 0, "wavefronts cover whole mask words");
  const uint32_t m = count[0];
  const uint32_t i = blockIdx.x * (uint32_t)kBlock + threadIdx.x;
  const uint32_t lane = __lane_id();
  if (i < n_tiles) tile_cnt[i] = 0;   // for the next step's counts
  const bool valid = i < n;
  const uint64_t word = valid ? mask[i >> 6] : 0ull;
  const bool kept = valid && ((word >> lane) & 1ull) == 0ull;
  const uint64_t km = __ballot(kept);
  if (km != 0ull) {
    const uint32_t key = kept ? knew[i] : 0u;
    const uint64_t e = ((uint64_t)key << 32) | i;
    // kept slots are in (key, slot) order, so the wavefront's bounds lie
    // between those of its first and last kept slot
    const int first = __builtin_ctzll(km), last = 63 - __builtin_clzll(km);
    const uint64_t ea = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)key, first) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)i, first);
    const uint64_t eb = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)key, last) << 32) |
                        (uint32_t)__builtin_amdgcn_readlane((int)i, last);
    const uint2 bnd = wave_lower_bounds(mov, m, ea, eb);
    if (kept) {
      const uint32_t before = word_off[i >> 6] + (uint32_t)__popcll(word & lanes_below());
      uint32_t lo = bnd.x, hi = bnd.y;   // moved entries below (key, i)
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (mov[mid] < e) lo = mid + 1;
        else hi = mid;
      }
      const uint32_t pos = i - before + lo;
      if (pos < n) {   // always, when the moved counts are consistent
        kout[pos] = key;
        perm[pos] = i;
        move_slot<T4, OMEGA, PROBE, MIXED>(io, i, pos, key);
      }
    }
  }
  // moved entries, spread over the blocks
  for (uint32_t b = threadIdx.x * gridDim.x + blockIdx.x; b < m; b += (uint32_t)kBlock * gridDim.x) {
    const uint64_t e = mov[b];
    const uint32_t key = (uint32_t)(e >> 32), s = (uint32_t)e;
    const uint32_t t = s < lb_old[key] ? lb_old[key] : lb_old[key + 1];
    const uint32_t pos = b + t - inc_moved_before(t, n, m, mask, word_off);
    if (pos >= n) continue;   // never, when the moved counts are consistent
    kout[pos] = key;
    perm[pos] = s;
    move_slot<T4, OMEGA, PROBE, MIXED>(io, s, pos, key);
  }
}

// Stable sort of keys[cur_key] from the previous build's order (keys[cur_key
// ^ 1] = its sorted keys, lb its cell table); same outputs as radix_sort.
static void incremental_sort(Ctx& c) {
  const uint32_t n = (uint32_t)c.n;
  const uint32_t n_tiles = (uint32_t)sort_tiles(n);
  const uint32_t* knew = c.keys[c.cur_key].as<uint32_t>();
  uint32_t* kold = c.keys[c.cur_key ^ 1].as<uint32_t>();   // overwritten with the sorted keys
  const uint64_t* mask = c.inc_mask.as<uint64_t>();         // written by the step kernel
  uint32_t* cnt = c.inc_count.as<uint32_t>();
  uint32_t* tile_cnt = c.inc_tiles.as<uint32_t>();      // n_tiles, from the step kernel
  uint32_t* tiles = tile_cnt + n_tiles;                  // tile offsets, n_tiles + 1
  uint32_t* woff = c.inc_woff.as<uint32_t>();            // moved slots before each mask word
  {
    LaunchScope ls(&c, "sort_inc_gather", 2.0 * n_tiles * n_tiles + 8.5 * n);
    k_inc_gather<<<n_tiles, kIncThreads, 0, c.stream>>>(n, n_tiles, tile_cnt, knew, mask, tiles, woff, cnt,
                                                        c.inc_host_dev, c.inc_mov.as<uint64_t>());
    PR_HIP(hipGetLastError());
  }
  {
    LaunchScope ls(&c, "sort_inc_sort", 0.0);
    k_inc_sort<<<kRankMax / 64u, kIncSortThreads, 0, c.stream>>>(cnt, key_bits(c.grid.n_cells),
                                                                 c.inc_mov.as<uint64_t>(),
                                                                 c.inc_tmp.as<uint64_t>());
    PR_HIP(hipGetLastError());
  }
  {
    // placement and the reorder in one pass: each slot's state goes straight
    // to its sorted position
    LaunchScope ls(&c, "sort_inc_place", (12.0 + reorder_bytes(c)) * n);
    const uint64_t* mov = c.inc_tmp.as<uint64_t>();   // sorted
    const uint32_t* lb_old = c.lb.as<uint32_t>();
    uint32_t* perm = c.vals[0].as<uint32_t>();
    with_reorder_form(c, c.cur, c.cur ^ 1, [&](auto form, auto io) {
      using F = decltype(form);
      k_inc_place<typename F::T, F::omega, F::probe, F::mixed><<<blocks_for(n), kBlock, 0, c.stream>>>(
          n, n_tiles, knew, mask, woff, mov, cnt, lb_old, tile_cnt, kold, perm, io);
    });
    PR_HIP(hipGetLastError());
  }
  c.cur_key ^= 1;
  c.perm_buf = 0;
}

// lb[c] = first slot with key >= c, for c = 0 .. n_cells.  Thread k < n owns
// the cells [lo, hi) whose lower bound is k: lo = key[k-1] + 1 (0 for k = 0),
// hi = key[k] + 1; the cells after the last key (lower bound n: the empty
// top of the domain) are spread over kTailBlocks extra blocks.  Every cell
// is written exactly once and no clearing pass is needed.  A lane writes a
// range of up to kOwnCells cells itself; longer ranges (the empty stretches
// of a row beside the walls) are written by the whole wavefront, 64 cells per
// store.
constexpr uint32_t kOwnCells = 4;
constexpr uint32_t kTailBlocks = 64;

__global__ __launch_bounds__(kBlock) void k_cell_lb(uint32_t n, uint32_t n_cells,
                                                    const uint32_t* __restrict__ skeys,
                                                    uint32_t* __restrict__ lb, uint32_t key_blocks) {
  if (blockIdx.x >= key_blocks) {
    const uint32_t first = n == 0 ? 0u : skeys[n - 1] + 1u;
    const uint32_t stride = (gridDim.x - key_blocks) * kBlock;
    for (uint32_t c = first + (blockIdx.x - key_blocks) * kBlock + threadIdx.x; c <= n_cells; c += stride) lb[c] = n;
    return;
  }
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lo = 0, hi = 0;
  if (k < n) {
    lo = k == 0 ? 0u : skeys[k - 1] + 1u;
    hi = skeys[k] + 1u;
  }
  const uint32_t len = hi > lo ? hi - lo : 0u;
  if (len <= kOwnCells)
    for (uint32_t c = lo; c < hi; ++c) lb[c] = k;
  // wave-uniform loop over the lanes with long ranges
  uint64_t m = __ballot(len > kOwnCells);
  const uint32_t lane = __lane_id();
  while (m) {
    const int l = __builtin_ctzll(m);
    m &= m - 1;
    const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)lo, l);
    const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)hi, l);
    const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)k, l);
    for (uint32_t c = a + lane; c < e; c += 64u) lb[c] = v;
  }
}

static void launch_cell_lb(Ctx& c, const uint32_t* skeys) {
  const uint32_t n = (uint32_t)c.n;
  const uint32_t key_blocks = (uint32_t)blocks_for(n);
  k_cell_lb<<<key_blocks + kTailBlocks, kBlock, 0, c.stream>>>(n, c.grid.n_cells, skeys, c.lb.as<uint32_t>(),
                                                               key_blocks);
  PR_HIP(hipGetLastError());
}

void neighbors_build(Ctx& c) {
  const uint32_t n = (uint32_t)c.n;
  if (!c.keys_valid) keys_from_positions(c);
  if (n == 0) {
    launch_cell_lb(c, nullptr);
    c.sorted = true; c.built = true;
    return;
  }
  // incremental sort when the state is still in the previous build's slot
  // order with fresh keys (right after a step) and the last moved count seen
  // was small; every few radix builds the moved count is measured again
  bool inc = false;
  if (c.inc_ready && c.sort_policy != Ctx::kSortRadix) {
    const uint32_t seen = *(volatile const uint32_t*)c.inc_host;
    inc = c.sort_policy == Ctx::kSortIncremental || seen <= kRankMax || ++c.inc_skipped >= 16u;
  }
  if (inc) {
    c.inc_skipped = 0;
    incremental_sort(c);
  } else {
    radix_sort(c);
  }
  c.inc_ready = false;
  const uint32_t* skeys = c.keys[c.cur_key].as<uint32_t>();
  const int src = c.cur, dst = c.cur ^ 1;
  if (!inc) {   // the incremental sort moved the state itself
    const uint32_t* perm = c.vals[c.perm_buf].as<uint32_t>();
    LaunchScope ls(&c, "reorder", reorder_bytes(c) * n);
    with_reorder_form(c, src, dst, [&](auto form, auto io) {
      using F = decltype(form);
      k_reorder<typename F::T, F::omega, F::probe, F::mixed><<<blocks_for(n), kBlock, 0, c.stream>>>(n, perm, skeys, io);
    });
    PR_HIP(hipGetLastError());
  }
  c.cur = dst;
  c.cur_ids ^= 1;
  {
    LaunchScope ls(&c, "cell_lb", 8.0 * n + 4.0 * c.grid.n_cells);
    launch_cell_lb(c, skeys);
  }
  c.sorted = true;
  c.built = true;
}

// ---------------------------------------------------------------------------
// id-order transfer helpers
// ---------------------------------------------------------------------------
// src_kind / dst_kind: 0 = double, 1 = float; comp and stride select one
// component of a packed element (stride 4 for double4/float4).

template <class T>
__global__ void k_scatter_ids(uint32_t n, const T* __restrict__ src, int stride, int comp,
                              const uint32_t* __restrict__ ids, double* __restrict__ out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[ids[k]] = (double)src[(size_t)k * stride + comp];
}

template <class T>
__global__ void k_gather_ids(uint32_t n, const double* __restrict__ in,
                             const uint32_t* __restrict__ ids, T* __restrict__ dst, int stride,
                             int comp) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[(size_t)k * stride + comp] = (T)in[ids[k]];
}

__global__ void k_scatter_keys(uint32_t n, const uint32_t* __restrict__ keys,
                               const uint32_t* __restrict__ ids, double* __restrict__ out) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[ids[k]] = (double)keys[k];
}

void scatter_to_ids(Ctx& c, const void* src, int src_kind, int comp, double* d_out) {
  const uint32_t n = (uint32_t)c.n;
  if (!n) return;
  const uint32_t* ids = c.ids[c.cur_ids].as<uint32_t>();
  int stride = comp >= 0 ? (comp >= 100 ? 2 : 4) : 1;
  int cc = comp >= 100 ? comp - 100 : (comp >= 0 ? comp : 0);
  if (src_kind == 2) {
    k_scatter_keys<<<blocks_for(n), kBlock, 0, c.stream>>>(n, (const uint32_t*)src, ids, d_out);
  } else if (src_kind == 0) {
    k_scatter_ids<double><<<blocks_for(n), kBlock, 0, c.stream>>>(n, (const double*)src, stride, cc, ids, d_out);
  } else {
    k_scatter_ids<float><<<blocks_for(n), kBlock, 0, c.stream>>>(n, (const float*)src, stride, cc, ids, d_out);
  }
  PR_HIP(hipGetLastError());
}

void gather_from_ids(Ctx& c, const double* d_in, void* dst, int dst_kind, int comp) {
  const uint32_t n = (uint32_t)c.n;
  if (!n) return;
  const uint32_t* ids = c.ids[c.cur_ids].as<uint32_t>();
  int stride = comp >= 0 ? (comp >= 100 ? 2 : 4) : 1;
  int cc = comp >= 100 ? comp - 100 : (comp >= 0 ? comp : 0);
  if (dst_kind == 0)
    k_gather_ids<double><<<blocks_for(n), kBlock, 0, c.stream>>>(n, d_in, ids, (double*)dst, stride, cc);
  else
    k_gather_ids<float><<<blocks_for(n), kBlock, 0, c.stream>>>(n, d_in, ids, (float*)dst, stride, cc);
  PR_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// eq1: force[i] += mass[j] (prestige/src/lib.rs:10) inside the loop of
// prestige/src/codegen/simple_cpu.rs:7-19.
// ---------------------------------------------------------------------------

// All pairs: thread t is particle id t, j runs over ids 0..n-1 in id order,
// exactly the generated loop.  force and mass are held in id order.
// The sum is one dependent FP64 add per pair (about 7 cycles each).  The
// masses pass through LDS a block-wide tile at a time, in two alternating
// buffers (one barrier per tile): the next tile is loaded from global memory
// while the current one is summed, and a full tile is read 32 masses ahead of
// the adds, enough to cover the LDS latency at one wavefront per SIMD.
__global__ __launch_bounds__(kBlock) void k_eq1_allpairs(uint32_t n,
                                                         const double* __restrict__ mass,
                                                         double* __restrict__ force) {
  __shared__ double2 tiles[2][kBlock / 2];
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = i < n ? force[i] : 0.0;
  double next = threadIdx.x < n ? mass[threadIdx.x] : 0.0;
  uint32_t buf = 0;
  for (uint32_t j0 = 0; j0 < n; j0 += kBlock, buf ^= 1u) {
    double2* tile = tiles[buf];
    double* tile1 = reinterpret_cast<double*>(tile);
    tile1[threadIdx.x] = next;
    __syncthreads();
    const uint32_t jn = j0 + kBlock + threadIdx.x;
    next = jn < n ? mass[jn] : 0.0;
    uint32_t m = n - j0 < (uint32_t)kBlock ? n - j0 : (uint32_t)kBlock;
    if (m == (uint32_t)kBlock) {
      double2 cur[16], nxt[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) cur[u] = tile[u];
      for (uint32_t t = 16; t < (uint32_t)kBlock / 2; t += 16) {
#pragma unroll
        for (int u = 0; u < 16; ++u) nxt[u] = tile[t + u];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
          acc += cur[u].x;
          acc += cur[u].y;
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) cur[u] = nxt[u];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        acc += cur[u].x;
        acc += cur[u].y;
      }
    } else {
      for (uint32_t t = 0; t < m; ++t) acc += tile1[t];
    }
  }
  if (i < n) force[i] = acc;
}

// Cells: j over the nine row runs of slot i, kept when |x_i - x_j|^2 < cutoff^2.
// force / mass are in id order; slot i updates force[ids[i]] only.
//
// One wavefront per block (64 consecutive slots).  Per row run the wavefront
// copies the union of its lanes' candidate ranges into LDS as FP32 offsets
// from its first slot in units of the cutoff (|q|^2 in .w), with their masses;
// each lane scans its own range with the dot form q.w - 2 s.q < t, whose
// threshold carries the rounding margin of the FP32 form (so the scan keeps a
// superset of the exact neighbours), then walks the set bits in ascending slot
// order and applies the exact FP64 test: the pairs and the order of the plain
// loop, so each force is the plain loop's sum bit for bit.  A union wider than
// the stage is scanned from global memory.
constexpr uint32_t kEq1Stage = 256;

template <class T4>
__global__ __launch_bounds__(64) void k_eq1_cells(uint32_t n, const T4* __restrict__ pos,
                                                  const uint32_t* __restrict__ ids,
                                                  const uint32_t* __restrict__ skeys,
                                                  const uint32_t* __restrict__ lb, Grid g, double c2,
                                                  double inv_cut, float thr, const double* __restrict__ mass,
                                                  double* __restrict__ force) {
#pragma clang fp contract(off)
  __shared__ float4 s_q[kEq1Stage + 4];   // 4 entries of padding for the scan's last trip
  __shared__ double s_m[kEq1Stage];        // mass of the staged slots, gathered once per row run
  const uint32_t lane = threadIdx.x, wbase = blockIdx.x * 64u, i = wbase + lane;
  const bool live = i < n;
  const T4 pi = pos[live ? i : wbase];
  const T4 p0 = pos[wbase];
  const double xr = (double)p0.x, yr = (double)p0.y, zr = (double)p0.z;
  auto offset = [&](const T4& p) {
    return make_float4((float)(((double)p.x - xr) * inv_cut), (float)(((double)p.y - yr) * inv_cut),
                       (float)(((double)p.z - zr) * inv_cut), 0.0f);
  };
  const uint32_t me = live ? ids[i] : 0u;
  double acc = live ? force[me] : 0.0;
  const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1], nz = (uint32_t)g.dims[2];
  const uint32_t key = live ? skeys[i] : 0u;
  const int cx = (int)(key % nx), cy = (int)((key / nx) % ny), cz = (int)(key / (nx * ny));
  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < (int)nx - 1 ? cx + 1 : (int)nx - 1;
  const float4 so = offset(pi);
  const float ax = -2.0f * so.x, ay = -2.0f * so.y, az = -2.0f * so.z;
  const float ss = __builtin_fmaf(so.x, so.x, __builtin_fmaf(so.y, so.y, so.z * so.z));
  const float ss_max = wave_max_pos(live ? ss : 0.0f);
  for (int dz = -1; dz <= 1; ++dz) {
    for (int dy = -1; dy <= 1; ++dy) {
      const int zz = cz + dz, yy = cy + dy;
      uint32_t j0 = 0u, len = 0u;
      if (live && zz >= 0 && zz < (int)nz && yy >= 0 && yy < (int)ny) {
        const uint32_t row = ((uint32_t)zz * ny + (uint32_t)yy) * nx;
        j0 = lb[row + (uint32_t)x0];
        len = lb[row + (uint32_t)x1 + 1] - j0;
      }
      const uint32_t base = wave_min(len ? j0 : 0xFFFFFFFFu);
      const uint32_t top = wave_max(len ? j0 + len : 0u);
      if (top <= base) continue;
      const bool staged = top - base <= kEq1Stage;
      float t = 0.0f;
      if (staged) {
        float mq = 0.0f;
        for (uint32_t kk = lane; kk < top - base; kk += 64u) {
          float4 q = offset(pos[base + kk]);
          q.w = __builtin_fmaf(q.x, q.x, __builtin_fmaf(q.y, q.y, q.z * q.z));
          s_q[kk] = q;
          s_m[kk] = mass[ids[base + kk]];
          mq = q.w > mq ? q.w : mq;
        }
        // rounding of the dot form with |s|, |q| <= M, as in the WCSPH scan
        const float m2 = fmaxf(ss_max, wave_max_pos(mq));
        t = thr + 0x1.0p-19f * (m2 + __builtin_sqrtf(m2) + 1.0f) - ss;
        __syncthreads();
      }
      // up to 128 candidates at a time: four scan words, then a walk of their set bits
      for (uint32_t c0 = 0u; c0 < len; c0 += 128u) {
        const uint32_t clen = len - c0 < 128u ? len - c0 : 128u;
        uint32_t wd[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (uint32_t w = 0u; w < 4u; ++w) {
          if (32u * w >= clen) break;
          const uint32_t cnt = clen - 32u * w < 32u ? clen - 32u * w : 32u;
          uint32_t word = 0u;
          for (uint32_t b = 0u; b < cnt; b += 4u) {
            uint32_t hb = 0u;
#pragma unroll
            for (uint32_t u = 0u; u < 4u; ++u) {
              const uint32_t jc = j0 + c0 + 32u * w + b + u;
              bool hit;
              if (staged) {
                const float4 q = s_q[jc - base];
                hit = __builtin_fmaf(ax, q.x, __builtin_fmaf(ay, q.y, __builtin_fmaf(az, q.z, q.w))) < t;
              } else {
                const float4 q = offset(pos[jc < n ? jc : n - 1u]);
                const float ddx = so.x - q.x, ddy = so.y - q.y, ddz = so.z - q.z;
                hit = __builtin_fmaf(ddx, ddx, __builtin_fmaf(ddy, ddy, ddz * ddz)) < thr;
              }
              hb |= (hit ? 1u : 0u) << u;
            }
            word |= hb << b;
          }
          wd[w] = cnt < 32u ? word & ((1u << cnt) - 1u) : word;
        }
        uint64_t m0 = (uint64_t)wd[0] | ((uint64_t)wd[1] << 32);
        uint64_t m1 = (uint64_t)wd[2] | ((uint64_t)wd[3] << 32);
        auto pop = [&]() -> uint32_t {
          uint32_t b;
          if (m0) {
            b = (uint32_t)__builtin_ctzll(m0);
            m0 &= m0 - 1;
          } else {
            b = 64u + (uint32_t)__builtin_ctzll(m1);
            m1 &= m1 - 1;
          }
          return b;
        };
        // two neighbours a trip, so their loads overlap; the sum still runs in slot order
        while (m0 | m1) {
          const uint32_t ja = j0 + c0 + pop();
          const bool two = (m0 | m1) != 0u;
          const uint32_t jb = two ? j0 + c0 + pop() : ja;
          const T4 pa = pos[ja], pb = pos[jb];
          const double ma = staged ? s_m[ja - base] : mass[ids[ja]];
          const double mb = staged ? s_m[jb - base] : mass[ids[jb]];
          const double ex = (double)pi.x - (double)pa.x, ey = (double)pi.y - (double)pa.y,
                       ez = (double)pi.z - (double)pa.z;
          const double fx = (double)pi.x - (double)pb.x, fy = (double)pi.y - (double)pb.y,
                       fz = (double)pi.z - (double)pb.z;
          if (ex * ex + ey * ey + ez * ez < c2) acc += ma;
          if (two && fx * fx + fy * fy + fz * fz < c2) acc += mb;
        }
      }
      if (staged) __syncthreads();
    }
  }
  if (live) force[me] = acc;
}

void eq1_eval(Ctx& c, int loop, double cutoff) {
  const uint32_t n = (uint32_t)c.n;
  PR_CHECK(c.have_force && c.have_mass, PRESTIGE_E_STATE, "eq1 needs fields 'force' and 'mass' uploaded");
  if (!n) return;
  if (loop == PRESTIGE_LOOP_ALLPAIRS) {
    LaunchScope ls(&c, "eq1_allpairs", 24.0 * n);
    k_eq1_allpairs<<<blocks_for(n), kBlock, 0, c.stream>>>(n, c.mass.as<double>(), c.force.as<double>());
    PR_HIP(hipGetLastError());
    return;
  }
  PR_CHECK(loop == PRESTIGE_LOOP_CELLS, PRESTIGE_E_INVALID, "unknown loop mode");
  PR_CHECK(cutoff > 0.0 && cutoff <= c.grid.cs, PRESTIGE_E_INVALID, "cutoff must be in (0, cell_size]");
  if (!c.sorted) neighbors_build(c);
  LaunchScope ls(&c, "eq1_cells", 56.0 * n);
  const uint32_t* ids = c.ids[c.cur_ids].as<uint32_t>();
  const float thr = cell_scan_threshold(c.grid, cutoff);
  if (c.desc.precision == PRESTIGE_F32)
    k_eq1_cells<float4><<<blocks_for(n, 64), 64, 0, c.stream>>>(
        n, c.pos[c.cur].as<float4>(), ids, c.keys[c.cur_key].as<uint32_t>(), c.lb.as<uint32_t>(),
        c.grid, cutoff * cutoff, 1.0 / cutoff, thr, c.mass.as<double>(), c.force.as<double>());
  else
    k_eq1_cells<double4><<<blocks_for(n, 64), 64, 0, c.stream>>>(
        n, c.pos[c.cur].as<double4>(), ids, c.keys[c.cur_key].as<uint32_t>(), c.lb.as<uint32_t>(),
        c.grid, cutoff * cutoff, 1.0 / cutoff, thr, c.mass.as<double>(), c.force.as<double>());
  PR_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// eq1 over the cell iterator without allocating (prestige_eq1_cells_noalloc)
// ---------------------------------------------------------------------------
// The same kernels as prestige_eval_equation(PRESTIGE_EQ1, PRESTIGE_LOOP_CELLS)
// on a fresh context (cell keys, radix sort, positions in key order, cell
// table, the eq1 gather), on the caller's device arrays, with every
// intermediate array in one file-scope device array: `scratch`.  It holds
// the sort's key and value ping-pong pairs (16 B per particle), the sorted
// positions (32 B), the cell table (4 B per cell + 4), the digit histogram
// and totals, and the error flags, each at a 256-byte boundary, sized for at
// most kNoAllocMaxN particles and kNoAllocMaxCells cells.  Each device holds
// its own copy; host code reaches it only through the address
// hipGetSymbolAddress gives on that device (noalloc_scratch), never through
// the array's name.  Every call on a device reuses that device's copy, so
// calls that overlap on one device race; each call clears what must start
// at zero (the flags) with a kernel.
constexpr uint64_t kNoAllocMaxN = 1ull << 20;
constexpr uint64_t kNoAllocMaxCells = 1ull << 22;

constexpr uint64_t na_round(uint64_t b) { return (b + 255ull) & ~255ull; }
constexpr uint64_t kNaKeys = 0;                                               // keys[2]: 2 x 4 N
constexpr uint64_t kNaVals = kNaKeys + 2 * na_round(4 * kNoAllocMaxN);        // vals[2]: 2 x 4 N
constexpr uint64_t kNaPos = kNaVals + 2 * na_round(4 * kNoAllocMaxN);         // double4 x N
constexpr uint64_t kNaLb = kNaPos + na_round(32 * kNoAllocMaxN);              // cells + 1
constexpr uint64_t kNaHist = kNaLb + na_round(4 * (kNoAllocMaxCells + 1));    // kBuckets x tiles
constexpr uint64_t kNaTot = kNaHist + na_round(4ull * kBuckets * ((kNoAllocMaxN + kTile - 1) / kTile));
constexpr uint64_t kNaFlags = kNaTot + na_round(4ull * kBuckets);
constexpr uint64_t kNoAllocBytes = kNaFlags + na_round(4ull * FLAG_COUNT);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

__global__ void k_noalloc_clear(uint32_t* flags) {
  if (threadIdx.x < (uint32_t)FLAG_COUNT) flags[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_cell_keys_soa(uint32_t n, const double* __restrict__ x,
                                                          const double* __restrict__ y, const double* __restrict__ z,
                                                          Grid g, uint32_t* __restrict__ keys, uint32_t* flags) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = cell_key_of(x[i], y[i], z[i], g, flags);
}

// Positions in key order: spos[k] = x[perm[k]] (the ids of a first build are
// the permutation itself).
__global__ __launch_bounds__(kBlock) void k_sorted_pos_soa(uint32_t n, const double* __restrict__ x,
                                                           const double* __restrict__ y, const double* __restrict__ z,
                                                           const uint32_t* __restrict__ perm,
                                                           double4* __restrict__ spos) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) {
    const uint32_t s = perm[k];
    spos[k] = make_double4(x[s], y[s], z[s], 0.0);
  }
}

// Device address of this device's scratch, looked up once per device with
// that device current.
static void* noalloc_scratch(int device) {
  static std::mutex mu;
  static void* addr[64] = {};
  PR_CHECK(device >= 0 && device < 64, PRESTIGE_E_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void* p = nullptr;
    PR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    PR_CHECK(p != nullptr, PRESTIGE_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    addr[device] = p;
  }
  return addr[device];
}

static void eq1_cells_noalloc(const double* x, const double* y, const double* z, const double* mass, double* force,
                              uint64_t n64, const double lo[3], const double hi[3], double cell_size, double cutoff,
                              hipStream_t st) {
  PR_CHECK(lo != nullptr && hi != nullptr, PRESTIGE_E_INVALID, "null grid bounds");
  const Grid g = make_grid(lo, hi, cell_size);
  PR_CHECK(cutoff > 0.0 && cutoff <= g.cs, PRESTIGE_E_INVALID, "cutoff must be in (0, cell_size]");
  PR_CHECK(n64 <= kNoAllocMaxN, PRESTIGE_E_CAPACITY,
           "n = " + std::to_string(n64) + " exceeds the " + std::to_string(kNoAllocMaxN) +
               " particles the allocation-free scratch holds");
  PR_CHECK(g.n_cells <= kNoAllocMaxCells, PRESTIGE_E_CAPACITY,
           "the grid has " + std::to_string(g.n_cells) + " cells, more than the " + std::to_string(kNoAllocMaxCells) +
               " the allocation-free scratch holds");
  if (n64 == 0) return;
  PR_CHECK(x && y && z && mass && force, PRESTIGE_E_INVALID, "null particle array");
  const uint32_t n = (uint32_t)n64;
  int dev = 0;
  PR_HIP(hipGetDevice(&dev));
  unsigned char* base = static_cast<unsigned char*>(noalloc_scratch(dev));
  uint32_t* flags = reinterpret_cast<uint32_t*>(base + kNaFlags);
  const SortBufs b{{reinterpret_cast<uint32_t*>(base + kNaKeys),
                    reinterpret_cast<uint32_t*>(base + kNaKeys + na_round(4 * kNoAllocMaxN))},
                   {reinterpret_cast<uint32_t*>(base + kNaVals),
                    reinterpret_cast<uint32_t*>(base + kNaVals + na_round(4 * kNoAllocMaxN))},
                   reinterpret_cast<uint32_t*>(base + kNaHist), reinterpret_cast<uint32_t*>(base + kNaTot), nullptr};
  double4* spos = reinterpret_cast<double4*>(base + kNaPos);
  uint32_t* lb = reinterpret_cast<uint32_t*>(base + kNaLb);
  k_noalloc_clear<<<1, 64, 0, st>>>(flags);
  PR_HIP(hipGetLastError());
  k_cell_keys_soa<<<blocks_for(n), kBlock, 0, st>>>(n, x, y, z, g, b.keys[0], flags);
  PR_HIP(hipGetLastError());
  int kb = 0, vb = 0;
  radix_sort_bufs(nullptr, b, 0, n, g.n_cells, st, &kb, &vb);
  const uint32_t* skeys = b.keys[kb];
  const uint32_t* perm = b.vals[vb];
  k_sorted_pos_soa<<<blocks_for(n), kBlock, 0, st>>>(n, x, y, z, perm, spos);
  PR_HIP(hipGetLastError());
  const uint32_t key_blocks = (uint32_t)blocks_for(n);
  k_cell_lb<<<key_blocks + kTailBlocks, kBlock, 0, st>>>(n, g.n_cells, skeys, lb, key_blocks);
  PR_HIP(hipGetLastError());
  k_eq1_cells<double4><<<blocks_for(n, 64), 64, 0, st>>>(n, spos, perm, skeys, lb, g, cutoff * cutoff, 1.0 / cutoff,
                                                          cell_scan_threshold(g, cutoff), mass, force);
  PR_HIP(hipGetLastError());
  uint32_t h[FLAG_COUNT] = {0, 0, 0, 0};
  PR_HIP(hipMemcpyAsync(h, flags, sizeof(h), hipMemcpyDeviceToHost, st));
  PR_HIP(hipStreamSynchronize(st));
  PR_CHECK(h[FLAG_DOMAIN] == 0u, PRESTIGE_E_DOMAIN, "a particle lies outside the cell grid");
}

}  // namespace prestige

extern "C" prestige_status prestige_eq1_cells_noalloc(const double* x, const double* y, const double* z,
                                                      const double* mass, double* force, uint64_t n,
                                                      const double grid_lo[3], const double grid_hi[3],
                                                      double cell_size, double cutoff, void* stream) {
  try {
    prestige::eq1_cells_noalloc(x, y, z, mass, force, n, grid_lo, grid_hi, cell_size, cutoff,
                                static_cast<hipStream_t>(stream));
    prestige::set_context_free_error("");
    return PRESTIGE_OK;
  } catch (const prestige::Error& e) {
    prestige::set_context_free_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    prestige::set_context_free_error(e.what());
    return PRESTIGE_E_INVALID;
  } catch (...) {
    prestige::set_context_free_error("unknown C++ exception");
    return PRESTIGE_E_INVALID;
  }
}
