// sph_kernels.hpp — WCSPH gather kernels, templated on the precision form.
//
// The pairwise bodies are the textbook WCSPH set (cubic spline of Monaghan &
// Lattanzio 1985, Tait EOS gamma = 7, Monaghan 1992 artificial viscosity)
// run inside the loop shape of prestige/src/codegen/simple_cpu.rs:7-19: i
// outer (one lane per sorted slot), j inner, gather into [i] only, so no
// atomics.
//
// Two passes per step, split where the fused bodies would read rho[j] and
// p[j] that the density body writes; both run one wavefront per block and
// visit the nine row runs of a particle's cell in lockstep, each run followed
// by the opposite row (sph_run):
//   density  (a) scan the run's candidates in FP32 (offsets from a point of
//            the wavefront with |q|^2, staged in LDS once per wavefront and
//            run), marking every candidate inside the cutoff plus a margin
//            in 32-bit words, which are stored with the run's (first slot,
//            count) for the force pass; (b) walk the set bits in the form's
//            own precision, keep the exact r < 2h pairs and sum the density.
//            Then the EOS.
//   force    stage the gather data of the wavefront's hits of a pair of
//            opposite runs in LDS and walk the stored masks: momentum +
//            viscosity, then symplectic Euler, the next step's cell key and
//            the moved-slot mask of the next incremental sort.
// Both passes take their 64-slot tiles from per-XCD queues (for_each_tile),
// and the density pass keeps the values of wall particles that had no moving
// particle nearby at this pass and the previous one (DryWalls).
// Scanning in FP32 halves the bytes each candidate costs, and the expensive
// pair bodies only run over the set bits.  Runs in the order of sph_run and
// bits in ascending order are the visiting order of the host oracle, so FP64
// sums are formed in exactly its order; the stored masks are also the frozen
// neighbour sets the adjoint reuses.
//
// Forms:
//   kF64   state and pair arithmetic in FP64, positions absolute.  The
//          including translation unit is compiled with -ffp-contract=off
//          and every expression keeps the operation order of the host
//          oracle, so the results are bit-identical to it.
//   kF32   state and pair arithmetic in FP32, positions absolute.
//   kMixed FP64 state (positions, velocities, integrator); pair arithmetic
//          in FP32 on the probe itself: each position is an offset from the
//          FP64 origin of its cell, in units of h, and the pair vector is
//          (o_i - o_j) - (c_j - c_i) * cs/h with c_j - c_i in {-1,0,1}^3, so
//          FP32 never holds an absolute coordinate; with cs = 2h every
//          spline constant is exact in FP32.  Density and force sums use
//          Kahan compensation; the EOS and the final scaling of both sums
//          are evaluated in FP64 from the compensated (sum, carry) pair.
#pragma once
#include "common.hpp"

namespace prestige {

enum Form : int { kF64 = 0, kF32 = 1, kMixed = 2 };

template <int F> struct SphForm;
template <> struct SphForm<kF64> {
  using real = double; using G4 = double4; using R2 = double2; using S4 = double4; using sreal = double;
};
template <> struct SphForm<kF32> {
  using real = float; using G4 = float4; using R2 = float2; using S4 = float4; using sreal = float;
};
template <> struct SphForm<kMixed> {
  using real = float; using G4 = float4; using R2 = float2; using S4 = double4; using sreal = double;
};

// Constants of one form, in the units its pair arithmetic uses.
struct SphUnits {
  double H2;       // cutoff^2: (2h)^2, or 4 in units of h
  double inv_h;    // q = r * inv_h (1 in units of h)
  double eps2;     // 0.01 h^2 (0.01 in units of h)
  double hv;       // viscosity mu = hv * (v.r) / (r^2 + eps2): h, or 1 in units of h
  float cs_u;      // cell size in probe units (cs / h)
  float H2_probe;  // scan cutoff^2 in units of h, with the margin the scan's rounding needs
  double scan_inv_h;  // 1/h: scan offsets are (x - reference) / h
  double unit_scale;  // d(pair units)/d(metres): 1/h for the mixed form, 1 otherwise
};

// Plain or compensated running sum.
template <class T, bool KAHAN> struct Sum {
  T s = T(0), c = T(0);
  __device__ inline void add(T x) {
    if constexpr (KAHAN) {
      T y = x - c;
      T t = s + y;
      c = (t - s) - y;
      s = t;
    } else {
      s = s + x;
    }
  }
  // value as FP64: s - c carries the compensation into the wider type
  __device__ inline double wide() const {
    if constexpr (KAHAN) return (double)s - (double)c;
    else return (double)s;
  }
};

struct CellRows {
  int cx, cy, cz, x0, x1;
};

__device__ inline CellRows decode_cell(uint32_t key, const Grid& g) {
  const uint32_t nx = (uint32_t)g.dims[0], ny = (uint32_t)g.dims[1];
  CellRows r;
  r.cx = (int)(key % nx);
  uint32_t rest = key / nx;
  r.cy = (int)(rest % ny);
  r.cz = (int)(rest / ny);
  r.x0 = r.cx > 0 ? r.cx - 1 : 0;
  r.x1 = r.cx < (int)nx - 1 ? r.cx + 1 : (int)nx - 1;
  return r;
}

// Visit the nine row runs of a cell: fn(j0, j1, dy, dz).
template <class Fn>
__device__ inline void for_each_run(const CellRows& cr, const Grid& g, const uint32_t* __restrict__ lb,
                                    Fn&& fn) {
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  for (int dz = -1; dz <= 1; ++dz) {
    int zz = cr.cz + dz;
    if (zz < 0 || zz >= nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      int yy = cr.cy + dy;
      if (yy < 0 || yy >= ny) continue;
      uint32_t row = ((uint32_t)zz * (uint32_t)ny + (uint32_t)yy) * (uint32_t)nx;
      uint32_t j0 = lb[row + (uint32_t)cr.x0], j1 = lb[row + (uint32_t)cr.x1 + 1];
      fn(j0, j1, dy, dz);
    }
  }
}

// Correctly rounded square root of a positive normal double (pair
// distances 0 < r^2 < 4 h^2; x = 0 would give NaN, and coordinates that
// differ by less than 1e-154 are outside the domain): the hardware
// reciprocal square root refined by the coupled Newton iteration for
// g ~ sqrt(x), hh ~ 1/(2 sqrt(x)) and two remainder corrections
// g += (x - g^2) hh (Markstein, IA-64 and Elementary Functions, 2000, ch. 8).
// sqrt() adds exponent scaling for tiny or huge x and special-value
// selects, which leave such x unchanged; without them the sequence is 7
// instructions shorter and gives the same bits.
__device__ inline double sqrt_pos(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double hh = y * 0.5;
  const double r = __builtin_fma(-hh, g, 0.5);
  g = __builtin_fma(g, r, g);
  hh = __builtin_fma(hh, r, hh);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, hh, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, hh, g);
}
// FP32 forms: the hardware square root (1 ulp) instead of the correctly
// rounded sequence (16 instructions); their pair terms carry ~1e-7 relative
// error from the FP32 geometry anyway.
__device__ inline float sqrt_pos(float x) { return __builtin_amdgcn_sqrtf(x); }

// Correctly rounded a / b for a positive normal b and an a whose quotient
// stays normal (the pair terms: b = r or r^2-scaled density, a bounded):
// reciprocal by two Newton steps from the hardware estimate, then the
// quotient a r corrected once by its remainder a - b q (Markstein 2000,
// ch. 8).  The division operator adds exponent pre-scaling (div_scale,
// div_fmas) and special-value repair (div_fixup), which leave such operands
// unchanged.  One difference: a zero a yields +0 where the operator gives
// -0 for a = -0; the pair sums start at +0 and never become -0, so adding
// either zero leaves them bit-identical.
__device__ inline double div_pos(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  const double rem = __builtin_fma(-b, q, a);
  return __builtin_fma(rem, r, q);
}
// FP32 forms: hardware reciprocal refined by one Newton step, times a (about
// 1.5 ulp; the correctly rounded division costs 10 instructions).
__device__ inline float div_pos(float a, float b) {
  float r = __builtin_amdgcn_rcpf(b);
  r = __builtin_fmaf(__builtin_fmaf(-b, r, 1.0f), r, r);
  return a * r;
}

// Mixed-form pair vector x_i - x_j from two probes (units of h).
__device__ inline void probe_pair(const float4& qi, const float4& qj, float cs_u, float& dx, float& dy,
                                  float& dz) {
  uint32_t wi = __float_as_uint(qi.w), wj = __float_as_uint(qj.w);
  dx = (qi.x - qj.x) - (float)cell_shift(wi, wj, 0) * cs_u;
  dy = (qi.y - qj.y) - (float)cell_shift(wi, wj, 1) * cs_u;
  dz = (qi.z - qj.z) - (float)cell_shift(wi, wj, 2) * cs_u;
}

// Mixed-form pair vector in the walks over stored row masks: (o_i - o_j) -
// (c_j - c_i) cs/h with the y and z cell shifts of a run group's two runs
// folded into per-run bases (by, bz; set by begin at each group) and only the
// x shift read from the cell codes.  The forward force pass and the reverse
// passes form it the same way.
struct ProbeRunBases {
  float by[2] = {0.0f, 0.0f}, bz[2] = {0.0f, 0.0f};
  __device__ void begin(const float4& pi, uint32_t grp, float cs_u);
  __device__ void pair(const float4& pi, const float4& pj, uint32_t t, float cs_u, float& dx, float& dy,
                       float& dz) const {
    dx = (pi.x - pj.x) - (float)cell_shift(__float_as_uint(pi.w), __float_as_uint(pj.w), 0) * cs_u;
    dy = (t ? by[1] : by[0]) - pj.y;
    dz = (t ? bz[1] : bz[0]) - pj.z;
  }
};

template <int F>
__device__ inline void form_pair(const typename SphForm<F>::G4& pi, const typename SphForm<F>::G4& pj,
                                 float cs_u, typename SphForm<F>::real& dx, typename SphForm<F>::real& dy,
                                 typename SphForm<F>::real& dz) {
  if constexpr (F == kMixed) {
    probe_pair(pi, pj, cs_u, dx, dy, dz);
  } else {
    dx = pi.x - pj.x;
    dy = pi.y - pj.y;
    dz = pi.z - pj.z;
  }
}

// Fused multiply-add, one rounding (a correctly rounded a b + c in either
// precision).  The WCSPH pair arithmetic writes its dot products and spline
// polynomials with it, in the same places as the oracle.
__device__ inline double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// a.b of two 3-vectors as fma(ax, bx, fma(ay, by, az bz)).
template <class real>
__device__ inline real dot3(real ax, real ay, real az, real bx, real by, real bz) {
  return fmad(ax, bx, fmad(ay, by, az * bz));
}

// Cubic spline shape f(q) for q < 2 (W = sigma f): 1 + q^2 (0.75 q - 1.5)
// below q = 1, 0.25 (2 - q)^3 above.
template <class real>
__device__ inline real spline_f(real q) {
  if (q < real(1.0)) return fmad(q * q, fmad(real(0.75), q, real(-1.5)), real(1.0));
  real t = real(2.0) - q;
  return real(0.25) * t * t * t;
}

// df/dq for q < 2: q (2.25 q - 3) below q = 1, -0.75 (2 - q)^2 above.
template <class real>
__device__ inline real spline_df(real q) {
  if (q < real(1.0)) return q * fmad(real(2.25), q, real(-3.0));
  real t = real(2.0) - q;
  return real(-0.75) * t * t;
}

// Loads kept in flight per lane.  The gathers are latency-bound (dependent
// loads, little arithmetic per byte), so each lane issues a batch of
// independent loads before it uses any of them: kScanUnr candidates in the
// density scan, kDensUnr neighbours in the density walk, kUnr neighbours in
// the walks of the force pass and the adjoint.
constexpr int kScanUnr = 2;
constexpr int kDensUnr = 2;
constexpr int kUnr = 2;
static_assert(kScanUnr == 1 || kScanUnr == 2 || kScanUnr == 4, "scan batch divides 32");
static_assert(kDensUnr >= 1 && kUnr >= 1, "walk batches hold at least one neighbour");
constexpr uint32_t kPad = 0xFFFFFFFFu;

// Candidate coordinates for the scan: offsets from a reference point of the
// wavefront (the position of its first slot) in units of h, rounded once to
// FP32.  Every neighbour lies within a cell of its particle, so the scan
// needs no per-candidate cell bookkeeping: x_i - x_j is one subtraction.
// The rounding is covered by the probe margin (sph_units), which grows with
// the largest coordinate magnitude the grid allows.
struct ScanPoint {
  float x, y, z;  // this slot's offset from the reference
  float H2p;      // probe cutoff^2 in units of h
};

// Scan cnt <= 32 candidates k0, k0 + 1, ... whose offsets ld(k) returns:
// bit b of the result is set when candidate k0 + b lies inside the probe
// cutoff.  Hits are shifted in from the right (word = 2 word + hit) and the
// word is bit-reversed once at the end.  A batch may read up to
// kScanUnr - 1 candidates past cnt; ld must accept them, and their bits are
// cleared.
template <class Ld>
__device__ inline uint32_t scan_word(Ld&& ld, uint32_t k0, uint32_t cnt, const ScanPoint& s) {
  uint32_t word = 0;
  for (uint32_t b = 0; b < cnt; b += kScanUnr) {
    float4 q[kScanUnr];
#pragma unroll
    for (int u = 0; u < kScanUnr; ++u) q[u] = ld(k0 + b + (uint32_t)u);
#pragma unroll
    for (int u = 0; u < kScanUnr; ++u) {
      const float dx = s.x - q[u].x;
      const float dy = s.y - q[u].y;
      const float dz = s.z - q[u].z;
      const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
      word = word + word + (r2 < s.H2p ? 1u : 0u);
    }
  }
  // word holds P = cnt rounded up to kScanUnr bits, candidate 0 in bit P - 1
  const uint32_t P = (cnt + (uint32_t)kScanUnr - 1u) & ~((uint32_t)kScanUnr - 1u);
  word = __builtin_bitreverse32(word) >> (32u - P);
  return cnt < 32u ? word & ((1u << cnt) - 1u) : word;
}

// word = 2 word + hit in one instruction: the wavefront's hit mask is the
// carry-in of an add with carry, each lane taking its own bit.
__device__ inline uint32_t shift_in(uint32_t w, bool hit) {
  const uint64_t m = __builtin_amdgcn_ballot_w64(hit);
  uint32_t r;
  asm("v_addc_co_u32 %0, vcc, %1, %1, %2" : "=v"(r) : "v"(w), "s"(m) : "vcc");
  return r;
}

// scan_word for LDS-staged candidates that carry |q|^2 in .w: |s - q|^2 <
// H2p is tested as q.w - 2 s.q < H2p - |s|^2, three FMAs per candidate with
// a = -2 s and t = H2p - |s|^2 precomputed.  The rounding of this form grows
// with the offsets' magnitude M, so t carries a margin in M (density kernel).
template <class Ld>
__device__ inline uint32_t scan_word_dot(Ld&& ld, uint32_t k0, uint32_t cnt, float ax, float ay, float az,
                                         float t) {
  uint32_t word = 0;
  for (uint32_t b = 0; b < cnt; b += kScanUnr) {
    float4 q[kScanUnr];
#pragma unroll
    for (int u = 0; u < kScanUnr; ++u) q[u] = ld(k0 + b + (uint32_t)u);
#pragma unroll
    for (int u = 0; u < kScanUnr; ++u) {
      const float v = __builtin_fmaf(ax, q[u].x, __builtin_fmaf(ay, q[u].y, __builtin_fmaf(az, q[u].z, q[u].w)));
      word = shift_in(word, v < t);
    }
  }
  const uint32_t P = (cnt + (uint32_t)kScanUnr - 1u) & ~((uint32_t)kScanUnr - 1u);
  word = __builtin_bitreverse32(word) >> (32u - P);
  return cnt < 32u ? word & ((1u << cnt) - 1u) : word;
}

// Row-run masks of the density pass, kept for the force pass and the
// adjoint.  Run r = 3 (dz + 1) + (dy + 1) of slot i (the order for_each_run
// visits rows): meta[r n + i] = (j0, len) of the run's candidate range (len
// 0 for a row outside the grid), and for 32 w < len word w of its hit mask at
// words[(4 r + w) n + i]; bit b of word w names slot j0 + 32 w + b.  Words at
// or past len are not written.
constexpr uint32_t kRuns = 9;

__device__ inline size_t meta_index(uint32_t r, uint32_t n, uint32_t i) { return (size_t)r * n + i; }

__device__ inline size_t word_index(uint32_t r, uint32_t w, uint32_t n, uint32_t i) {
  return (size_t)(4u * r + w) * n + i;
}

// Scan one row run [j0, j0 + len), len <= 128: bit b of the 128-bit mask
// (m0, m1) is set when candidate j0 + b lies inside the probe cutoff;
// wf(k0, cnt) scans cnt <= 32 candidates from slot k0 into a word.  The run's
// words are also stored as run r of slot i.
template <class Wf>
__device__ inline void scan_run(Wf&& wf, uint32_t j0, uint32_t len, uint32_t r, uint32_t* __restrict__ words,
                                uint32_t n, uint32_t i, uint64_t& m0, uint64_t& m1) {
  uint32_t wd[4] = {0u, 0u, 0u, 0u};
#pragma unroll
  for (uint32_t w = 0; w < 4u; ++w) {
    if (len > 32u * w) {
      const uint32_t cnt = len - 32u * w < 32u ? len - 32u * w : 32u;
      wd[w] = wf(j0 + 32u * w, cnt);
      words[word_index(r, w, n, i)] = wd[w];
    }
  }
  m0 = (uint64_t)wd[0] | ((uint64_t)wd[1] << 32);
  m1 = (uint64_t)wd[2] | ((uint64_t)wd[3] << 32);
}

// Mask of run r of slot i as stored by the density pass.
__device__ inline void load_run(const uint2* __restrict__ meta, const uint32_t* __restrict__ words, uint32_t r,
                                uint32_t n, uint32_t i, uint32_t& j0, uint64_t& m0, uint64_t& m1) {
  const uint2 mt = meta[meta_index(r, n, i)];
  j0 = mt.x;
  uint32_t wd[4];
#pragma unroll
  for (uint32_t w = 0; w < 4u; ++w) wd[w] = mt.y > 32u * w ? words[word_index(r, w, n, i)] : 0u;
  m0 = (uint64_t)wd[0] | ((uint64_t)wd[1] << 32);
  m1 = (uint64_t)wd[2] | ((uint64_t)wd[3] << 32);
}

// Visit the set bits of the W-word mask m in ascending word order, B at a
// time: fn(jj) receives wb[w] + b for bit b of word w, kPad for unused batch
// entries.  The nonzero words are first moved to the front (order kept), and
// every bit is then taken from the front word, the queue moving up one word
// when it empties: a pop costs the same few instructions whichever word it
// comes from.
template <int B, int W, class Fn>
__device__ inline void walk_bits_q(uint64_t (&m)[W], uint32_t (&wb)[W], Fn&& fn) {
#pragma unroll
  for (int p = 0; p < W - 1; ++p) {
#pragma unroll
    for (int k = 0; k < W - 1 - p; ++k) {
      if (m[k] == 0u) {
        m[k] = m[k + 1];
        wb[k] = wb[k + 1];
        m[k + 1] = 0u;
      }
    }
  }
  while (m[0]) {
    uint32_t jj[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      jj[u] = kPad;
      if (m[0]) {
        jj[u] = wb[0] + (uint32_t)__builtin_ctzll(m[0]);
        m[0] &= m[0] - 1u;
        if (m[0] == 0u) {
#pragma unroll
          for (int k = 0; k < W - 1; ++k) {
            m[k] = m[k + 1];
            wb[k] = wb[k + 1];
          }
          m[W - 1] = 0u;
        }
      }
    }
    fn(jj);
  }
}

// Order of the row runs in both WCSPH passes: each run is followed by the
// opposite row (run 8 - r), the centre row last, and a wavefront walks each
// pair as one stream.  Lanes near one corner of their cell have many hits in
// one row of a pair and few in the other, so the pair's per-lane totals are
// far more even than either row's, and lockstep walks idle fewer lanes.
constexpr uint32_t kGroups = 5;  // {0, 8}, {1, 7}, {2, 6}, {3, 5}, {4}
__device__ inline uint32_t sph_run(uint32_t k) {
  // k-th run of the order 0, 8, 1, 7, 2, 6, 3, 5, 4
  return (k & 1u) ? 8u - (k >> 1) : (k >> 1);
}

__device__ inline void ProbeRunBases::begin(const float4& pi, uint32_t grp, float cs_u) {
#pragma unroll
  for (uint32_t t = 0; t < 2u; ++t) {
    const uint32_t r = sph_run(2u * grp + t);
    by[t] = pi.y - (float)((int)(r % 3u) - 1) * cs_u;
    bz[t] = pi.z - (float)((int)(r / 3u) - 1) * cs_u;
  }
}

// Walk of a wavefront over its lanes' stored masks (one wavefront per block;
// every lane calls it, live or not), run groups in the WCSPH order.  For each
// group begin(grp) runs first; then the wavefront copies the gather data of
// the union of its lanes' hit ranges of the group's runs into LDS with
// stage(k, j) (slot j -> stage entry k), then each lane walks its own masks B
// at a time with lds(jj, tt) (jj: stage entries, kPad for unused batch
// entries, jj[0] always valid; tt: which run of the group, 0 or 1, each entry
// belongs to).  A group whose unions do not fit CAP entries together is
// walked run by run, and a run whose union alone exceeds CAP with glob(jj,
// tt) (jj: slots).  Lanes visit their neighbours in group order and scan
// order either way; end_run() follows each group.
template <int B, uint32_t CAP, bool PF = false, class Begin, class Stage, class Lds, class Glob, class End>
__device__ inline void walk_runs_staged(const uint2* __restrict__ meta, const uint32_t* __restrict__ words,
                                        uint32_t n, uint32_t i, bool live, uint32_t lane, Begin&& begin,
                                        Stage&& stage, Lds&& lds, Glob&& glob, End&& end_run) {
  // PF: the next group's row metadata and the first two mask words of each of
  // its runs (candidates 0..63, nearly every run) are loaded one group ahead,
  // masked by the run's length on use, so their latency overlaps the current
  // group's walk (8 registers); words 2 and 3 are read when a run needs them.
  uint2 pm[2] = {make_uint2(0u, 0u), make_uint2(0u, 0u)};
  uint32_t pw[4] = {0u, 0u, 0u, 0u};
  auto prefetch = [&](uint32_t g) {
#pragma unroll
    for (uint32_t t = 0; t < 2u; ++t) {
      const uint32_t k = 2u * g + t;
      pm[t] = make_uint2(0u, 0u);
      if (live && k < kRuns) {
        const uint32_t r = sph_run(k);
        pm[t] = meta[meta_index(r, n, i)];
#pragma unroll
        for (uint32_t w = 0; w < 2u; ++w) pw[2u * t + w] = words[word_index(r, w, n, i)];
      }
    }
  };
  if constexpr (PF) prefetch(0u);
  for (uint32_t grp = 0; grp < kGroups; ++grp) {
    begin(grp);
    const uint32_t nrun = grp + 1u == kGroups ? 1u : 2u;
    uint64_t m[4] = {0u, 0u, 0u, 0u};
    uint32_t j0[2] = {0u, 0u}, base[2] = {0u, 0u}, span[2] = {0u, 0u};
    if constexpr (PF) {
#pragma unroll
      for (uint32_t t = 0; t < 2u; ++t) {
        if (t >= nrun) break;
        j0[t] = pm[t].x;
        uint32_t wd[4];
        wd[0] = pm[t].y > 0u ? pw[2u * t] : 0u;
        wd[1] = pm[t].y > 32u ? pw[2u * t + 1u] : 0u;
        const uint32_t r = sph_run(2u * grp + t);
        wd[2] = pm[t].y > 64u ? words[word_index(r, 2u, n, i)] : 0u;
        wd[3] = pm[t].y > 96u ? words[word_index(r, 3u, n, i)] : 0u;
        m[2u * t] = (uint64_t)wd[0] | ((uint64_t)wd[1] << 32);
        m[2u * t + 1u] = (uint64_t)wd[2] | ((uint64_t)wd[3] << 32);
      }
      if (grp + 1u < kGroups) prefetch(grp + 1u);
    }
#pragma unroll
    for (uint32_t t = 0; t < 2u; ++t) {
      if (t >= nrun) break;
      if constexpr (!PF)
        if (live) load_run(meta, words, sph_run(2u * grp + t), n, i, j0[t], m[2u * t], m[2u * t + 1u]);
      const uint64_t lo = m[2u * t], hi = m[2u * t + 1u];
      uint32_t a = 0xFFFFFFFFu, b = 0u;
      if (lo | hi) {
        a = j0[t] + (lo ? (uint32_t)__builtin_ctzll(lo) : 64u + (uint32_t)__builtin_ctzll(hi));
        b = j0[t] + (hi ? 128u - (uint32_t)__builtin_clzll(hi) : 64u - (uint32_t)__builtin_clzll(lo));
      }
      base[t] = wave_min(a);
      const uint32_t top = wave_max(b);
      span[t] = top > base[t] ? top - base[t] : 0u;
    }
    const uint32_t total = span[0] + span[1];
    if (total == 0u) {
      end_run();
      continue;
    }
    if (total <= CAP) {
      // both runs' unions side by side: run 0 at [0, span0), run 1 after it
      for (uint32_t k = lane; k < total; k += 64u)
        stage(k, k < span[0] ? base[0] + k : base[1] + (k - span[0]));
      __syncthreads();
      // stage entry of bit 0 of each mask word
      uint32_t wb[4] = {j0[0] - base[0], j0[0] - base[0] + 64u, span[0] + j0[1] - base[1],
                        span[0] + j0[1] - base[1] + 64u};
      const uint32_t s0 = span[0];
      walk_bits_q<B>(m, wb, [&](const uint32_t* jj) {
        uint32_t tt[B];
#pragma unroll
        for (int v = 0; v < B; ++v) tt[v] = jj[v] != kPad && jj[v] >= s0 ? 1u : 0u;
        lds(jj, tt);
      });
      __syncthreads();  // the next group's staging overwrites the stage
    } else {
#pragma unroll
      for (uint32_t t = 0; t < 2u; ++t) {
        if (span[t] == 0u) continue;
        uint32_t tt[B];
#pragma unroll
        for (int v = 0; v < B; ++v) tt[v] = t;
        uint64_t mm[2] = {m[2u * t], m[2u * t + 1u]};
        if (span[t] <= CAP) {
          for (uint32_t k = lane; k < span[t]; k += 64u) stage(k, base[t] + k);
          __syncthreads();
          uint32_t wb[2] = {j0[t] - base[t], j0[t] - base[t] + 64u};
          walk_bits_q<B>(mm, wb, [&](const uint32_t* jj) { lds(jj, tt); });
          __syncthreads();
        } else {
          uint32_t wb[2] = {j0[t], j0[t] + 64u};
          walk_bits_q<B>(mm, wb, [&](const uint32_t* jj) { glob(jj, tt); });
        }
      }
    }
    end_run();
  }
}

// ---------------------------------------------------------------------------
// density + EOS: rho_i = m sigma sum_j f(|x_ij|/h) over j with |x_ij| < 2h
// (j == i included), p_i = max(B((rho_i/rho0)^7 - 1), 0), pr_i = p_i / rho_i^2.
// Per row run: probe scan -> run mask (stored for the force pass); per run
// and its opposite row: one walk of the two masks' set bits in the form's
// precision.
//
// One wavefront per block.  For each row run the wavefront first copies the
// union of its lanes' candidate ranges into LDS as FP32 scan offsets (one
// coalesced pass; the lanes of a wavefront share most candidates), and every
// lane scans its own range from there.  A union larger than kStageCap slots
// (a wavefront spread over distant rows) is scanned from global memory.
// ---------------------------------------------------------------------------
// Wall particles far from the fluid.  A fixed particle whose 27 cells held
// no moving particle at this pass and at the previous one has the same
// candidates as then (fixed particles never move, and the stable sort keeps
// their order), so its density, pressure and p/rho^2 are the previous pass's
// bit for bit; a wavefront made only of such particles keeps them instead of
// scanning (FP64 form: rho and p/rho^2 travel in the .w lanes of the state
// and p follows from rho; FP32 and mixed forms: each pass leaves a copy of
// the (rho, p/rho^2) and p of its dry particles by particle id).
// movcell[c] = pass when cell c holds a moving particle (the force pass
// stamps the next pass), dry_pass[id] = the last pass the particle was seen
// dry.  Passes count up; an upload or a restore skips one, so nothing from
// before it is reused.  Its row masks are not rewritten: the forward force
// pass never walks a fixed particle's masks.  Steps recorded for the adjoint
// give it masks with no candidates (zero_masks): the adjoint force pass walks
// every slot, and a dry particle's terms there are exactly zero (no moving
// particle lies within its cells, so its pressure and density reach no
// moving particle's force and every pair it has is between fixed particles).
struct DryWalls {
  const uint32_t* ids;
  uint64_t n_moving;
  const uint32_t* movcell;
  uint32_t* dry_pass;
  void* rp_id;     // FP32 / mixed: (rho, p/rho^2) by id (SphForm::R2)
  void* pres_id;   // FP32 / mixed: p by id (SphForm::real)
  uint32_t pass;
  bool stamps;       // movcell describes this pass
  bool reuse;        // densities may be kept
  bool zero_masks;   // kept particles get empty row masks (recorded steps)
};

constexpr int kDensBlock = 64;  // one wavefront: block barriers are wave barriers
// Wavefronts per SIMD the register budget is held to: the tile loop lets the
// compiler keep more values live, and these are the budgets that measured
// fastest (density: 7 in every form, one fewer than the occupancy before the
// loop, which avoids spills).
template <int F>
constexpr int kDensWaves = 7;
constexpr uint32_t kStageCap = 256;   // scan offsets staged per run (4 KB)

// Tait pressure p = max(B ((rho / rho0)^7 - 1), 0), one operation order for
// every caller.
__device__ inline double tait_p(double rho, const SphConst& kc) {
  const double ratio = rho * kc.inv_rho0;
  const double r2_ = ratio * ratio;
  const double r4 = r2_ * r2_;
  const double r7 = r4 * r2_ * ratio;
  double p = kc.B * (r7 - 1.0);
  if (p < 0.0) p = 0.0;
  return p;
}

// True when none of the 27 cells around cr held a moving particle at this
// pass (cell stamps written by the previous step's force pass).
__device__ inline bool dry_cells(const CellRows& cr, const Grid& g, const DryWalls& dw) {
  const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
  bool dry = true;
  for (int dz = -1; dz <= 1; ++dz) {
    const int zz = cr.cz + dz;
    if (zz < 0 || zz >= nz) continue;
    for (int dy = -1; dy <= 1; ++dy) {
      const int yy = cr.cy + dy;
      if (yy < 0 || yy >= ny) continue;
      const uint32_t row = ((uint32_t)zz * (uint32_t)ny + (uint32_t)yy) * (uint32_t)nx;
      for (int x = cr.x0; x <= cr.x1; ++x) dry &= dw.movcell[row + (uint32_t)x] != dw.pass;
    }
  }
  return dry;
}

template <int F>
__global__ __launch_bounds__(kDensBlock, kDensWaves<F>) void k_sph_density(
    uint32_t n, const typename SphForm<F>::S4* __restrict__ spos, const typename SphForm<F>::G4* pos,
    const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ lb, Grid g, SphConst kc, SphUnits u,
    uint2* __restrict__ meta, uint32_t* __restrict__ words, typename SphForm<F>::R2* __restrict__ rp,
    typename SphForm<F>::G4* wpos, typename SphForm<F>::G4* wvel,
    typename SphForm<F>::real* __restrict__ pres, uint32_t* flags, TileQueue tq, DryWalls dw) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  using S4 = typename SphForm<F>::S4;
  __shared__ float4 s_q[kStageCap + kScanUnr];
  const uint32_t lane = threadIdx.x;
  for_each_tile(tq, [&](uint32_t tile) {
    const uint32_t wbase = tile * kDensBlock;
    const uint32_t i = wbase + lane;
    const bool live = i < n;
    // reference point of the wavefront (n >= 1 whenever the kernel runs)
    const S4 p0 = spos[wbase < n ? wbase : n - 1];
    const double xr = (double)p0.x, yr = (double)p0.y, zr = (double)p0.z, ih = u.scan_inv_h;
    auto offset = [&](const S4& p) {
      return make_float4((float)(((double)p.x - xr) * ih), (float)(((double)p.y - yr) * ih),
                         (float)(((double)p.z - zr) * ih), 0.0f);
    };
    ScanPoint sp{0.0f, 0.0f, 0.0f, u.H2_probe};
    CellRows cr{};
    G4 pi{};
    if (live) {
      cr = decode_cell(skeys[i], g);
      pi = pos[i];
    }
    // Mixed form, wavefront within one row of cells (the common case): scan
    // offsets are taken from the 16-byte probes and the cell keys, in a frame
    // at the origin of cell (cx_ref, row) in units of h, instead of from the
    // 32-byte FP64 state.  Every candidate of run r then lies in the run's
    // row, cells cx_ref + (key - row base), shifted by the run's (dy, dz).
    bool probe_frame = false;
    uint32_t cx_ref = 0, row_ref = 0;
    if constexpr (F == kMixed) {
      const uint32_t rowid = (uint32_t)cr.cz * (uint32_t)g.dims[1] + (uint32_t)cr.cy;
      const uint32_t rmin = wave_min(live ? rowid : 0xFFFFFFFFu);
      probe_frame = rmin == wave_max(live ? rowid : 0u);
      row_ref = rmin;
      cx_ref = wave_min(live ? (uint32_t)cr.cx : 0xFFFFFFFFu);
    }
    auto probe_offset = [&](uint32_t j, uint32_t rowbase, float oy, float oz) {
      const float4 q = reinterpret_cast<const float4*>(pos)[j];
      const int cx = (int)(skeys[j] - rowbase) - (int)cx_ref;
      return make_float4(q.x + (float)cx * u.cs_u, q.y + oy, q.z + oz, 0.0f);
    };
    if (live) {
      if (probe_frame) {
        sp.x = (float)pi.x + (float)(cr.cx - (int)cx_ref) * u.cs_u;
        sp.y = (float)pi.y;
        sp.z = (float)pi.z;
      } else {
        const float4 o = offset(spos[i]);
        sp.x = o.x;
        sp.y = o.y;
        sp.z = o.z;
      }
    }
    // a wavefront of fixed particles whose densities cannot have changed
    // since the last pass keeps them (see DryWalls)
    using R2 = typename SphForm<F>::R2;
    uint64_t dry_mask = 0;   // lanes seen dry at this pass
    if (dw.stamps && __ballot(live && (uint64_t)dw.ids[i] < dw.n_moving) == 0ull) {
      bool keep = true, dry = false;
      uint32_t id = 0;
      if (live) {
        id = dw.ids[i];
        dry = dry_cells(cr, g, dw);
        keep = dry && dw.reuse && dw.dry_pass[id] == dw.pass - 1u;
        if (dry) dw.dry_pass[id] = dw.pass;
      }
      dry_mask = __ballot(dry);
      if (__ballot(!keep) == 0ull) {
        if (live && dw.zero_masks)
          for (uint32_t r = 0; r < kRuns; ++r) meta[meta_index(r, n, i)] = make_uint2(0u, 0u);
        if (live) {
          if constexpr (F == kF64) {
            const double rho = pi.w, pr = reinterpret_cast<const double*>(wvel + i)[3];
            rp[i] = make_double2(rho, pr);
            pres[i] = tait_p(rho, kc);
          } else {
            rp[i] = static_cast<const R2*>(dw.rp_id)[id];
            pres[i] = static_cast<const real*>(dw.pres_id)[id];
          }
        }
        return;
      }
    }
    // dot-form scan of staged runs: a = -2 s, and |s|^2 for the threshold
    const float ax = -2.0f * sp.x, ay = -2.0f * sp.y, az = -2.0f * sp.z;
    const float ss = __builtin_fmaf(sp.x, sp.x, __builtin_fmaf(sp.y, sp.y, sp.z * sp.z));
    const float ss_max = wave_max_pos(ss);
    const real H2 = (real)u.H2, inv_h = (real)u.inv_h;
    Sum<real, F == kMixed> acc;
    bool overflow = false;
    const int nx = g.dims[0], ny = g.dims[1], nz = g.dims[2];
    // Runs in the WCSPH order, scanned one at a time and walked a group at a
    // time: a run and its opposite row ({0, 8}, {1, 7}, {2, 6}, {3, 5}, {4})
    // as one stream of set bits, so a lane with few hits in one row and many
    // in the other idles less of the lockstep walk (at 4M: FP64 1147 -> 1106
    // us a pass, mixed 941 -> 892, FP32 904 -> 844).  The order of the pairs
    // stays the scan order.
    constexpr uint32_t G = 2u;   // runs per group
    for (uint32_t grp = 0; grp < kGroups; ++grp) {
      const uint32_t nrun = grp + 1u < kGroups ? 2u : 1u;
      uint64_t m[2 * G];
      uint32_t gj0[G];
      float gby[G], gbz[G];   // mixed form: per-run bases of the pair vector
#pragma unroll
      for (uint32_t t = 0; t < G; ++t) {
        m[2u * t] = m[2u * t + 1u] = 0u;
        gj0[t] = 0u;
        gby[t] = gbz[t] = 0.0f;
      }
#pragma unroll
      for (uint32_t t = 0; t < G; ++t) {
        if (t >= nrun) break;
        const uint32_t r = sph_run(G * grp + t);
        const int zz = cr.cz + (int)(r / 3u) - 1, yy = cr.cy + (int)(r % 3u) - 1;
        const bool row_ok = live && zz >= 0 && zz < nz && yy >= 0 && yy < ny;
        uint32_t j0 = 0, len = 0;
        if (row_ok) {
          const uint32_t row = ((uint32_t)zz * (uint32_t)ny + (uint32_t)yy) * (uint32_t)nx;
          j0 = lb[row + (uint32_t)cr.x0];
          len = lb[row + (uint32_t)cr.x1 + 1] - j0;
          if (len > 128u) {
            overflow = true;
            len = 128u;
          }
        }
        gj0[t] = j0;
        // probe pair vector (o_i - o_j) - (c_j - c_i) cs/h: the y and z shifts
        // are the run's (folded into bases), x comes from the cell codes
        if constexpr (F == kMixed) {
          gby[t] = (float)pi.y - (float)((int)(r % 3u) - 1) * u.cs_u;
          gbz[t] = (float)pi.z - (float)((int)(r / 3u) - 1) * u.cs_u;
        }
        if (live) meta[meta_index(r, n, i)] = make_uint2(j0, len);
        // union of the wavefront's candidate ranges of this run
        const uint32_t base = wave_min(len ? j0 : 0xFFFFFFFFu);
        const uint32_t top = wave_max(len ? j0 + len : 0u);
        if (top <= base) continue;
        const bool staged = top - base <= kStageCap;
        // scan offset of candidate j in this wavefront's frame
        const int rdy = (int)(r % 3u) - 1, rdz = (int)(r / 3u) - 1;
        const uint32_t rowbase = (row_ref + (uint32_t)(rdz * g.dims[1] + rdy)) * (uint32_t)nx;
        const float ofy = (float)rdy * u.cs_u, ofz = (float)rdz * u.cs_u;
        auto cand = [&](uint32_t j) {
          return probe_frame ? probe_offset(j, rowbase, ofy, ofz) : offset(spos[j]);
        };
        float thr = 0.0f;
        if (staged) {
          float mq = 0.0f;
          for (uint32_t kk = lane; kk < top - base; kk += kDensBlock) {
            float4 q = cand(base + kk);
            q.w = __builtin_fmaf(q.x, q.x, __builtin_fmaf(q.y, q.y, q.z * q.z));
            s_q[kk] = q;
            mq = q.w > mq ? q.w : mq;
          }
          // Rounding of the dot form with |s|, |q| <= M: |q|^2 (3 roundings),
          // the three FMAs (partial sums up to 3 M^2), |s|^2 and the threshold,
          // and the FP32 offsets themselves (2 u M per component on a pair vector
          // of length < 2) stay below u (22 M^2 + 14 M + 5); the margin is
          // 32 u (M^2 + M + 1), u = 2^-24, on top of the common probe margin.
          const float m2 = fmaxf(ss_max, wave_max_pos(mq));
          const float h2p = u.H2_probe + 0x1.0p-19f * (m2 + __builtin_sqrtf(m2) + 1.0f);
          thr = h2p - ss;
          __syncthreads();
        }
        if (len) {
          if (staged) {
            scan_run([&](uint32_t k0, uint32_t cnt) {
                       return scan_word_dot([&](uint32_t j) { return s_q[j - base]; }, k0, cnt, ax, ay, az, thr);
                     },
                     j0, len, r, words, n, i, m[2u * t], m[2u * t + 1u]);
          } else {
            scan_run([&](uint32_t k0, uint32_t cnt) {
                       return scan_word([&](uint32_t j) { return cand(j < n ? j : n - 1); }, k0, cnt, sp);
                     },
                     j0, len, r, words, n, i, m[2u * t], m[2u * t + 1u]);
          }
        }
        if (staged) __syncthreads();  // the next run's staging overwrites s_q
      }
      // the group's words in stream order, each with the slot of its bit 0
      uint32_t wb[2 * G];
#pragma unroll
      for (uint32_t t = 0; t < G; ++t) {
        // mixed form: bit 31 names the run of the group (slots stay below 2^31)
        const uint32_t tag = F == kMixed ? t << 31 : 0u;
        wb[2u * t] = gj0[t] | tag;
        wb[2u * t + 1u] = (gj0[t] + 64u) | tag;
      }
      walk_bits_q<kDensUnr>(m, wb, [&](const uint32_t* jj) {
        uint32_t js[kDensUnr], tv[kDensUnr];
#pragma unroll
        for (int v = 0; v < kDensUnr; ++v) {
          const uint32_t q = jj[v] == kPad ? jj[0] : jj[v];
          tv[v] = F == kMixed ? q >> 31 : 0u;
          js[v] = F == kMixed ? q & 0x7FFFFFFFu : q;
        }
        G4 pj[kDensUnr];
#pragma unroll
        for (int v = 0; v < kDensUnr; ++v) pj[v] = pos[js[v]];
#pragma unroll
        for (int v = 0; v < kDensUnr; ++v) {
          if (jj[v] == kPad) continue;
          real dx, dy, dz;
          if constexpr (F == kMixed) {
            dx = (pi.x - pj[v].x) - (float)cell_shift(__float_as_uint(pi.w), __float_as_uint(pj[v].w), 0) * u.cs_u;
            dy = (tv[v] ? gby[1] : gby[0]) - pj[v].y;
            dz = (tv[v] ? gbz[1] : gbz[0]) - pj[v].z;
          } else {
            form_pair<F>(pi, pj[v], u.cs_u, dx, dy, dz);
          }
          real r2 = dot3(dx, dy, dz, dx, dy, dz);
          if (r2 < H2) {
            // the self pair (and coincident slots) has r2 = 0, q = 0
            real q = r2 > real(0.0) ? sqrt_pos(r2) * inv_h : real(0.0);
            acc.add(spline_f(q));
          }
        }
      });
    }
    if (!live) return;
    if (overflow) atomicOr(&flags[FLAG_CAPACITY], 1u);
    if constexpr (F == kF64) {
      const double rho = kc.msig * acc.s;
      const double p = tait_p(rho, kc);
      const double pr = p / (rho * rho);
      rp[i] = make_double2(rho, pr);
      reinterpret_cast<double*>(wpos + i)[3] = rho;
      reinterpret_cast<double*>(wvel + i)[3] = pr;
      pres[i] = p;
    } else if constexpr (F == kF32) {
      float rho = (float)kc.msig * acc.s;
      float ratio = rho * (float)kc.inv_rho0;
      float r2_ = ratio * ratio;
      float r4 = r2_ * r2_;
      float r7 = r4 * r2_ * ratio;
      float p = (float)kc.B * (r7 - 1.0f);
      if (p < 0.0f) p = 0.0f;
      const float pr = p / (rho * rho);
      rp[i] = make_float2(rho, pr);
      reinterpret_cast<float*>(wpos + i)[3] = rho;
      reinterpret_cast<float*>(wvel + i)[3] = pr;
      pres[i] = p;
      if ((dry_mask >> lane) & 1ull) {
        const uint32_t id = dw.ids[i];
        static_cast<float2*>(dw.rp_id)[id] = make_float2(rho, pr);
        static_cast<float*>(dw.pres_id)[id] = p;
      }
    } else {
      // EOS in FP64 from the compensated sum: gamma = 7 amplifies the relative
      // error of rho by 7 rho^7 / (rho^7 - rho0^7), so rho is never rounded to
      // FP32 before p is formed.
      double rho = kc.msig * acc.wide();
      double ratio = rho * kc.inv_rho0;
      double r2_ = ratio * ratio;
      double r4 = r2_ * r2_;
      double r7 = r4 * r2_ * ratio;
      double p = kc.B * (r7 - 1.0);
      if (p < 0.0) p = 0.0;
      const float2 rpi = make_float2((float)rho, (float)(p / (rho * rho)));
      rp[i] = rpi;
      pres[i] = (float)p;
      if ((dry_mask >> lane) & 1ull) {
        const uint32_t id = dw.ids[i];
        static_cast<float2*>(dw.rp_id)[id] = rpi;
        static_cast<float*>(dw.pres_id)[id] = (float)p;
      }
    }
  });
}

// ---------------------------------------------------------------------------
// momentum + artificial viscosity + symplectic Euler + next cell key.
//   a_i = g - m sigma/h sum_j (pr_i + pr_j + Pi_ij) f'(q)/r x_ij,
//   Pi_ij = -alpha c0 mu_ij / (0.5 (rho_i + rho_j)) if v_ij.x_ij < 0,
//   mu_ij = h v_ij.x_ij / (r^2 + 0.01 h^2);
//   v' = v + dt a, x' = x + dt v'.  Fixed particles (id >= n_moving) copy
//   through.  State is read from (spos, svel) and written to (opos, ovel):
//   neighbours keep reading the pre-step state.
//
// One wavefront per block, walking the row runs in lockstep.  For each run
// the wavefront first copies the gather data of the union of its lanes' hit
// ranges into LDS (coalesced; the lanes of a wavefront share most
// neighbours), then every lane walks its own mask with LDS reads instead of
// dependent global loads.  A union wider than kForceCap slots is walked from
// global memory.
// ---------------------------------------------------------------------------
// LDS array of staged 4-vectors.  A double4 entry is kept as two 16-byte
// halves in two arrays: lanes reading nearby entries of a run then touch
// nearby 16-byte slots, which spreads a wavefront's reads over more LDS banks
// than 32-byte entries do.  float4 entries are 16 bytes already.
template <class T4, uint32_t N>
struct StageArr {
  T4 v[N];
  __device__ void put(uint32_t k, const T4& x) { v[k] = x; }
  __device__ T4 get(uint32_t k) const { return v[k]; }
};
template <uint32_t N>
struct StageArr<double4, N> {
  double2 lo[N], hi[N];
  __device__ void put(uint32_t k, const double4& x) {
    lo[k] = make_double2(x.x, x.y);
    hi[k] = make_double2(x.z, x.w);
  }
  __device__ double4 get(uint32_t k) const {
    const double2 a = lo[k], b = hi[k];
    return make_double4(a.x, a.y, b.x, b.y);
  }
};

constexpr int kForceBlock = 64;  // one wavefront: block barriers are wave barriers
template <int F>
constexpr int kForceWaves = F == kF64 ? 4 : 5;
constexpr uint32_t kForceCap = 160;
// Mask words loaded one run group ahead in the force walk: in the FP64 form
// (1152 -> 1098 us at 4M); the mixed form, at 5 wavefronts per SIMD, has no
// registers to spare for them (888 -> 951 us), and the FP32 form gains
// nothing (745 -> 752 us).
template <int F>
constexpr bool kForcePrefetch = F == kF64;

// One momentum pair: sc with a_i -= (m sigma / h) sc x_ij for 0 < r < 2h;
// false outside (then nothing is added).  sc = (pr_i + pr_j + Pi_ij) f'(q) / r
// with Pi_ij = -alpha c0 h (v.r) / ((r^2 + eps2) rhobar) for approaching
// pairs, put over one common denominator: a single division per pair.  ach is
// alpha c0 times hv.
template <int F>
__device__ inline bool force_pair_sc(typename SphForm<F>::real dx, typename SphForm<F>::real dy,
                                     typename SphForm<F>::real dz, const typename SphForm<F>::G4& vi,
                                     typename SphForm<F>::real rhoi, typename SphForm<F>::real pri,
                                     const typename SphForm<F>::G4& vj, typename SphForm<F>::real rhoj,
                                     typename SphForm<F>::real prj, const SphUnits& u,
                                     typename SphForm<F>::real ach, typename SphForm<F>::real& sc) {
  using real = typename SphForm<F>::real;
  const real H2 = (real)u.H2, inv_h = (real)u.inv_h, eps2 = (real)u.eps2;
  real r2 = dot3(dx, dy, dz, dx, dy, dz);
  if (!(r2 < H2 && r2 > real(0.0))) return false;
  real rr = sqrt_pos(r2);
  real q = rr * inv_h;
  real dw = spline_df(q);
  real dvx = vi.x - vj.x, dvy = vi.y - vj.y, dvz = vi.z - vj.z;
  real vdotr = dot3(dvx, dvy, dvz, dx, dy, dz);
  real num, den;
  if (vdotr < real(0.0)) {
    real drb = (r2 + eps2) * (real(0.5) * (rhoi + rhoj));
    num = fmad(pri + prj, drb, -(ach * vdotr)) * dw;
    den = drb * rr;
  } else {
    num = (pri + prj) * dw;
    den = rr;
  }
  sc = div_pos(num, den);
  return true;
}

template <int F>
__global__ __launch_bounds__(kForceBlock, kForceWaves<F>) void k_sph_force(
    uint32_t n, uint64_t n_moving, const typename SphForm<F>::G4* __restrict__ gpos,
    const typename SphForm<F>::G4* __restrict__ gvel, const typename SphForm<F>::R2* __restrict__ rp,
    const uint32_t* __restrict__ ids, const uint32_t* __restrict__ skeys,
    const uint2* __restrict__ meta, const uint32_t* __restrict__ words, Grid g, SphConst kc,
    SphUnits u, const typename SphForm<F>::S4* __restrict__ spos,
    const typename SphForm<F>::S4* __restrict__ svel, typename SphForm<F>::S4* __restrict__ opos,
    typename SphForm<F>::S4* __restrict__ ovel, uint32_t* __restrict__ okeys,
    typename SphForm<F>::S4* __restrict__ oacc, MovedOut moved, uint32_t* flags, TileQueue tq,
    uint32_t* __restrict__ movcell, uint32_t next_pass) {
  using real = typename SphForm<F>::real;
  using G4 = typename SphForm<F>::G4;
  using R2 = typename SphForm<F>::R2;
  using S4 = typename SphForm<F>::S4;
  using sreal = typename SphForm<F>::sreal;
  __shared__ StageArr<G4, kForceCap> s_p, s_v;
  __shared__ R2 s_rp[F == kMixed ? kForceCap : 1];
  const uint32_t lane = threadIdx.x;
  for_each_tile(tq, [&](uint32_t tile) {
    const uint32_t i = tile * kForceBlock + lane;
    const bool live = i < n && (uint64_t)ids[i] < n_moving;
    // (rho, p/rho^2) of a slot: the density pass also left them in the .w
    // lanes of the FP64 / FP32 gather copies
    G4 pi{}, vi{};
    real rhoi = real(0.0), pri = real(0.0);
    if (live) {
      pi = gpos[i];
      vi = gvel[i];
      if constexpr (F == kMixed) {
        const R2 rpi = rp[i];
        rhoi = rpi.x;
        pri = rpi.y;
      } else {
        rhoi = pi.w;
        pri = vi.w;
      }
    }
    const real alpha_c0 = (real)kc.alpha_c0, ach = alpha_c0 * (real)u.hv;
    Sum<real, F == kMixed> sx, sy, sz;
    ProbeRunBases rb;   // mixed form: per-run bases of the pair vector
    auto begin_group = [&](uint32_t grp) {
      if constexpr (F == kMixed) rb.begin(pi, grp, u.cs_u);
    };
    // pair bodies over a batch of neighbours fetched by get(j, p, v, rho, pr)
    auto body = [&](const uint32_t* jj, const uint32_t* tt, auto&& get) {
      G4 pj[kUnr], vj[kUnr];
      real rhoj[kUnr], prj[kUnr];
  #pragma unroll
      for (int v = 0; v < kUnr; ++v) get(jj[v] == kPad ? jj[0] : jj[v], pj[v], vj[v], rhoj[v], prj[v]);
  #pragma unroll
      for (int v = 0; v < kUnr; ++v) {
        if (jj[v] == kPad) continue;
        real dx, dy, dz, sc;
        if constexpr (F == kMixed) rb.pair(pi, pj[v], tt[v], u.cs_u, dx, dy, dz);
        else form_pair<F>(pi, pj[v], u.cs_u, dx, dy, dz);
        if (force_pair_sc<F>(dx, dy, dz, vi, rhoi, pri, vj[v], rhoj[v], prj[v], u, ach, sc)) {
          if constexpr (F == kMixed) {
            // compensated sums of the FP32 terms
            sx.add(sc * dx);
            sy.add(sc * dy);
            sz.add(sc * dz);
          } else {
            sx.s = fmad(sc, dx, sx.s);
            sy.s = fmad(sc, dy, sy.s);
            sz.s = fmad(sc, dz, sz.s);
          }
        }
      }
    };
    walk_runs_staged<kUnr, kForceCap, kForcePrefetch<F>>(
        meta, words, n, i, live, lane, begin_group,
        [&](uint32_t k, uint32_t j) {
          s_p.put(k, gpos[j]);
          s_v.put(k, gvel[j]);
          if constexpr (F == kMixed) s_rp[k] = rp[j];
        },
        [&](const uint32_t* jj, const uint32_t* tt) {
          body(jj, tt, [&](uint32_t k, G4& p, G4& v, real& rho, real& pr) {
            p = s_p.get(k);
            v = s_v.get(k);
            if constexpr (F == kMixed) {
              const R2 t = s_rp[k];
              rho = t.x;
              pr = t.y;
            } else {
              rho = p.w;
              pr = v.w;
            }
          });
        },
        [&](const uint32_t* jj, const uint32_t* tt) {
          body(jj, tt, [&](uint32_t j, G4& p, G4& v, real& rho, real& pr) {
            p = gpos[j];
            v = gvel[j];
            if constexpr (F == kMixed) {
              const R2 t = rp[j];
              rho = t.x;
              pr = t.y;
            } else {
              rho = p.w;
              pr = v.w;
            }
          });
        },
        [] {});
    if (i >= n) return;
    const uint64_t live_mask = __ballot(live);
    if (!live) {
      opos[i] = spos[i];
      ovel[i] = svel[i];
      okeys[i] = skeys[i];
      if (oacc) oacc[i] = S4{0, 0, 0, 0};
      moved_word_fixed(moved, i, live_mask);
      return;
    }
    const S4 xs = spos[i];
    const S4 vs = svel[i];
    sreal ax, ay, az;
    if constexpr (F == kF32) {
      const float msh = (float)kc.msh;
      ax = (float)kc.g[0] - msh * sx.s;
      ay = (float)kc.g[1] - msh * sy.s;
      az = (float)kc.g[2] - msh * sz.s;
    } else {
      ax = kc.g[0] - kc.msh * sx.wide();
      ay = kc.g[1] - kc.msh * sy.wide();
      az = kc.g[2] - kc.msh * sz.wide();
    }
    const sreal dt = (sreal)kc.dt;
    S4 v2 = vs, x2 = xs;
    v2.x = vs.x + dt * ax;
    v2.y = vs.y + dt * ay;
    v2.z = vs.z + dt * az;
    x2.x = xs.x + dt * v2.x;
    x2.y = xs.y + dt * v2.y;
    x2.z = xs.z + dt * v2.z;
    opos[i] = x2;
    ovel[i] = v2;
    if (oacc) oacc[i] = S4{ax, ay, az, 0};
    if (!isfinite(ax) || !isfinite(ay) || !isfinite(az)) atomicOr(&flags[FLAG_NONFINITE], 1u);
    if ((double)v2.x * (double)v2.x + (double)v2.y * (double)v2.y + (double)v2.z * (double)v2.z > kc.cfl_v2)
      atomicOr(&flags[FLAG_CFL], 1u);
    // next step's cell key, same FP64 floor as the key kernel
    double fx = floor(((double)x2.x - g.lo[0]) * g.inv_cs);
    double fy = floor(((double)x2.y - g.lo[1]) * g.inv_cs);
    double fz = floor(((double)x2.z - g.lo[2]) * g.inv_cs);
    bool ok = fx >= 0.0 && fy >= 0.0 && fz >= 0.0 && fx < (double)g.dims[0] &&
              fy < (double)g.dims[1] && fz < (double)g.dims[2];
    uint32_t key = 0;
    if (ok)
      key = ((uint32_t)fz * (uint32_t)g.dims[1] + (uint32_t)fy) * (uint32_t)g.dims[0] + (uint32_t)fx;
    else
      atomicOr(&flags[FLAG_DOMAIN], 1u);
    okeys[i] = key;
    movcell[key] = next_pass;   // this cell holds a moving particle at the next density pass
    moved_word_live(moved, i, key != skeys[i], live_mask);
  });
}

template <int F>
SphUnits sph_units(const Ctx& c) {
  SphUnits u;
  // The scan only has to keep every pair the exact test of the walk keeps.
  // Its FP32 offsets are rounded from values of magnitude at most E (the
  // grid's extent plus its largest coordinate, in units of h; the second term
  // covers the absolute FP32 coordinates of the F32 form's own test), so an
  // r^2 near the cutoff 4 is off by at most about 4 E 2^-24 relative.  The
  // margin is 2^-10 plus four times that bound; the walk then applies the
  // exact test.
  double E = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double lo = c.grid.lo[a], hi = lo + c.grid.dims[a] * c.grid.cs;
    const double e = (std::fabs(lo) > std::fabs(hi) ? std::fabs(lo) : std::fabs(hi)) + (hi - lo);
    E = e > E ? e : E;
  }
  E *= c.sph.inv_h;
  const double margin = 1.0 + 1.0 / 1024.0 + 16.0 * E * 0x1.0p-24;
  const double cs_h = c.grid.cs * c.sph.inv_h;
  u.cs_u = (float)cs_h;
  u.H2_probe = (float)(4.0 * margin);
  u.scan_inv_h = c.sph.inv_h;
  if constexpr (F == kMixed) {
    u.H2 = 4.0;
    u.inv_h = 1.0;
    u.eps2 = 0.01;
    u.hv = 1.0;
    u.unit_scale = c.sph.inv_h;
  } else {
    u.H2 = c.sph.H2;
    u.inv_h = c.sph.inv_h;
    u.eps2 = c.sph.eps2;
    u.hv = c.sph.h;
    u.unit_scale = 1.0;
  }
  return u;
}

// Per-step inputs of both gathers of form F.
template <int F>
struct SphStepIO {
  SphUnits u;
  const typename SphForm<F>::G4* gpos;
  const typename SphForm<F>::G4* gvel;
  const uint32_t* skeys;
  TileQueue tq;
  DryWalls dw;
  uint32_t pass;
};

template <int F>
SphStepIO<F> sph_step_io(Ctx& c) {
  using G4 = typename SphForm<F>::G4;
  SphStepIO<F> io;
  io.u = sph_units<F>(c);
  io.gpos = (F == kMixed) ? c.off32.as<G4>() : c.pos[c.cur].as<G4>();
  io.gvel = (F == kMixed) ? c.vel32.as<G4>() : c.vel[c.cur].as<G4>();
  io.skeys = c.keys[c.cur_key].as<uint32_t>();
  static_assert(kDensBlock == kForceBlock, "both gathers walk the same tiles");
  const uint32_t n_tiles = blocks_for((uint32_t)c.n, kDensBlock);
  io.tq = TileQueue{c.tile_q.as<uint32_t>(), n_tiles, (n_tiles + 7u) / 8u};
  io.pass = c.dens_pass;
  io.dw = DryWalls{c.ids[c.cur_ids].as<uint32_t>(), c.n_moving, c.movcell.as<uint32_t>(),
                   c.dry_pass.as<uint32_t>(), c.rp_id.p, c.pres_id.p, io.pass, c.stamp_pass == io.pass,
                   c.dry_reuse, c.record_all};
  return io;
}

// Density + EOS pass of form F into rp / pres (and, for F64 / F32, the .w
// lanes of the state's gather copies wpos / wvel).
template <int F>
void sph_density_launch(Ctx& c, const SphStepIO<F>& io, typename SphForm<F>::R2* rp,
                        typename SphForm<F>::real* pres, typename SphForm<F>::G4* wpos,
                        typename SphForm<F>::G4* wvel) {
  using S4 = typename SphForm<F>::S4;
  using real = typename SphForm<F>::real;
  const uint32_t n = (uint32_t)c.n;
  const double ge = (double)sizeof(real);     // bytes per scalar of the pair arithmetic
  // algorithmic bytes per particle: position (3 scalars; the FP32 probe for
  // the mixed form) read once, (rho, p/rho^2, p) written once
  double per = (F == kMixed ? 12.0 : 3.0 * ge) + 3.0 * ge;
  LaunchScope ls(&c, F == kF64 ? "sph_density_f64" : (F == kF32 ? "sph_density_f32" : "sph_density_mixed"),
                 per * n);
  static const int per_cu = blocks_per_cu(k_sph_density<F>, kDensBlock);
  const uint32_t dgrid = queue_grid(per_cu, c.n_cus, io.tq.n_tiles);
  k_sph_density<F><<<dgrid, kDensBlock, 0, c.stream>>>(n, c.pos[c.cur].as<S4>(), io.gpos, io.skeys,
                                                      c.lb.as<uint32_t>(), c.grid, c.sph, io.u, c.nmeta.as<uint2>(),
                                                      c.nword.as<uint32_t>(), rp, wpos, wvel, pres,
                                                      c.flags.as<uint32_t>(), io.tq, io.dw);
  PR_HIP(hipGetLastError());
}

// Host driver of one WCSPH step in form F: rebuild, density, force.
template <int F>
void sph_step_impl(Ctx& c) {
  using G4 = typename SphForm<F>::G4;
  using R2 = typename SphForm<F>::R2;
  using S4 = typename SphForm<F>::S4;
  using real = typename SphForm<F>::real;
  neighbors_build(c);
  const uint32_t n = (uint32_t)c.n;
  if (!n) return;
  ++c.dens_pass;
  const SphStepIO<F> io = sph_step_io<F>(c);
  const SphUnits& u = io.u;
  const G4* gpos = io.gpos;
  const G4* gvel = io.gvel;
  const uint32_t* skeys = io.skeys;
  const TileQueue& tq = io.tq;
  const uint32_t pass = io.pass;
  const double se = (F == kF32) ? 4.0 : 8.0;   // bytes per state scalar
  const double ge = (double)sizeof(real);     // bytes per scalar of the pair arithmetic
  if (F == kMixed && (c.desc.flags & PRESTIGE_FLAG_DENSITY_F64)) {
    sph_density_f64_for_mixed(c);
  } else {
    // .w lanes of the FP64 / FP32 gather copies receive (rho, p/rho^2)
    G4* wpos = (F == kMixed) ? nullptr : c.pos[c.cur].as<G4>();
    G4* wvel = (F == kMixed) ? nullptr : c.vel[c.cur].as<G4>();
    sph_density_launch<F>(c, io, c.rp.as<R2>(), c.pres.as<real>(), wpos, wvel);
  }
  {
    // algorithmic bytes per moving particle: position, velocity (3 + 3
    // scalars of the pair arithmetic), (rho, p/rho^2), state in (mixed: the
    // FP64 copy) and out (6 + 6 state scalars), next key
    const double nm = (double)c.n_moving;
    double per = 8.0 * ge + (F == kMixed ? 6.0 * se : 0.0) + 6.0 * se + 4.0;
    LaunchScope ls(&c, F == kF64 ? "sph_force_f64" : (F == kF32 ? "sph_force_f32" : "sph_force_mixed"), per * nm);
    const int dst = c.cur ^ 1;
    const int kdst = c.cur_key ^ 1;
    static const int per_cu = blocks_per_cu(k_sph_force<F>, kForceBlock);
    const uint32_t fgrid = queue_grid(per_cu, c.n_cus, tq.n_tiles);
    k_sph_force<F><<<fgrid, kForceBlock, 0, c.stream>>>(
        n, c.n_moving, gpos, gvel, c.rp.as<R2>(), c.ids[c.cur_ids].as<uint32_t>(), skeys,
        c.nmeta.as<uint2>(), c.nword.as<uint32_t>(), c.grid, c.sph, u, c.pos[c.cur].as<S4>(), c.vel[c.cur].as<S4>(),
        c.pos[dst].as<S4>(), c.vel[dst].as<S4>(), c.keys[kdst].as<uint32_t>(),
        c.keep_acc ? c.acc.as<S4>() : nullptr, MovedOut{c.inc_mask.as<uint64_t>(), c.inc_tiles.as<uint32_t>()},
        c.flags.as<uint32_t>(), tq, c.movcell.as<uint32_t>(), pass + 1u);
    PR_HIP(hipGetLastError());
    c.cur = dst;
    c.cur_key = kdst;
    c.sorted = false;
    c.keys_valid = true;
    c.inc_ready = true;   // keys[cur_key ^ 1] and lb still describe the slots the step kept
    c.stamp_pass = pass + 1u;
  }
}

}  // namespace prestige
