// jit.hip — runs the gather passes prestige_generate_hip writes for fused
// equations: hiprtc compiles the program for the device's architecture (once
// per context and source text), fields are bound by name, and the passes run
// in order on the context stream.  The HIP counterpart of executing what
// prestige/src/codegen/simple_cpu.rs:3-22 prints.
#include <hip/hiprtc.h>

#include <cmath>

#include <string>
#include <vector>

#include "codegen.hpp"
#include "common.hpp"

namespace prestige {

namespace {

// positions in particle-id order (the all-pairs form walks ids)
template <class T4>
__global__ __launch_bounds__(kBlock) void k_pos_to_ids(uint32_t n, const T4* __restrict__ pos,
                                                       const uint32_t* __restrict__ ids, T4* __restrict__ out) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[ids[k]] = pos[k];
}

// a field in slot order for the cell loop's [j] reads: out[k] = f[ids[k]]
__global__ __launch_bounds__(kBlock) void k_field_to_slots(uint32_t n, const uint32_t* __restrict__ ids,
                                                           const double* __restrict__ f,
                                                           double* __restrict__ out) {
  const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k < n) out[k] = f[ids[k]];
}

// Layout of prestige_consts in the generated source.
struct GenConsts {
  double h, dt, rho0, c0, alpha, kn, kt, mu, gx, gy, gz;
  double two_zeta, neg_inv_kt;
};

DevBuf& kept_buf(std::map<std::string, std::unique_ptr<DevBuf>>& m, const std::string& name, size_t bytes) {
  std::unique_ptr<DevBuf>& b = m[name];
  if (!b) b.reset(new DevBuf());
  b->alloc(bytes);
  return *b;
}

hipModule_t module_for(Ctx& c, const std::string& src) {
  auto it = c.jit.mods.find(src);
  if (it != c.jit.mods.end()) return it->second;
  hipDeviceProp_t prop;
  PR_HIP(hipGetDeviceProperties(&prop, c.device));
  const std::string arch = std::string("--offload-arch=") + prop.gcnArchName;
  const char* opts[] = {arch.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcProgram prog;
  PR_CHECK(hiprtcCreateProgram(&prog, src.c_str(), "prestige_generated.hip", 0, nullptr, nullptr) ==
               HIPRTC_SUCCESS,
           PRESTIGE_E_HIP, "hiprtcCreateProgram failed");
  const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(PRESTIGE_E_INVALID, "generated equation code does not compile: " + log);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  hipModule_t mod;
  PR_HIP(hipModuleLoadData(&mod, code.data()));
  c.jit.mods.emplace(src, mod);
  return mod;
}

// Device array of a field the passes use (FP64, id order): force and mass
// are eq1's arrays, every other name is a named field of the context.
// A missing field that the passes write is created zero-filled.
double* field_ptr(Ctx& c, const GenField& f) {
  const size_t bytes = (c.n ? c.n : 1) * sizeof(double);
  if (f.name == "force" || f.name == "mass") {
    DevBuf& b = f.name == "force" ? c.force : c.mass;
    bool& have = f.name == "force" ? c.have_force : c.have_mass;
    if (!have) {
      PR_CHECK(f.written, PRESTIGE_E_STATE, "field '" + f.name + "' was never uploaded");
      b.alloc(bytes);
      PR_HIP(hipMemsetAsync(b.p, 0, bytes, c.stream));
      have = true;
    }
    return b.as<double>();
  }
  auto it = c.gfields.find(f.name);
  if (it == c.gfields.end() && !f.written) {
    // a field of the context's model (velocities, radii, WCSPH outputs, ...), read only
    DevBuf& b = kept_buf(c.gmodel, f.name, bytes);
    if (model_field_ids(c, f.name, b.as<double>())) return b.as<double>();
  }
  if (it == c.gfields.end()) {
    PR_CHECK(f.written, PRESTIGE_E_STATE, "field '" + f.name + "' was never uploaded");
    PR_CHECK(generic_field_name(c, f.name), PRESTIGE_E_INVALID,
             "field '" + f.name + "' is a model field and cannot be written by an equation");
    std::unique_ptr<DevBuf> b(new DevBuf());
    b->alloc(bytes);
    PR_HIP(hipMemsetAsync(b->p, 0, bytes, c.stream));
    it = c.gfields.emplace(f.name, std::move(b)).first;
  }
  return it->second->as<double>();
}

DevBuf& field_buf(Ctx& c, const std::string& name) {
  if (name == "force") return c.force;
  if (name == "mass") return c.mass;
  return *c.gfields.at(name);
}

}  // namespace

void run_equations(Ctx& c, const int32_t* eq_ids, uint32_t neq, int32_t loop, double cutoff) {
  PR_CHECK(neq == 0 || eq_ids, PRESTIGE_E_INVALID, "null equation list");
  const bool f32 = c.desc.precision == PRESTIGE_F32;
  GenProgram prog;
  try {
    prog = generate_hip_program(eq_ids, neq, loop, f32);
  } catch (const GenError& g) {
    throw Error(g.code, g.msg);
  }
  uint32_t n = (uint32_t)c.n;
  if (!n || prog.passes.empty()) return;
  if (loop == PRESTIGE_LOOP_CELLS) {
    PR_CHECK(cutoff > 0.0 && cutoff <= c.grid.cs, PRESTIGE_E_INVALID, "cutoff must be in (0, cell_size]");
    if (!c.sorted) neighbors_build(c);
  }
  const hipModule_t mod = module_for(c, prog.source);
  const size_t pe = f32 ? sizeof(float4) : sizeof(double4);
  const void* pos = c.pos[c.cur].p;
  if (loop == PRESTIGE_LOOP_ALLPAIRS) {
    c.posid.alloc((size_t)n * pe);
    if (f32)
      k_pos_to_ids<float4><<<blocks_for(n), kBlock, 0, c.stream>>>(n, c.pos[c.cur].as<float4>(),
                                                                   c.ids[c.cur_ids].as<uint32_t>(),
                                                                   c.posid.as<float4>());
    else
      k_pos_to_ids<double4><<<blocks_for(n), kBlock, 0, c.stream>>>(n, c.pos[c.cur].as<double4>(),
                                                                    c.ids[c.cur_ids].as<uint32_t>(),
                                                                    c.posid.as<double4>());
    PR_HIP(hipGetLastError());
    pos = c.posid.p;
  }
  const prestige_desc& d = c.desc;
  GenConsts k{d.h,  d.dt, d.rho0, d.c0, d.alpha, d.kn, d.kt, d.mu, d.gravity[0], d.gravity[1], d.gravity[2],
              c.dem.two_zeta, c.dem.neg_inv_kt};
  const uint32_t* ids = c.ids[c.cur_ids].as<uint32_t>();
  const uint32_t* skeys = c.keys[c.cur_key].as<uint32_t>();
  const uint32_t* lb = c.lb.as<uint32_t>();
  int nx = c.grid.dims[0], ny = c.grid.dims[1], nz = c.grid.dims[2];
  double c2 = cutoff * cutoff;
  // scan threshold of the cell walk in units of the cutoff: 1 plus the margin
  // that keeps the FP32 scan a superset of the exact FP64 test (rounding of
  // offsets up to the grid's extent E, in cutoffs, and of the test itself)
  double inv_cut = loop == PRESTIGE_LOOP_CELLS ? 1.0 / cutoff : 0.0;
  float thr = loop == PRESTIGE_LOOP_CELLS ? cell_scan_threshold(c.grid, cutoff) : 0.0f;
  for (const GenPass& gp : prog.passes) {
    hipFunction_t fn;
    PR_HIP(hipModuleGetFunction(&fn, mod, gp.kernel.c_str()));
    // argument values must stay at fixed addresses until the launch
    std::vector<void*> fptr;
    fptr.reserve(2 * gp.fields.size());
    std::vector<std::string> swap_names;
    for (const GenField& f : gp.fields) {
      fptr.push_back(field_ptr(c, f));
      if (f.split) {
        fptr.push_back(kept_buf(c.gsplit, f.name, (size_t)n * sizeof(double)).p);
        swap_names.push_back(f.name);
      }
    }
    // slot-order copies of the fields read at [j] (cell loop), taken from the
    // pass's inputs before it runs
    std::vector<void*> gptr;
    gptr.reserve(gp.slot_fields.size());
    for (const GenField& f : gp.slot_fields) {
      const double* src = field_ptr(c, f);
      DevBuf& o = kept_buf(c.gslots, f.name, (size_t)n * sizeof(double));
      {
        LaunchScope ls(&c, "generated_slot_copy", 20.0 * n);
        k_field_to_slots<<<blocks_for(n), kBlock, 0, c.stream>>>(n, ids, src, o.as<double>());
        PR_HIP(hipGetLastError());
      }
      gptr.push_back(o.p);
    }
    std::vector<void*> args;
    args.push_back(&n);
    args.push_back(&pos);
    if (loop == PRESTIGE_LOOP_CELLS) {
      args.push_back(&ids);
      args.push_back(&skeys);
      args.push_back(&lb);
      args.push_back(&nx);
      args.push_back(&ny);
      args.push_back(&nz);
      args.push_back(&c2);
      args.push_back(&inv_cut);
      args.push_back(&thr);
    }
    args.push_back(&k);
    for (void*& p : fptr) args.push_back(&p);
    for (void*& p : gptr) args.push_back(&p);
    {
      LaunchScope ls(&c, "generated_pass", 0.0);
      // cell passes: one wavefront per block (their scan stages in LDS per wavefront)
      const int bs = loop == PRESTIGE_LOOP_CELLS ? 64 : kBlock;
      PR_HIP(hipModuleLaunchKernel(fn, blocks_for(n, bs), 1, 1, bs, 1, 1, 0, c.stream, args.data(), nullptr));
    }
    PR_HIP(hipStreamSynchronize(c.stream));
    // the pass wrote the split fields' new values to their outputs
    for (const std::string& nm : swap_names) std::swap(field_buf(c, nm).p, c.gsplit.at(nm)->p);
  }
}

}  // namespace prestige
