// equations.cpp — the reference's equation interface, carried over as data:
//   EquationIR {name, reads, writes, body}   prestige/src/equations/ir.rs:3-14
//   #[equation] read/write inference          prestige_macros/src/lib.rs:64-134
//   fuse(Vec<EquationIR>) -> FusedEquations   prestige/src/equations/fuse.rs:14-40
//   generate_simple_cpu(&FusedEquations)      prestige/src/codegen/simple_cpu.rs:3-22
//
// Each built-in equation is stored as the Rust body the #[equation] macro
// would capture, plus the read/write sets that macro infers from it (every
// `X[..]` is a read, LHS included; `X[..] =` and `X[..] op=` are writes) and
// the i/j role of each read that the macro does not record.  The pass
// scheduler is the part the reference leaves out: fused bodies run per (i, j)
// pair in order, so a body that reads `F[j]` of a field an earlier body of
// the same pass writes would see partial sums; such an equation starts a new
// gather pass.  A body that only writes fields earlier bodies read at [j]
// stays in the pass, because the device passes write their outputs to a
// second buffer.  i-only bodies (no [j] read) commute to the end of the j
// loop, which is why the EOS fuses into the density pass.
#include <algorithm>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "../../include/prestige_hip.h"

namespace {

struct Eq {
  const char* name;
  const char* body;                                   // as captured by quote!(#body)
  std::vector<std::pair<std::string, unsigned>> reads;   // name, role bits
  std::vector<std::string> writes;
};

const std::vector<Eq>& table() {
  static const std::vector<Eq> t = {
      {"eq1", "{ force [i] += mass [j] ; }",
       {{"force", PRESTIGE_ROLE_I}, {"mass", PRESTIGE_ROLE_J}},
       {"force"}},
      {"sph_density",
       "{ rho [i] += mass [j] * kernel_w (x [i] - x [j] , y [i] - y [j] , z [i] - z [j] , h) ; }",
       {{"mass", PRESTIGE_ROLE_J},
        {"rho", PRESTIGE_ROLE_I},
        {"x", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"y", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"z", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J}},
       {"rho"}},
      {"sph_eos", "{ p [i] = tait (rho [i] , rho0 , c0) ; }",
       {{"p", PRESTIGE_ROLE_I}, {"rho", PRESTIGE_ROLE_I}},
       {"p"}},
      {"sph_momentum",
       "{ let f = momentum (x [i] - x [j] , y [i] - y [j] , z [i] - z [j] , vx [i] - vx [j] , "
       "vy [i] - vy [j] , vz [i] - vz [j] , p [i] / (rho [i] * rho [i]) + p [j] / (rho [j] * rho [j]) , "
       "rho [i] + rho [j] , mass [j]) ; ax [i] += f . 0 ; ay [i] += f . 1 ; az [i] += f . 2 ; }",
       {{"ax", PRESTIGE_ROLE_I},
        {"ay", PRESTIGE_ROLE_I},
        {"az", PRESTIGE_ROLE_I},
        {"mass", PRESTIGE_ROLE_J},
        {"p", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"rho", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"vx", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"vy", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"vz", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"x", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"y", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"z", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J}},
       {"ax", "ay", "az"}},
      {"integrate",
       "{ vx [i] += dt * ax [i] ; vy [i] += dt * ay [i] ; vz [i] += dt * az [i] ; "
       "x [i] += dt * vx [i] ; y [i] += dt * vy [i] ; z [i] += dt * vz [i] ; }",
       {{"ax", PRESTIGE_ROLE_I},
        {"ay", PRESTIGE_ROLE_I},
        {"az", PRESTIGE_ROLE_I},
        {"vx", PRESTIGE_ROLE_I},
        {"vy", PRESTIGE_ROLE_I},
        {"vz", PRESTIGE_ROLE_I},
        {"x", PRESTIGE_ROLE_I},
        {"y", PRESTIGE_ROLE_I},
        {"z", PRESTIGE_ROLE_I}},
       {"vx", "vy", "vz", "x", "y", "z"}},
      {"dem_contact",
       "{ let c = contact (x [i] - x [j] , y [i] - y [j] , z [i] - z [j] , radius [i] + radius [j] , "
       "vx [i] - vx [j] , vy [i] - vy [j] , vz [i] - vz [j] , wx [i] , wy [i] , wz [i] , wx [j] , wy [j] , "
       "wz [j] , mass [i] , mass [j]) ; fx [i] += c . 0 ; fy [i] += c . 1 ; fz [i] += c . 2 ; "
       "tx [i] += c . 3 ; ty [i] += c . 4 ; tz [i] += c . 5 ; }",
       {{"fx", PRESTIGE_ROLE_I},
        {"fy", PRESTIGE_ROLE_I},
        {"fz", PRESTIGE_ROLE_I},
        {"mass", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"radius", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"tx", PRESTIGE_ROLE_I},
        {"ty", PRESTIGE_ROLE_I},
        {"tz", PRESTIGE_ROLE_I},
        {"vx", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"vy", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"vz", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"wx", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"wy", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"wz", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"x", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"y", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"z", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J}},
       {"fx", "fy", "fz", "tx", "ty", "tz"}},
  };
  return t;
}

void copy_out(const std::string& s, char* dst, size_t cap) {
  if (!dst || cap == 0) return;
  size_t k = std::min(cap - 1, s.size());
  std::memcpy(dst, s.data(), k);
  dst[k] = '\0';
}

bool valid_ids(const int32_t* ids, uint32_t n) {
  if (n && !ids) return false;
  for (uint32_t k = 0; k < n; ++k)
    if (ids[k] < 0 || ids[k] >= (int32_t)table().size()) return false;
  return true;
}

}  // namespace

extern "C" {

prestige_status prestige_equation_describe(int32_t eq_id, char* name, size_t name_cap, char* reads,
                                           size_t reads_cap, char* writes, size_t writes_cap,
                                           uint8_t* read_roles, size_t roles_cap, char* body,
                                           size_t body_cap) {
  if (eq_id < 0 || eq_id >= (int32_t)table().size()) return PRESTIGE_E_INVALID;
  const Eq& e = table()[(size_t)eq_id];
  copy_out(e.name, name, name_cap);
  std::string r, w;
  for (size_t k = 0; k < e.reads.size(); ++k) {
    if (k) r += ",";
    r += e.reads[k].first;
    if (read_roles && k < roles_cap) read_roles[k] = (uint8_t)e.reads[k].second;
  }
  for (size_t k = 0; k < e.writes.size(); ++k) {
    if (k) w += ",";
    w += e.writes[k];
  }
  copy_out(r, reads, reads_cap);
  copy_out(w, writes, writes_cap);
  copy_out(e.body, body, body_cap);
  return PRESTIGE_OK;
}

prestige_status prestige_fuse_schedule(const int32_t* eq_ids, uint32_t n, uint32_t* pass_of, uint32_t* n_passes) {
  if (!valid_ids(eq_ids, n) || (n && !pass_of) || !n_passes) return PRESTIGE_E_INVALID;
  uint32_t pass = 0;
  std::set<std::string> written;   // fields written so far in the current pass
  bool any = false;
  for (uint32_t k = 0; k < n; ++k) {
    const Eq& e = table()[(size_t)eq_ids[k]];
    bool hazard = false;
    for (const auto& rd : e.reads)
      if ((rd.second & PRESTIGE_ROLE_J) && written.count(rd.first)) hazard = true;
    if (hazard) {
      ++pass;
      written.clear();
    }
    for (const auto& wr : e.writes) written.insert(wr);
    pass_of[k] = pass;
    any = true;
  }
  *n_passes = any ? pass + 1 : 0;
  return PRESTIGE_OK;
}

prestige_status prestige_generate_simple_cpu(const int32_t* eq_ids, uint32_t n, char* out, size_t cap, size_t* len) {
  if (!valid_ids(eq_ids, n)) return PRESTIGE_E_INVALID;
  // text layout of prestige/src/codegen/simple_cpu.rs:7-19
  std::string code;
  code += "for i in 0..n {\n";
  code += "    for j in 0..n {\n";
  for (uint32_t k = 0; k < n; ++k) {
    code += "        ";
    code += table()[(size_t)eq_ids[k]].body;
    code += "\n";
  }
  code += "    }\n";
  code += "}\n";
  if (len) *len = code.size();
  copy_out(code, out, cap);
  return PRESTIGE_OK;
}

}  // extern "C"
