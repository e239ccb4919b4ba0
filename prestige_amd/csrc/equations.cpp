// equations.cpp — the reference's equation interface, carried over as data:
//   EquationIR {name, reads, writes, body}   prestige/src/equations/ir.rs:3-14
//   #[equation] read/write inference          prestige_macros/src/lib.rs:64-134
//   fuse(Vec<EquationIR>) -> FusedEquations   prestige/src/equations/fuse.rs:14-40
//   generate_simple_cpu(&FusedEquations)      prestige/src/codegen/simple_cpu.rs:3-22
// and the HIP code generator that is the GPU sibling of generate_simple_cpu.
//
// Each equation is the Rust body the #[equation] macro captures
// (quote!(#body) text), the read/write sets that macro infers from it (every
// `X[..]` is a read, LHS included; `X[..] =` and `X[..] op=` are writes) and
// the i/j role of each field, which the macro does not record: the index
// inside the brackets.  Built-in equations are listed below with their sets
// written out; prestige_equation_define registers further bodies and infers
// the sets with the same rule.
//
// Pass scheduling is the part the reference leaves out: fused bodies run per
// (i, j) pair in order, so a body that reads `F[j]` of a field an earlier body
// of the same pass writes would see partial sums; such an equation starts a
// new gather pass.  A body that only writes fields earlier bodies read at [j]
// stays in the pass: generated passes give such a field a separate output
// array.  i-only bodies (no [j] access) commute to the end of the j loop and
// run once per i, which is why the EOS fuses into the density pass; a body
// with [j] accesses that reads a field an earlier i-only body of the pass
// writes starts a new pass, so it sees the final value.
#include <algorithm>
#include <cctype>
#include <cstring>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/prestige_hip.h"
#include "codegen.hpp"

namespace prestige {
namespace {

struct Eq {
  std::string name;
  std::string body;                                     // as captured by quote!(#body)
  std::vector<std::pair<std::string, unsigned>> reads;  // name, role bits
  std::vector<std::string> writes;
};

std::vector<Eq> builtins() {
  return {
      {"eq1", "{ force [i] += mass [j] ; }", {{"force", PRESTIGE_ROLE_I}, {"mass", PRESTIGE_ROLE_J}}, {"force"}},
      {"sph_density",
       "{ rho [i] += mass [j] * kernel_w (x [i] - x [j] , y [i] - y [j] , z [i] - z [j] , h) ; }",
       {{"mass", PRESTIGE_ROLE_J},
        {"rho", PRESTIGE_ROLE_I},
        {"x", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"y", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J},
        {"z", PRESTIGE_ROLE_I | PRESTIGE_ROLE_J}},
       {"rho"}},
      {"sph_eos", "{ p [i] = tait (rho [i] , rho0 , c0) ; }", {{"p", PRESTIGE_ROLE_I}, {"rho", PRESTIGE_ROLE_I}},
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
       "{ let c = contact (x [i] - x [j] , y [i] - y [j] , z [i] - z [j] , radius [i] , radius [j] , "
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
}

// Built-ins first (ids 0 .. PRESTIGE_EQ_COUNT - 1), then defined equations.
std::mutex g_mutex;
std::vector<Eq>& registry() {
  static std::vector<Eq> r = builtins();
  return r;
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
    if (ids[k] < 0 || ids[k] >= (int32_t)registry().size()) return false;
  return true;
}

bool is_ident(const std::string& s) {
  if (s.empty() || !(std::isalpha((unsigned char)s[0]) || s[0] == '_')) return false;
  for (char ch : s)
    if (!(std::isalnum((unsigned char)ch) || ch == '_')) return false;
  return true;
}

// ---------------------------------------------------------------------------
// body tokens
// ---------------------------------------------------------------------------
enum TokKind { kIdent, kNumber, kPunct };
struct Tok {
  TokKind kind;
  std::string s;
};

std::vector<Tok> tokenize(const std::string& b) {
  static const char* multi[] = {"<<=", ">>=", "+=", "-=", "*=", "/=", "%=", "^=", "&=", "|=", "==", "!=",
                                "<=",  ">=",  "&&", "||", "::", "->", "<<", ">>", ".."};
  std::vector<Tok> t;
  size_t k = 0;
  while (k < b.size()) {
    const char ch = b[k];
    if (std::isspace((unsigned char)ch)) {
      ++k;
      continue;
    }
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      size_t e = k;
      while (e < b.size() && (std::isalnum((unsigned char)b[e]) || b[e] == '_')) ++e;
      t.push_back({kIdent, b.substr(k, e - k)});
      k = e;
      continue;
    }
    if (std::isdigit((unsigned char)ch)) {
      // digits [. digits] [e[+-]digits] [f64 | f32]; underscores dropped; a
      // '.' not followed by a digit ends the number (a method call or field)
      std::string num;
      size_t e = k;
      while (e < b.size() && (std::isdigit((unsigned char)b[e]) || b[e] == '_')) {
        if (b[e] != '_') num += b[e];
        ++e;
      }
      if (e + 1 < b.size() && b[e] == '.' && std::isdigit((unsigned char)b[e + 1])) {
        num += '.';
        ++e;
        while (e < b.size() && (std::isdigit((unsigned char)b[e]) || b[e] == '_')) {
          if (b[e] != '_') num += b[e];
          ++e;
        }
      } else if (e < b.size() && b[e] == '.' && (e + 1 >= b.size() || !std::isalpha((unsigned char)b[e + 1]))) {
        num += ".0";  // "1." is a float literal
        ++e;
      }
      if (e < b.size() && (b[e] == 'e' || b[e] == 'E')) {
        size_t f = e + 1;
        if (f < b.size() && (b[f] == '+' || b[f] == '-')) ++f;
        if (f < b.size() && std::isdigit((unsigned char)b[f])) {
          num += b.substr(e, f - e);
          e = f;
          while (e < b.size() && std::isdigit((unsigned char)b[e])) num += b[e++];
        }
      }
      for (const char* suf : {"_f64", "_f32", "f64", "f32"}) {
        const size_t L = std::strlen(suf);
        if (b.compare(e, L, suf) == 0) {
          e += L;
          break;
        }
      }
      t.push_back({kNumber, num});
      k = e;
      continue;
    }
    bool found = false;
    for (const char* m : multi) {
      const size_t L = std::strlen(m);
      if (b.compare(k, L, m) == 0) {
        t.push_back({kPunct, m});
        k += L;
        found = true;
        break;
      }
    }
    if (found) continue;
    if (std::strchr("{}[](),;.+-*/%<>=!&|^:", ch)) {
      t.push_back({kPunct, std::string(1, ch)});
      ++k;
      continue;
    }
    throw GenError{PRESTIGE_E_INVALID, std::string("unexpected character '") + ch + "' in equation body"};
  }
  return t;
}

bool is_assign_op(const std::string& s) {
  return s == "=" || s == "+=" || s == "-=" || s == "*=" || s == "/=" || s == "%=" || s == "^=" ||
         s == "&=" || s == "|=" || s == "<<=" || s == ">>=";
}

// Index expression "X [ ... ]" at token k: returns the position of the
// matching ']' or 0 when k does not start one.
size_t index_close(const std::vector<Tok>& t, size_t k) {
  if (k + 1 >= t.size() || t[k].kind != kIdent || t[k + 1].s != "[") return 0;
  int depth = 0;
  for (size_t m = k + 1; m < t.size(); ++m) {
    if (t[m].s == "[") ++depth;
    if (t[m].s == "]" && --depth == 0) return m;
  }
  throw GenError{PRESTIGE_E_INVALID, "unbalanced '[' in equation body"};
}

// Name the macro records for the indexed expression ending at identifier k
// (prestige_macros/src/lib.rs:75-77, 105-107, 125-127): only a path counts,
// and of a path only its first segment, so `x[i]` records x,
// `crate::m::x[i]` records crate and a field `self.x[i]` (not a path)
// records nothing (empty result).
std::string indexed_path_root(const std::vector<Tok>& t, size_t k) {
  if (k > 0 && t[k - 1].s == ".") return std::string();
  size_t s = k;
  while (s >= 2 && t[s - 1].s == "::" && t[s - 2].kind == kIdent) s -= 2;
  return t[s].s;
}

// The macro's rule (prestige_macros/src/lib.rs:71-133): every indexed path is
// a read, an indexed path on the left of `=` or `op=` is a write; plus the
// role of each access: the index inside the brackets (i, j, or other).
void infer(const std::vector<Tok>& t, Eq& e) {
  std::map<std::string, unsigned> roles;
  std::set<std::string> writes;
  for (size_t k = 0; k < t.size(); ++k) {
    const size_t c = index_close(t, k);
    if (!c) continue;
    const std::string name = indexed_path_root(t, k);
    if (name.empty()) continue;
    unsigned role = 0;
    if (c == k + 3 && t[k + 2].s == "i") role = PRESTIGE_ROLE_I;
    if (c == k + 3 && t[k + 2].s == "j") role = PRESTIGE_ROLE_J;
    roles[name] |= role;
    if (c + 1 < t.size() && is_assign_op(t[c + 1].s)) writes.insert(name);
  }
  e.reads.assign(roles.begin(), roles.end());
  e.writes.assign(writes.begin(), writes.end());
}

// ---------------------------------------------------------------------------
// HIP translation of one body
// ---------------------------------------------------------------------------
bool is_pos(const std::string& s) { return s == "x" || s == "y" || s == "z"; }

// Descriptor constants a body may name.
const std::set<std::string>& const_names() {
  static const std::set<std::string> s = {"h", "dt", "rho0", "c0", "alpha", "kn", "kt", "mu", "gx", "gy", "gz"};
  return s;
}

// Functions a body may call (Rust method or path form is rewritten to these).
const std::map<std::string, std::string>& fn_map() {
  static const std::map<std::string, std::string> m = {
      {"sqrt", "sqrt"}, {"abs", "fabs"}, {"exp", "exp"},     {"ln", "log"},     {"log10", "log10"},
      {"log2", "log2"}, {"powi", "pow"}, {"powf", "pow"},    {"sin", "sin"},    {"cos", "cos"},
      {"tan", "tan"},   {"min", "fmin"}, {"max", "fmax"},    {"floor", "floor"}, {"ceil", "ceil"},
      {"hypot", "hypot"}, {"atan2", "atan2"}};
  return m;
}

// Device helpers provided to generated code.
const std::set<std::string>& helper_names() {
  static const std::set<std::string> s = {"kernel_w", "tait", "momentum", "contact", "pt_make", "pt_get"};
  return s;
}

size_t match_close(const std::vector<Tok>& t, size_t open) {
  const std::string o = t[open].s;
  const std::string c = o == "(" ? ")" : (o == "[" ? "]" : "}");
  int depth = 0;
  for (size_t m = open; m < t.size(); ++m) {
    if (t[m].s == o) ++depth;
    if (t[m].s == c && --depth == 0) return m;
  }
  throw GenError{PRESTIGE_E_INVALID, "unbalanced '" + o + "' in equation body"};
}

size_t match_open(const std::vector<Tok>& t, size_t close) {
  const std::string c = t[close].s;
  const std::string o = c == ")" ? "(" : (c == "]" ? "[" : "{");
  int depth = 0;
  for (size_t m = close + 1; m-- > 0;) {
    if (t[m].s == c) ++depth;
    if (t[m].s == o && --depth == 0) return m;
  }
  throw GenError{PRESTIGE_E_INVALID, "unbalanced '" + c + "' in equation body"};
}

// Rewrite Rust-only syntax into the C subset: `recv.m(args)` -> `f(recv,
// args)`, `f64::m(..)` -> `f(..)`, `std::f64::consts::PI` -> literal,
// `x as f64` -> `x`, `let [mut] v [: T] =` -> `double v =`, `if c {` ->
// `if (c) {`, and a `;` before a closing brace where Rust allows none.
std::vector<Tok> rewrite(std::vector<Tok> t, const std::string& eq) {
  auto bad = [&](const std::string& what) {
    return GenError{PRESTIGE_E_UNSUPPORTED, "equation '" + eq + "': " + what + " has no HIP translation"};
  };
  // paths
  for (size_t k = 0; k + 2 < t.size(); ++k) {
    if (t[k].s == "std" && t[k + 1].s == "::") {
      t.erase(t.begin() + (long)k, t.begin() + (long)k + 2);
      --k;
      continue;
    }
    if ((t[k].s == "f64" || t[k].s == "f32") && t[k + 1].s == "::") {
      if (t[k + 2].s == "consts" && k + 4 < t.size() && t[k + 3].s == "::") {
        const std::string c = t[k + 4].s;
        const char* v = c == "PI" ? "3.141592653589793" : (c == "E" ? "2.718281828459045" : nullptr);
        if (!v) throw bad("constant '" + c + "'");
        t.erase(t.begin() + (long)k, t.begin() + (long)k + 4);
        t[k] = {kNumber, v};
        continue;
      }
      if (!fn_map().count(t[k + 2].s)) throw bad("'" + t[k].s + "::" + t[k + 2].s + "'");
      t.erase(t.begin() + (long)k, t.begin() + (long)k + 2);
      t[k].s = fn_map().at(t[k].s);
      continue;
    }
  }
  // casts
  for (size_t k = 0; k + 1 < t.size(); ++k) {
    if (t[k].kind == kIdent && t[k].s == "as") {
      if (t[k + 1].s != "f64" && t[k + 1].s != "f32") throw bad("cast to '" + t[k + 1].s + "'");
      t.erase(t.begin() + (long)k, t.begin() + (long)k + 2);
      --k;
    }
  }
  // method calls, innermost receivers first (left to right)
  for (;;) {
    size_t dot = 0;
    bool found = false;
    for (size_t k = 1; k + 2 < t.size(); ++k) {
      if (t[k].s == "." && t[k + 1].kind == kIdent && t[k + 2].s == "(") {
        dot = k;
        found = true;
        break;
      }
    }
    if (!found) break;
    const std::string m = t[dot + 1].s;
    if (!fn_map().count(m)) throw bad("method '" + m + "'");
    size_t start = dot - 1;
    if (t[start].s == ")" || t[start].s == "]") {
      start = match_open(t, start);
      if (start > 0 && t[start - 1].kind == kIdent) --start;  // f(..).m() or field[..].m()
    } else if (t[start].kind != kIdent && t[start].kind != kNumber) {
      throw bad("method call on '" + t[start].s + "'");
    }
    const size_t close = match_close(t, dot + 2);
    std::vector<Tok> repl;
    repl.push_back({kIdent, fn_map().at(m)});
    repl.push_back({kPunct, "("});
    repl.insert(repl.end(), t.begin() + (long)start, t.begin() + (long)dot);
    if (close > dot + 3) {
      repl.push_back({kPunct, ","});
      repl.insert(repl.end(), t.begin() + (long)dot + 3, t.begin() + (long)close);
    }
    repl.push_back({kPunct, ")"});
    t.erase(t.begin() + (long)start, t.begin() + (long)close + 1);
    t.insert(t.begin() + (long)start, repl.begin(), repl.end());
  }
  // if expressions (an `if` where a value is expected) become immediately
  // invoked lambdas whose branches return their blocks' last expressions
  for (;;) {
    size_t k = 0;
    bool found = false;
    for (; k < t.size(); ++k) {
      if (t[k].s != "if") continue;
      const std::string prev = k ? t[k - 1].s : "{";
      if (prev != "{" && prev != ";" && prev != "}" && prev != "else") {
        found = true;
        break;
      }
    }
    if (!found) break;
    std::vector<Tok> lam = {{kPunct, "("}, {kPunct, "["}, {kPunct, "&"}, {kPunct, "]"}, {kPunct, "("},
                            {kPunct, ")"}, {kPunct, "{"}};
    size_t at = k;
    for (;;) {   // one `if cond { block }` of the chain per turn
      size_t b = at + 1;
      int depth = 0;
      for (; b < t.size(); ++b) {
        if (t[b].s == "(" || t[b].s == "[") ++depth;
        if (t[b].s == ")" || t[b].s == "]") --depth;
        if (t[b].s == "{" && depth == 0) break;
      }
      if (b >= t.size()) throw bad("if without block");
      const size_t e = match_close(t, b);
      lam.push_back({kIdent, "if"});
      lam.push_back({kPunct, "("});
      lam.insert(lam.end(), t.begin() + (long)at + 1, t.begin() + (long)b);
      lam.push_back({kPunct, ")"});
      // the block with its last expression returned
      size_t last = b + 1;
      int d2 = 0;
      for (size_t m = b + 1; m < e; ++m) {
        if (t[m].s == "(" || t[m].s == "[" || t[m].s == "{") ++d2;
        if (t[m].s == ")" || t[m].s == "]" || t[m].s == "}") --d2;
        if (t[m].s == ";" && d2 == 0) last = m + 1;
      }
      if (last >= e) throw bad("if expression branch without a value");
      lam.push_back({kPunct, "{"});
      lam.insert(lam.end(), t.begin() + (long)b + 1, t.begin() + (long)last);
      lam.push_back({kIdent, "return"});
      lam.insert(lam.end(), t.begin() + (long)last, t.begin() + (long)e);
      lam.push_back({kPunct, ";"});
      lam.push_back({kPunct, "}"});
      if (e + 1 >= t.size() || t[e + 1].s != "else") throw bad("if expression without else");
      lam.push_back({kIdent, "else"});
      if (e + 2 < t.size() && t[e + 2].s == "if") {
        at = e + 2;
        continue;
      }
      if (e + 2 >= t.size() || t[e + 2].s != "{") throw bad("else without block");
      const size_t e2 = match_close(t, e + 2);
      size_t last2 = e + 3;
      int d3 = 0;
      for (size_t m = e + 3; m < e2; ++m) {
        if (t[m].s == "(" || t[m].s == "[" || t[m].s == "{") ++d3;
        if (t[m].s == ")" || t[m].s == "]" || t[m].s == "}") --d3;
        if (t[m].s == ";" && d3 == 0) last2 = m + 1;
      }
      if (last2 >= e2) throw bad("if expression branch without a value");
      lam.push_back({kPunct, "{"});
      lam.insert(lam.end(), t.begin() + (long)e + 3, t.begin() + (long)last2);
      lam.push_back({kIdent, "return"});
      lam.insert(lam.end(), t.begin() + (long)last2, t.begin() + (long)e2);
      lam.push_back({kPunct, ";"});
      lam.push_back({kPunct, "}"});
      lam.push_back({kPunct, "}"});
      lam.push_back({kPunct, ")"});
      lam.push_back({kPunct, "("});
      lam.push_back({kPunct, ")"});
      t.erase(t.begin() + (long)k, t.begin() + (long)e2 + 1);
      t.insert(t.begin() + (long)k, lam.begin(), lam.end());
      break;
    }
  }
  // tuple fields: t.0 -> pt_get<0>(t) (receiver: a name or a parenthesised /
  // called expression)
  for (size_t k = 1; k + 1 < t.size(); ++k) {
    if (!(t[k].s == "." && t[k + 1].kind == kNumber)) continue;
    if (t[k + 1].s.find_first_not_of("0123456789") != std::string::npos) throw bad("tuple field '" + t[k + 1].s + "'");
    size_t start = k - 1;
    if (t[start].s == ")") {
      start = match_open(t, start);
      if (start > 0 && t[start - 1].kind == kIdent) --start;
    } else if (t[start].kind != kIdent) {
      throw bad("tuple field of '" + t[start].s + "'");
    }
    std::vector<Tok> repl = {{kIdent, "pt_get"}, {kPunct, "<"}, t[k + 1], {kPunct, ">"}, {kPunct, "("}};
    repl.insert(repl.end(), t.begin() + (long)start, t.begin() + (long)k);
    repl.push_back({kPunct, ")"});
    t.erase(t.begin() + (long)start, t.begin() + (long)k + 2);
    t.insert(t.begin() + (long)start, repl.begin(), repl.end());
    k = start + repl.size() - 1;
  }
  // tuple expressions (a, b, ..) -> pt_make(a, b, ..): a parenthesis that is
  // no call and holds a comma at its own level
  for (size_t k = 0; k < t.size(); ++k) {
    if (t[k].s != "(") continue;
    if (k > 0 && (t[k - 1].kind == kIdent || t[k - 1].s == ">" || t[k - 1].s == ")" || t[k - 1].s == "]")) continue;
    if (k > 0 && t[k - 1].s == "let") continue;   // destructuring pattern, below
    const size_t e = match_close(t, k);
    int depth = 0;
    bool comma = false;
    for (size_t m = k + 1; m < e; ++m) {
      if (t[m].s == "(" || t[m].s == "[" || t[m].s == "{") ++depth;
      if (t[m].s == ")" || t[m].s == "]" || t[m].s == "}") --depth;
      if (t[m].s == "," && depth == 0) comma = true;
    }
    if (comma) t.insert(t.begin() + (long)k, {kIdent, "pt_make"});
  }
  // let: `let (a, b) = e;` -> auto __tN = e; double a = pt_get<0>(__tN); ...
  // and `let v = e;` -> double v = e (auto when e is a tuple)
  int tmp_id = 0;
  for (size_t k = 0; k < t.size(); ++k) {
    if (t[k].s != "let") continue;
    size_t v = k + 1;
    if (v < t.size() && t[v].s == "mut") t.erase(t.begin() + (long)v);
    if (v < t.size() && t[v].s == "(") {
      const size_t e = match_close(t, v);
      std::vector<std::string> names;
      for (size_t m = v + 1; m < e; ++m) {
        if (t[m].s == "mut") continue;
        if (t[m].kind == kIdent) names.push_back(t[m].s);
        else if (t[m].s != ",") throw bad("pattern in let");
      }
      if (e + 1 >= t.size() || t[e + 1].s != "=") throw bad("let pattern without value");
      size_t semi = e + 2;
      int depth = 0;
      for (; semi < t.size(); ++semi) {
        if (t[semi].s == "(" || t[semi].s == "[" || t[semi].s == "{") ++depth;
        if (t[semi].s == ")" || t[semi].s == "]" || t[semi].s == "}") --depth;
        if (t[semi].s == ";" && depth == 0) break;
      }
      if (semi >= t.size()) throw bad("let without ';'");
      const std::string tn = "__t" + std::to_string(tmp_id++);
      std::vector<Tok> repl = {{kIdent, "auto"}, {kIdent, tn}, {kPunct, "="}};
      repl.insert(repl.end(), t.begin() + (long)e + 2, t.begin() + (long)semi);
      repl.push_back({kPunct, ";"});
      for (size_t q = 0; q < names.size(); ++q) {
        if (names[q] == "_") continue;
        const std::vector<Tok> d = {{kIdent, "double"}, {kIdent, names[q]}, {kPunct, "="}, {kIdent, "pt_get"},
                                    {kPunct, "<"}, {kNumber, std::to_string(q)}, {kPunct, ">"}, {kPunct, "("},
                                    {kIdent, tn}, {kPunct, ")"}, {kPunct, ";"}};
        repl.insert(repl.end(), d.begin(), d.end());
      }
      t.erase(t.begin() + (long)k, t.begin() + (long)semi + 1);
      t.insert(t.begin() + (long)k, repl.begin(), repl.end());
      k += repl.size() - 1;
      continue;
    }
    if (v >= t.size() || t[v].kind != kIdent) throw bad("pattern in let");
    if (v + 1 < t.size() && t[v + 1].s == ":") {
      if (v + 2 >= t.size() || (t[v + 2].s != "f64" && t[v + 2].s != "f32")) throw bad("non-float let binding");
      t.erase(t.begin() + (long)v + 1, t.begin() + (long)v + 3);
    }
    // a value that is a tuple keeps its type: a tuple expression, a helper
    // returning one, or an if expression (whose branches decide)
    const bool tuple = v + 2 < t.size() && t[v + 1].s == "=" &&
                       (t[v + 2].s == "pt_make" || t[v + 2].s == "momentum" || t[v + 2].s == "contact" ||
                        (t[v + 2].s == "(" && v + 3 < t.size() && t[v + 3].s == "["));
    t[k].s = tuple ? "auto" : "double";
  }
  // if statements
  for (size_t k = 0; k < t.size(); ++k) {
    if (t[k].s != "if") continue;
    size_t b = k + 1;
    int depth = 0;
    for (; b < t.size(); ++b) {
      if (t[b].s == "(" || t[b].s == "[") ++depth;
      if (t[b].s == ")" || t[b].s == "]") --depth;
      if (t[b].s == "{" && depth == 0) break;
    }
    if (b >= t.size()) throw bad("if without block");
    t.insert(t.begin() + (long)b, {kPunct, ")"});
    t.insert(t.begin() + (long)k + 1, {kPunct, "("});
    ++k;
  }
  // helpers that read the descriptor constants get them as a last argument
  for (size_t k = 0; k + 1 < t.size(); ++k) {
    if ((t[k].s == "momentum" || t[k].s == "contact") && t[k + 1].s == "(") {
      const size_t e = match_close(t, k + 1);
      t.insert(t.begin() + (long)e, {kIdent, "prestige_k"});
      t.insert(t.begin() + (long)e, {kPunct, ","});
    }
  }
  // a block's last expression may omit its ';'
  for (size_t k = 1; k < t.size(); ++k) {
    if (t[k].s != "}") continue;
    const std::string& p = t[k - 1].s;
    if (p != ";" && p != "{" && p != "}") {
      t.insert(t.begin() + (long)k, {kPunct, ";"});
      ++k;
    }
  }
  return t;
}

// Field-access and name mapping of one body inside a generated pass.
struct PassNames {
  std::set<std::string> split;  // fields read at [j] and written in this pass
  bool slot_j = false;          // [j] reads come from slot-order copies g_<name>[js]
};

std::string emit_body(const Eq& e, const PassNames& pn, std::set<std::string>* consts_used) {
  auto bad = [&](const std::string& what) {
    return GenError{PRESTIGE_E_UNSUPPORTED, "equation '" + e.name + "': " + what};
  };
  std::vector<Tok> t = rewrite(tokenize(e.body), e.name);
  std::set<std::string> locals;
  for (size_t k = 0; k + 1 < t.size(); ++k)
    if ((t[k].s == "double" || t[k].s == "auto") && t[k + 1].kind == kIdent) locals.insert(t[k + 1].s);
  std::string out;
  for (size_t k = 0; k < t.size(); ++k) {
    const Tok& tk = t[k];
    std::string piece = tk.s;
    if (tk.kind == kIdent) {
      const size_t c = index_close(t, k);
      if (c) {
        if (c != k + 3 || (t[k + 2].s != "i" && t[k + 2].s != "j"))
          throw bad("field '" + tk.s + "' indexed by something other than i or j");
        const bool at_i = t[k + 2].s == "i";
        const bool write = c + 1 < t.size() && is_assign_op(t[c + 1].s);
        if (write && !at_i) throw bad("writes '" + tk.s + "[j]' (gather passes write through [i] only)");
        if (is_pos(tk.s)) {
          if (write) throw bad("writes the position '" + tk.s + "' that drives the neighbour search");
          piece = std::string("(double)p") + (at_i ? "i." : "j.") + tk.s;
        } else if (at_i) {
          piece = tk.s + "_i";
        } else {
          piece = pn.slot_j ? "g_" + tk.s + "[js]" : "f_" + tk.s + (pn.split.count(tk.s) ? "_in" : "") + "[jd]";
        }
        out += piece + " ";
        k = c;
        continue;
      }
      if (tk.s == "i") piece = "me";
      else if (tk.s == "j") piece = "jd";
      else if (const_names().count(tk.s)) {
        if (consts_used) consts_used->insert(tk.s);
      } else if (tk.s == "double" || tk.s == "auto" || tk.s == "return" || tk.s == "prestige_k" ||
                 tk.s == "if" || tk.s == "else" || tk.s == "true" || tk.s == "false" || locals.count(tk.s) ||
                 helper_names().count(tk.s)) {
      } else {
        bool fn = false;
        for (const auto& kv : fn_map())
          if (kv.second == tk.s) fn = true;
        if (!fn) throw bad("unknown name '" + tk.s + "'");
      }
    }
    out += piece + " ";
  }
  return out;
}

bool has_j(const Eq& e, const std::vector<Tok>& t) {
  for (const auto& r : e.reads)
    if (r.second & PRESTIGE_ROLE_J) return true;
  for (const auto& tk : t)
    if (tk.kind == kIdent && tk.s == "j") return true;
  return false;
}

// Pass index of every equation of the fused list (the schedule rule above).
std::vector<uint32_t> schedule(const int32_t* ids, uint32_t n) {
  std::vector<uint32_t> pass(n);
  uint32_t p = 0;
  std::set<std::string> written, written_ionly;
  for (uint32_t k = 0; k < n; ++k) {
    const Eq& e = registry()[(size_t)ids[k]];
    const bool jb = has_j(e, tokenize(e.body));
    bool hazard = false;
    for (const auto& rd : e.reads) {
      if ((rd.second & PRESTIGE_ROLE_J) && written.count(rd.first)) hazard = true;
      if (jb && written_ionly.count(rd.first)) hazard = true;
    }
    if (hazard) {
      ++p;
      written.clear();
      written_ionly.clear();
    }
    for (const auto& w : e.writes) {
      written.insert(w);
      if (!jb) written_ionly.insert(w);
    }
    pass[k] = p;
  }
  return pass;
}

const char* kPreamble = R"(// Device helpers available to equation bodies.
// pg_sqrt: correctly rounded square root of a positive normal x, the hardware
// reciprocal square root refined by Newton steps with a final remainder
// correction (Markstein 2000); sqrt() adds exponent scaling for denormal and
// special arguments, which pair distances never are.
__device__ inline double pg_sqrt(double x) {
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
// kernel_w: cubic spline W(r, h) = sigma f(r / h), sigma = 1 / (pi h^3),
// support 2h (Monaghan & Lattanzio 1985).  Both pieces are evaluated and
// selected (lanes of a wavefront straddle r = h); 1 / h and sigma depend on h
// only and leave the pair loop.
__device__ inline double kernel_w(double dx, double dy, double dz, double h) {
  // |x|^2 as the cell walk's exact test forms it, so the two share one evaluation
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double inv_h = 1.0 / h;
  const double s = 1.0 / (3.141592653589793 * h * h * h);
  const double q = r2 > 0.0 ? pg_sqrt(r2) * inv_h : 0.0;
  const double inner = __builtin_fma(q * q, __builtin_fma(0.75, q, -1.5), 1.0);
  const double t = 2.0 - q;
  const double outer = 0.25 * t * t * t;
  return s * (q < 1.0 ? inner : (q < 2.0 ? outer : 0.0));
}
// tait: p = max(rho0 c0^2 / 7 ((rho / rho0)^7 - 1), 0)
__device__ inline double tait(double rho, double rho0, double c0) {
  const double r = rho / rho0, r2 = r * r, r4 = r2 * r2;
  const double p = rho0 * c0 * c0 / 7.0 * (r4 * r2 * r - 1.0);
  return p > 0.0 ? p : 0.0;
}
struct prestige_consts {
  double h, dt, rho0, c0, alpha, kn, kt, mu, gx, gy, gz;
  double two_zeta, neg_inv_kt;   // DEM: 2 zeta from the restitution, -1 / kt
};
// Cell-loop passes run one wavefront per block over 64 consecutive slots.  Per row
// run the wavefront stages the union of its lanes' candidate ranges in LDS as FP32
// offsets from its first slot in units of the cutoff (|q|^2 in .w); each lane scans
// its own range into words of 32 candidates with the dot form q.w - 2 s.q < t, up
// to 128 candidates at a time, and walks their set bits in ascending slot order, two
// neighbours per trip, applying the exact FP64 test before the bodies.  The threshold t carries the rounding margin of the FP32 form, so the
// scan keeps a superset of the exact neighbours and the bodies see exactly the
// pairs and the order of the plain loop.
__device__ inline unsigned pg_wave_min(unsigned v) {
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned t = __shfl_xor(v, d, 64);
    v = t < v ? t : v;
  }
  return (unsigned)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ inline unsigned pg_wave_max(unsigned v) {
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned t = __shfl_xor(v, d, 64);
    v = t > v ? t : v;
  }
  return (unsigned)__builtin_amdgcn_readfirstlane((int)v);
}
// Tuples of doubles: (a, b, ..) in a body is pt_make(a, b, ..), t.k is pt_get<k>(t).
template <int N> struct pt_tuple { double v[N]; };
template <class... A> __device__ inline pt_tuple<(int)sizeof...(A)> pt_make(A... a) {
  return pt_tuple<(int)sizeof...(A)>{{(double)a...}};
}
template <int K, int N> __device__ inline double pt_get(const pt_tuple<N>& t) { return t.v[K]; }
// momentum: acceleration of i from j, -m_j sigma / h (P_i + P_j + Pi_ij) f'(q) / r x_ij,
// for 0 < r < 2h (else 0), with P = p / rho^2 summed by the caller (psum), rho_i + rho_j
// (rsum) and Monaghan's Pi_ij = -alpha c0 h (v.r) / ((r^2 + 0.01 h^2) rhobar) for
// approaching pairs, over one common denominator; the WCSPH force pass's operation order.
__device__ inline pt_tuple<3> momentum(double dx, double dy, double dz, double dvx, double dvy, double dvz,
                                       double psum, double rsum, double mj, const prestige_consts& k) {
  const double inv_h = 1.0 / k.h;
  const double H2 = (2.0 * k.h) * (2.0 * k.h);
  const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
  if (!(r2 < H2 && r2 > 0.0)) return pt_tuple<3>{{0.0, 0.0, 0.0}};
  const double eps2 = (0.01 * k.h) * k.h;
  const double rr = pg_sqrt(r2);
  const double q = rr * inv_h;
  double dw;
  if (q < 1.0) {
    dw = q * __builtin_fma(2.25, q, -3.0);
  } else {
    const double t = 2.0 - q;
    dw = -0.75 * t * t;
  }
  const double vdotr = __builtin_fma(dvx, dx, __builtin_fma(dvy, dy, dvz * dz));
  const double ach = (k.alpha * k.c0) * k.h;
  double num, den;
  if (vdotr < 0.0) {
    const double drb = (r2 + eps2) * (0.5 * rsum);
    num = __builtin_fma(psum, drb, -(ach * vdotr)) * dw;
    den = drb * rr;
  } else {
    num = psum * dw;
    den = rr;
  }
  const double sc = num / den;
  const double sigma = 1.0 / (((3.141592653589793 * k.h) * k.h) * k.h);
  const double msh = (mj * sigma) * inv_h;
  return pt_tuple<3>{{-(msh * (sc * dx)), -(msh * (sc * dy)), -(msh * (sc * dz))}};
}
// contact: force and torque on sphere i from sphere j (linear spring-dashpot normal
// force, tangential spring from a fresh history advanced by v_t dt and Coulomb-capped,
// torque (-r_i n) x F_t), zero when the spheres do not overlap; the DEM contact pass's
// operation order.  Returns (Fx, Fy, Fz, Tx, Ty, Tz).
__device__ inline pt_tuple<6> contact(double dx, double dy, double dz, double ri, double rj, double dvx,
                                      double dvy, double dvz, double wxi, double wyi, double wzi, double wxj,
                                      double wyj, double wzj, double mi, double mj, const prestige_consts& k) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double R = ri + rj;
  if (!(r2 < R * R && r2 > 0.0)) return pt_tuple<6>{{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}};
  const double d = pg_sqrt(r2);
  const double invd = 1.0 / d;
  const double nx = dx * invd, ny = dy * invd, nz = dz * invd;
  const double delta = R - d;
  const double wsx = ri * wxi + rj * wxj, wsy = ri * wyi + rj * wyj, wsz = ri * wzi + rj * wzj;
  const double cx = wsy * nz - wsz * ny, cy = wsz * nx - wsx * nz, cz = wsx * ny - wsy * nx;
  const double vrx = dvx - cx, vry = dvy - cy, vrz = dvz - cz;
  const double vn = vrx * nx + vry * ny + vrz * nz;
  const double vtx = vrx - vn * nx, vty = vry - vn * ny, vtz = vrz - vn * nz;
  const double meff = mi * mj / (mi + mj);
  const double cn = k.two_zeta * sqrt(meff * k.kn);
  const double fn = k.kn * delta - cn * vn;
  const double xo = 0.0;
  const double xdn = xo * nx + xo * ny + xo * nz;
  const double xpx = xo - xdn * nx, xpy = xo - xdn * ny, xpz = xo - xdn * nz;
  double xnx = xpx + vtx * k.dt, xny = xpy + vty * k.dt, xnz = xpz + vtz * k.dt;
  double ftx = -k.kt * xnx, fty = -k.kt * xny, ftz = -k.kt * xnz;
  const double ftn = sqrt(ftx * ftx + fty * fty + ftz * ftz);
  const double fcap = k.mu * (fn < 0.0 ? -fn : fn);
  if (ftn > fcap) {
    const double sc = fcap / ftn;
    ftx = ftx * sc;
    fty = fty * sc;
    ftz = ftz * sc;
  }
  const double tx = ny * ftz - nz * fty, ty = nz * ftx - nx * ftz, tz = nx * fty - ny * ftx;
  return pt_tuple<6>{{fn * nx + ftx, fn * ny + fty, fn * nz + ftz, -(ri * tx), -(ri * ty), -(ri * tz)}};
}
)";

GenProgram build_program(const int32_t* ids, uint32_t n, int32_t loop, bool f32_pos) {
  if (!valid_ids(ids, n)) throw GenError{PRESTIGE_E_INVALID, "unknown equation id"};
  if (loop != PRESTIGE_LOOP_ALLPAIRS && loop != PRESTIGE_LOOP_CELLS) throw GenError{PRESTIGE_E_INVALID, "unknown loop"};
  const std::vector<uint32_t> pass = schedule(ids, n);
  const uint32_t npass = n ? pass.back() + 1 : 0;
  GenProgram prog;
  std::string& s = prog.source;
  s += "// Gather passes generated by prestige_generate_hip from the fused equations";
  for (uint32_t k = 0; k < n; ++k) s += std::string(k ? ", " : " ") + registry()[(size_t)ids[k]].name;
  s += ".\n// Loop form: ";
  s += loop == PRESTIGE_LOOP_CELLS ? "cell iterator (slot i over the nine row runs of its cell, |x_i - x_j|^2 < c2)"
                                   : "all pairs (id i, j over ids 0..n)";
  s += ".\n// Fields are FP64 arrays in particle-id order; [i] fields live in registers for the\n"
       "// whole j loop and are stored once at the end.\n";
  s += "#if !defined(__HIPCC_RTC__)\n#include <hip/hip_runtime.h>\n#endif\n";
  s += std::string("typedef ") + (f32_pos ? "float4" : "double4") + " prestige_pos;\n";
  s += kPreamble;
  for (uint32_t p = 0; p < npass; ++p) {
    std::vector<const Eq*> jb, ib;
    std::map<std::string, GenField> fields;  // ordered by name
    std::set<std::string> at_i, read_j, written;
    for (uint32_t k = 0; k < n; ++k) {
      if (pass[k] != p) continue;
      const Eq& e = registry()[(size_t)ids[k]];
      (has_j(e, tokenize(e.body)) ? jb : ib).push_back(&e);
      for (const auto& rd : e.reads) {
        if (is_pos(rd.first)) continue;
        GenField& f = fields[rd.first];
        f.name = rd.first;
        f.read = true;
        if (rd.second & PRESTIGE_ROLE_I) at_i.insert(rd.first);
        if (rd.second & PRESTIGE_ROLE_J) read_j.insert(rd.first);
      }
      for (const auto& w : e.writes) {
        if (is_pos(w)) continue;
        fields[w].name = w;
        fields[w].written = true;
        written.insert(w);
        at_i.insert(w);
      }
    }
    PassNames pn;
    pn.slot_j = loop == PRESTIGE_LOOP_CELLS;
    for (auto& kv : fields) {
      kv.second.split = kv.second.written && read_j.count(kv.first);
      if (kv.second.split) pn.split.insert(kv.first);
    }
    std::set<std::string> consts;
    std::string jcode, icode;
    for (const Eq* e : jb) jcode += "      { " + emit_body(*e, pn, &consts) + "}\n";
    for (const Eq* e : ib) icode += "  { " + emit_body(*e, pn, &consts) + "}\n";
    GenPass gp;
    // kernel name: the pass index and its equations (identifier characters only)
    gp.kernel = "prestige_pass_" + std::to_string(p);
    for (const std::vector<const Eq*>* v : {&jb, &ib})
      for (const Eq* e : *v) {
        std::string nm;
        for (char ch : e->name) nm += std::isalnum((unsigned char)ch) ? ch : '_';
        if (gp.kernel.size() + nm.size() < 96) gp.kernel += "_" + nm;
      }
    std::string params;
    for (auto& kv : fields) {
      const GenField& f = kv.second;
      gp.fields.push_back(f);
      if (f.split)
        params += ",\n    const double* __restrict__ f_" + f.name + "_in, double* __restrict__ f_" + f.name + "_out";
      else if (f.written)
        params += ",\n    double* __restrict__ f_" + f.name;
      else
        params += ",\n    const double* __restrict__ f_" + f.name;
    }
    if (pn.slot_j)
      for (const std::string& f : read_j) {
        params += ",\n    const double* __restrict__ g_" + f;
        gp.slot_fields.push_back(fields[f]);
      }
    s += "\n// pass " + std::to_string(p) + ": fields";
    for (auto& kv : fields) s += " " + kv.first + (kv.second.split ? "(in,out)" : "");
    s += "\n";
    if (loop == PRESTIGE_LOOP_CELLS)
      // 8 wavefronts per SIMD (64 VGPRs): the pair loop is latency-bound
      s += "extern \"C\" __global__ void __launch_bounds__(64, 8) " + gp.kernel +
           "(\n    unsigned n, const prestige_pos* __restrict__ pos, const unsigned* __restrict__ ids,\n"
           "    const unsigned* __restrict__ skeys, const unsigned* __restrict__ lb, int nx, int ny, int nz,\n"
           "    double c2, double inv_cut, float thr, prestige_consts prestige_k" + params + ") {\n";
    else
      s += "extern \"C\" __global__ void __launch_bounds__(256) " + gp.kernel +
           "(\n    unsigned n, const prestige_pos* __restrict__ pos, prestige_consts prestige_k" + params + ") {\n";
    s += "#pragma clang fp contract(off)\n";
    if (loop == PRESTIGE_LOOP_CELLS) {
      s += "  __shared__ float4 s_q[256 + 4];\n"
           "  const unsigned lane = threadIdx.x, wbase = blockIdx.x * 64u, i = wbase + lane;\n"
           "  const bool live = i < n;\n"
           "  const unsigned me = live ? ids[i] : 0u;\n"
           "  const prestige_pos pi = pos[live ? i : wbase];\n"
           "  const prestige_pos p0 = pos[wbase];\n"
           "  const double xr = (double)p0.x, yr = (double)p0.y, zr = (double)p0.z;\n"
           "  auto offset = [&](const prestige_pos& p) {\n"
           "    return make_float4((float)(((double)p.x - xr) * inv_cut), (float)(((double)p.y - yr) * inv_cut),\n"
           "                       (float)(((double)p.z - zr) * inv_cut), 0.0f);\n"
           "  };\n";
    } else {
      s += "  const unsigned me = blockIdx.x * blockDim.x + threadIdx.x;\n  if (me >= n) return;\n";
      s += "  const prestige_pos pi = pos[me];\n";
    }
    for (const auto& c : consts) s += "  const double " + c + " = prestige_k." + c + ";\n";
    for (const auto& f : at_i) {
      const bool sp = pn.split.count(f) > 0;
      s += "  double " + f + "_i = f_" + f + (sp ? "_in" : "") + "[me];\n";
    }
    if (loop == PRESTIGE_LOOP_CELLS && jb.empty()) s += "  if (!live) return;\n";
    if (!jb.empty()) {
      if (loop == PRESTIGE_LOOP_CELLS) {
        s += "  const unsigned key = live ? skeys[i] : 0u;\n"
             "  const int cx = (int)(key % (unsigned)nx), cy = (int)((key / (unsigned)nx) % (unsigned)ny),\n"
             "            cz = (int)(key / ((unsigned)nx * (unsigned)ny));\n"
             "  const int x0 = cx > 0 ? cx - 1 : 0, x1 = cx < nx - 1 ? cx + 1 : nx - 1;\n"
             "  const float4 so = offset(pi);\n"
             "  const float ax_ = -2.0f * so.x, ay_ = -2.0f * so.y, az_ = -2.0f * so.z;\n"
             "  const float ss = __builtin_fmaf(so.x, so.x, __builtin_fmaf(so.y, so.y, so.z * so.z));\n"
             "  const unsigned ss_max = pg_wave_max(live ? __float_as_uint(ss) : 0u);\n"
             "  for (int dz = -1; dz <= 1; ++dz) {\n"
             "    for (int dy = -1; dy <= 1; ++dy) {\n"
             "      const int zz = cz + dz, yy = cy + dy;\n"
             "      unsigned j0 = 0u, len = 0u;\n"
             "      if (live && zz >= 0 && zz < nz && yy >= 0 && yy < ny) {\n"
             "        const unsigned row = ((unsigned)zz * (unsigned)ny + (unsigned)yy) * (unsigned)nx;\n"
             "        j0 = lb[row + (unsigned)x0];\n"
             "        len = lb[row + (unsigned)x1 + 1] - j0;\n"
             "      }\n"
             "      const unsigned base = pg_wave_min(len ? j0 : 0xFFFFFFFFu);\n"
             "      const unsigned top = pg_wave_max(len ? j0 + len : 0u);\n"
             "      if (top <= base) continue;\n"
             "      const bool staged = top - base <= 256u;\n"
             "      float t = 0.0f;\n"
             "      if (staged) {\n"
             "        float mq = 0.0f;\n"
             "        for (unsigned kk = lane; kk < top - base; kk += 64u) {\n"
             "          float4 q = offset(pos[base + kk]);\n"
             "          q.w = __builtin_fmaf(q.x, q.x, __builtin_fmaf(q.y, q.y, q.z * q.z));\n"
             "          s_q[kk] = q;\n"
             "          mq = q.w > mq ? q.w : mq;\n"
             "        }\n"
             "        const float m2 = fmaxf(__uint_as_float(ss_max), __uint_as_float(pg_wave_max(__float_as_uint(mq))));\n"
             "        t = thr + 0x1.0p-19f * (m2 + __builtin_sqrtf(m2) + 1.0f) - ss;\n"
             "        __syncthreads();\n"
             "      }\n"
             "      // up to 128 candidates at a time: four scan words, then one walk of their set bits\n"
             "      for (unsigned c0 = 0u; c0 < len; c0 += 128u) {\n"
             "        const unsigned clen = len - c0 < 128u ? len - c0 : 128u;\n"
             "        unsigned wd[4] = {0u, 0u, 0u, 0u};\n"
             "#pragma unroll\n"
             "        for (unsigned w = 0u; w < 4u; ++w) {\n"
             "          if (32u * w >= clen) break;\n"
             "          const unsigned cnt = clen - 32u * w < 32u ? clen - 32u * w : 32u;\n"
             "          unsigned word = 0u;\n"
             "          // four candidates per trip; those past cnt (stage padding, clamped slots) are masked off\n"
             "          for (unsigned b = 0u; b < cnt; b += 4u) {\n"
             "            unsigned hb = 0u;\n"
             "#pragma unroll\n"
             "            for (unsigned u = 0u; u < 4u; ++u) {\n"
             "              const unsigned jc = j0 + c0 + 32u * w + b + u;\n"
             "              bool hit;\n"
             "              if (staged) {\n"
             "                const float4 q = s_q[jc - base];\n"
             "                hit = __builtin_fmaf(ax_, q.x, __builtin_fmaf(ay_, q.y, __builtin_fmaf(az_, q.z, q.w))) < t;\n"
             "              } else {\n"
             "                const float4 q = offset(pos[jc < n ? jc : n - 1u]);\n"
             "                const float ddx = so.x - q.x, ddy = so.y - q.y, ddz = so.z - q.z;\n"
             "                hit = __builtin_fmaf(ddx, ddx, __builtin_fmaf(ddy, ddy, ddz * ddz)) < thr;\n"
             "              }\n"
             "              hb |= (hit ? 1u : 0u) << u;\n"
             "            }\n"
             "            word |= hb << b;\n"
             "          }\n"
             "          wd[w] = cnt < 32u ? word & ((1u << cnt) - 1u) : word;\n"
             "        }\n"
             "        unsigned long long m0 = (unsigned long long)wd[0] | ((unsigned long long)wd[1] << 32);\n"
             "        unsigned long long m1 = (unsigned long long)wd[2] | ((unsigned long long)wd[3] << 32);\n"
             "        auto pop = [&]() -> unsigned {\n"
             "          if (m0) {\n"
             "            const unsigned b = (unsigned)__builtin_ctzll(m0);\n"
             "            m0 &= m0 - 1ull;\n"
             "            return b;\n"
             "          }\n"
             "          const unsigned b = 64u + (unsigned)__builtin_ctzll(m1);\n"
             "          m1 &= m1 - 1ull;\n"
             "          return b;\n"
             "        };\n"
             "        while (m0 | m1) {\n"
             "          // two neighbours per trip, both positions loaded before either body runs\n"
             "          unsigned jsv[2];\n"
             "          jsv[0] = j0 + c0 + pop();\n"
             "          const bool two = (m0 | m1) != 0ull;\n"
             "          jsv[1] = two ? j0 + c0 + pop() : jsv[0];\n"
             "          const prestige_pos pjv[2] = {pos[jsv[0]], pos[jsv[1]]};\n"
             "#pragma unroll\n"
             "          for (int u_ = 0; u_ < 2; ++u_) {\n"
             "          if (u_ == 1 && !two) break;\n"
             "          const unsigned js = jsv[u_];\n"
             "          const prestige_pos pj = pjv[u_];\n"
             "          const double ex = (double)pi.x - (double)pj.x, ey = (double)pi.y - (double)pj.y,\n"
             "                       ez = (double)pi.z - (double)pj.z;\n"
             "          if (!(ex * ex + ey * ey + ez * ez < c2)) continue;\n"
             "          const unsigned jd = ids[js];\n";
        s += jcode;
        s += "          }\n        }\n      }\n      if (staged) __syncthreads();\n    }\n  }\n  if (!live) return;\n";
      } else {
        s += "  for (unsigned jd = 0; jd < n; ++jd) {\n      const prestige_pos pj = pos[jd];\n";
        s += jcode;
        s += "  }\n";
      }
    }
    s += icode;
    for (const auto& f : written) {
      const bool sp = pn.split.count(f) > 0;
      s += "  f_" + f + (sp ? "_out" : "") + "[me] = " + f + "_i;\n";
    }
    s += "}\n";
    prog.passes.push_back(gp);
  }
  return prog;
}

}  // namespace

GenProgram generate_hip_program(const int32_t* eq_ids, uint32_t n, int32_t loop, bool f32_pos) {
  std::lock_guard<std::mutex> lk(g_mutex);
  return build_program(eq_ids, n, loop, f32_pos);
}

}  // namespace prestige

using prestige::GenError;
using prestige::registry;

extern "C" {

prestige_status prestige_equation_describe(int32_t eq_id, char* name, size_t name_cap, char* reads,
                                           size_t reads_cap, char* writes, size_t writes_cap,
                                           uint8_t* read_roles, size_t roles_cap, char* body,
                                           size_t body_cap) {
  std::lock_guard<std::mutex> lk(prestige::g_mutex);
  if (eq_id < 0 || eq_id >= (int32_t)registry().size()) return PRESTIGE_E_INVALID;
  const prestige::Eq& e = registry()[(size_t)eq_id];
  prestige::copy_out(e.name, name, name_cap);
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
  prestige::copy_out(r, reads, reads_cap);
  prestige::copy_out(w, writes, writes_cap);
  prestige::copy_out(e.body, body, body_cap);
  return PRESTIGE_OK;
}

prestige_status prestige_equation_define(const char* name, const char* body, int32_t* eq_id) {
  if (!name || !body || !eq_id || !prestige::is_ident(name)) {
    prestige::set_context_free_error("equation name must be an identifier; body and eq_id non-null");
    return PRESTIGE_E_INVALID;
  }
  try {
    prestige::Eq e;
    e.name = name;
    e.body = body;
    const auto t = prestige::tokenize(e.body);
    if (t.empty() || t.front().s != "{" || prestige::match_close(t, 0) != t.size() - 1) {
      prestige::set_context_free_error("an equation body is one block: { ... }");
      return PRESTIGE_E_INVALID;
    }
    prestige::infer(t, e);
    std::lock_guard<std::mutex> lk(prestige::g_mutex);
    registry().push_back(e);
    *eq_id = (int32_t)registry().size() - 1;
    return PRESTIGE_OK;
  } catch (const GenError& g) {
    prestige::set_context_free_error(g.msg);
    return g.code;
  }
}

prestige_status prestige_fuse_schedule(const int32_t* eq_ids, uint32_t n, uint32_t* pass_of, uint32_t* n_passes) {
  std::lock_guard<std::mutex> lk(prestige::g_mutex);
  if (!prestige::valid_ids(eq_ids, n) || (n && !pass_of) || !n_passes) {
    prestige::set_context_free_error("unknown equation id or null output");
    return PRESTIGE_E_INVALID;
  }
  try {
    const std::vector<uint32_t> p = prestige::schedule(eq_ids, n);
    for (uint32_t k = 0; k < n; ++k) pass_of[k] = p[k];
    *n_passes = n ? p.back() + 1 : 0;
    return PRESTIGE_OK;
  } catch (const GenError& g) {
    prestige::set_context_free_error(g.msg);
    return g.code;
  }
}

prestige_status prestige_generate_simple_cpu(const int32_t* eq_ids, uint32_t n, char* out, size_t cap, size_t* len) {
  std::lock_guard<std::mutex> lk(prestige::g_mutex);
  if (!prestige::valid_ids(eq_ids, n)) {
    prestige::set_context_free_error("unknown equation id");
    return PRESTIGE_E_INVALID;
  }
  // text layout of prestige/src/codegen/simple_cpu.rs:7-19
  std::string code;
  code += "for i in 0..n {\n";
  code += "    for j in 0..n {\n";
  for (uint32_t k = 0; k < n; ++k) {
    code += "        ";
    code += registry()[(size_t)eq_ids[k]].body;
    code += "\n";
  }
  code += "    }\n";
  code += "}\n";
  if (len) *len = code.size();
  prestige::copy_out(code, out, cap);
  return PRESTIGE_OK;
}

prestige_status prestige_generate_hip(const int32_t* eq_ids, uint32_t n, int32_t loop, char* out, size_t cap,
                                      size_t* len) {
  std::lock_guard<std::mutex> lk(prestige::g_mutex);
  try {
    const prestige::GenProgram p = prestige::build_program(eq_ids, n, loop, false);
    if (len) *len = p.source.size();
    prestige::copy_out(p.source, out, cap);
    return PRESTIGE_OK;
  } catch (const GenError& g) {
    prestige::set_context_free_error(g.msg);
    return g.code;
  }
}

}  // extern "C"
