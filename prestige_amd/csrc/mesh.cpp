// mesh.cpp — Gmsh surface meshes as fixed boundary particles (host only).
//
// Reads ASCII MSH 2.x and 4.x files: nodes plus the surface elements
// (3- and 6-node triangles, 4-, 8- and 9-node quadrangles; corner nodes only,
// quadrangles split along the diagonal from their first node).  Points,
// lines and volume elements are skipped.  The surface is sampled at spacing
// dx, each point once:
//   * planar patches: coplanar triangles joined across shared edges;
//   * feature lines: edges between two patches, or with one or more than two
//     triangles (open rims, non-manifold seams);
//   * corners: feature-line nodes where the line is not straight (any
//     number of feature edges but two, or two that are not collinear);
//   * each straight run of feature edges between two corners, at
//     n = ceil(L / dx) equal steps (n = L / dx when that is an integer to
//     1e-9), its interior points;
//   * each patch, the points of a square dx lattice in its plane that are
//     inside the patch and at least dx / 2 from its feature edges.  Axes:
//     e1 along the projection of the world axis least aligned with the
//     normal, e2 = n x e1; origin at the patch's corner with the smallest
//     node tag.  Coplanar patches of a box whose sides are multiples of dx
//     thus sample exactly the lattice points of its surface.
// Further layers l = 1 .. layers-1 resample the surface offset by l*dx:
// every node moves by side * l * dx * d_v, where d_v solves n_f . d_v = 1
// over the distinct planes f meeting at the node (minimum-norm least
// squares), which moves each of them by exactly l*dx where at most three
// independent planes meet.  Face normal: (b - a) x (c - a) over a triangle's
// nodes a, b, c.
#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "../../include/prestige_hip.h"
#include "codegen.hpp"

namespace {

struct V3 {
  double x, y, z;
};
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }
V3 unit(V3 a) {
  const double l = norm(a);
  return l > 0.0 ? (1.0 / l) * a : V3{0.0, 0.0, 0.0};
}

struct MeshError {
  prestige_status code;
  std::string msg;
};

struct Mesh {
  std::vector<long long> tag;  // node tag of each dense node index
  std::vector<V3> pos;
  std::vector<std::array<int, 3>> tris;
};

// Advance to the line after "$Name"; false at end of file.
bool seek_section(std::istream& in, const std::string& name) {
  std::string line;
  while (std::getline(in, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    if (line == "$" + name) return true;
  }
  return false;
}

// Nodes per element of the MSH element types; corners = the leading nodes
// that span the element (2 = triangle, 3 = quadrangle surface types).
int nodes_of(int type) {
  switch (type) {
    case 1: return 2;    // line
    case 2: return 3;    // triangle
    case 3: return 4;    // quadrangle
    case 4: return 4;    // tetrahedron
    case 5: return 8;    // hexahedron
    case 6: return 6;    // prism
    case 7: return 5;    // pyramid
    case 8: return 3;    // 3-node line
    case 9: return 6;    // 6-node triangle
    case 10: return 9;   // 9-node quadrangle
    case 11: return 10;  // 10-node tetrahedron
    case 15: return 1;   // point
    case 16: return 8;   // 8-node quadrangle
    default: return -1;
  }
}

void add_element(std::vector<std::array<long long, 3>>& tris, int type, const std::vector<long long>& nd) {
  if (type == 2 || type == 9) tris.push_back({nd[0], nd[1], nd[2]});
  if (type == 3 || type == 10 || type == 16) {
    tris.push_back({nd[0], nd[1], nd[2]});
    tris.push_back({nd[0], nd[2], nd[3]});
  }
}

Mesh read_msh(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw MeshError{PRESTIGE_E_INVALID, "cannot open '" + path + "'"};
  if (!seek_section(in, "MeshFormat")) throw MeshError{PRESTIGE_E_INVALID, "no $MeshFormat section"};
  double version = 0.0;
  int binary = 0, dsize = 0;
  in >> version >> binary >> dsize;
  if (!in) throw MeshError{PRESTIGE_E_INVALID, "malformed $MeshFormat section"};
  if (binary) throw MeshError{PRESTIGE_E_UNSUPPORTED, "binary MSH files are not read; save as ASCII"};
  const bool v4 = version >= 4.0 && version < 5.0;
  if (!v4 && !(version >= 2.0 && version < 3.0))
    throw MeshError{PRESTIGE_E_UNSUPPORTED, "MSH version " + std::to_string(version) + " (2.x and 4.x are read)"};

  std::map<long long, V3> nodes;
  in.clear();
  in.seekg(0);
  if (!seek_section(in, "Nodes")) throw MeshError{PRESTIGE_E_INVALID, "no $Nodes section"};
  if (v4) {
    long long nblocks = 0, nn, mn, mx;
    in >> nblocks >> nn >> mn >> mx;
    for (long long b = 0; in && b < nblocks; ++b) {
      int edim = 0, etag, param = 0;
      long long cnt = 0;
      in >> edim >> etag >> param >> cnt;
      if (!in || cnt < 0) break;
      std::vector<long long> tags((size_t)cnt);
      for (auto& t : tags) in >> t;
      for (long long k = 0; in && k < cnt; ++k) {
        V3 p;
        in >> p.x >> p.y >> p.z;
        double u;
        if (param)
          for (int q = 0; q < edim; ++q) in >> u;  // parametric coordinates
        nodes[tags[(size_t)k]] = p;
      }
    }
  } else {
    long long nn = 0;
    in >> nn;
    for (long long k = 0; in && k < nn; ++k) {
      long long id;
      V3 p;
      in >> id >> p.x >> p.y >> p.z;
      nodes[id] = p;
    }
  }
  if (!in) throw MeshError{PRESTIGE_E_INVALID, "malformed $Nodes section"};

  std::vector<std::array<long long, 3>> tris;
  in.clear();
  in.seekg(0);
  if (!seek_section(in, "Elements")) throw MeshError{PRESTIGE_E_INVALID, "no $Elements section"};
  auto element = [&](int type) {
    const int nv = nodes_of(type);
    if (nv < 0) throw MeshError{PRESTIGE_E_UNSUPPORTED, "element type " + std::to_string(type)};
    std::vector<long long> nd((size_t)nv);
    for (auto& v : nd) in >> v;
    add_element(tris, type, nd);
  };
  if (v4) {
    long long nblocks = 0, ne, mn, mx;
    in >> nblocks >> ne >> mn >> mx;
    for (long long b = 0; in && b < nblocks; ++b) {
      int edim, etag, type = 0;
      long long cnt = 0;
      in >> edim >> etag >> type >> cnt;
      for (long long k = 0; in && k < cnt; ++k) {
        long long tag;
        in >> tag;
        element(type);
      }
    }
  } else {
    long long ne = 0;
    in >> ne;
    for (long long k = 0; in && k < ne; ++k) {
      long long id, tag;
      int type = 0, ntags = 0;
      in >> id >> type >> ntags;
      for (int t = 0; t < ntags; ++t) in >> tag;
      element(type);
    }
  }
  if (!in) throw MeshError{PRESTIGE_E_INVALID, "malformed $Elements section"};

  Mesh m;
  std::map<long long, int> dense;
  auto index = [&](long long t) {
    auto it = dense.find(t);
    if (it != dense.end()) return it->second;
    auto nt = nodes.find(t);
    if (nt == nodes.end()) throw MeshError{PRESTIGE_E_INVALID, "element refers to missing node " + std::to_string(t)};
    const int k = (int)m.pos.size();
    dense.emplace(t, k);
    m.tag.push_back(t);
    m.pos.push_back(nt->second);
    return k;
  };
  for (const auto& t : tris) {
    std::array<int, 3> d{index(t[0]), index(t[1]), index(t[2])};
    if (norm(cross(m.pos[d[1]] - m.pos[d[0]], m.pos[d[2]] - m.pos[d[0]])) > 0.0) m.tris.push_back(d);
  }
  if (m.tris.empty()) throw MeshError{PRESTIGE_E_INVALID, "no triangle or quadrangle elements of nonzero area"};
  return m;
}

// Patches, feature lines and corners: fixed by the mesh, shared by every layer.
struct Topology {
  std::vector<int> patch_of;             // per triangle
  std::vector<std::vector<int>> patches; // triangles of each patch, ascending
  std::vector<std::vector<std::pair<int, int>>> patch_edges;  // feature edges bounding each patch
  std::vector<int> corners;              // node indices, by ascending tag
  std::vector<std::pair<int, int>> runs; // straight feature runs, corner to corner
  std::vector<int> anchor;               // lattice origin node of each patch
  std::vector<V3> offset;                // d_v per node
};

int find_root(std::vector<int>& p, int a) {
  while (p[a] != a) a = p[a] = p[p[a]];
  return a;
}

constexpr double kFlat = 1e-10;  // 1 - cos of the largest angle taken as flat or straight

// Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations: A = V diag(w) V^T,
// eigenvectors in the columns of V.
void sym_eig3(double A[3][3], double w[3], double V[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) V[i][j] = i == j ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    const double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
    if (off <= 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (A[p][q] == 0.0) continue;
        const double th = 0.5 * (A[q][q] - A[p][p]) / A[p][q];
        const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 3; ++k) {  // A <- A J
          const double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {  // A <- J^T A
          const double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 3; ++k) {  // V <- V J
          const double vkp = V[k][p], vkq = V[k][q];
          V[k][p] = c * vkp - s * vkq;
          V[k][q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < 3; ++i) w[i] = A[i][i];
}

// Displacement d with n . d = 1 for every unit normal n of the planes meeting
// at a node: the minimum-norm least-squares solution of the normal equations
// (sum n n^T) d = sum n, exact where at most three independent planes meet;
// |d| is capped at 5 (wedges sharper than about 23 degrees).
V3 plane_offset(const std::vector<V3>& ns) {
  if (ns.empty()) return {0.0, 0.0, 0.0};
  double A[3][3] = {}, b[3] = {};
  for (const V3& n : ns) {
    const double c[3] = {n.x, n.y, n.z};
    for (int i = 0; i < 3; ++i) {
      b[i] += c[i];
      for (int j = 0; j < 3; ++j) A[i][j] += c[i] * c[j];
    }
  }
  double w[3], V[3][3];
  sym_eig3(A, w, V);
  const double wmax = std::fmax(w[0], std::fmax(w[1], w[2]));
  double d[3] = {};
  for (int k = 0; k < 3; ++k) {
    if (!(w[k] > 1e-8 * wmax)) continue;
    const double proj = (V[0][k] * b[0] + V[1][k] * b[1] + V[2][k] * b[2]) / w[k];
    for (int i = 0; i < 3; ++i) d[i] += proj * V[i][k];
  }
  V3 r{d[0], d[1], d[2]};
  const double len = norm(r);
  return len > 5.0 ? (5.0 / len) * r : r;
}

Topology topology(const Mesh& m) {
  const size_t nt = m.tris.size(), nn = m.pos.size();
  std::vector<V3> fn(nt);
  for (size_t t = 0; t < nt; ++t) {
    const V3 a = m.pos[m.tris[t][0]], b = m.pos[m.tris[t][1]], c = m.pos[m.tris[t][2]];
    fn[t] = unit(cross(b - a, c - a));
  }
  std::map<std::pair<int, int>, std::vector<int>> edges;
  for (size_t t = 0; t < nt; ++t)
    for (int e = 0; e < 3; ++e) {
      int u = m.tris[t][e], v = m.tris[t][(e + 1) % 3];
      if (u > v) std::swap(u, v);
      edges[{u, v}].push_back((int)t);
    }
  std::vector<int> parent(nt);
  for (size_t t = 0; t < nt; ++t) parent[t] = (int)t;
  for (const auto& kv : edges)
    if (kv.second.size() == 2 && dot(fn[kv.second[0]], fn[kv.second[1]]) >= 1.0 - kFlat) {
      const int a = find_root(parent, kv.second[0]), b = find_root(parent, kv.second[1]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  Topology tp;
  tp.patch_of.assign(nt, -1);
  std::map<int, int> root_patch;
  for (size_t t = 0; t < nt; ++t) {
    const int r = find_root(parent, (int)t);
    auto it = root_patch.find(r);
    if (it == root_patch.end()) {
      it = root_patch.emplace(r, (int)tp.patches.size()).first;
      tp.patches.emplace_back();
    }
    tp.patch_of[t] = it->second;
    tp.patches[it->second].push_back((int)t);
  }
  tp.patch_edges.resize(tp.patches.size());
  std::vector<std::vector<int>> fadj(nn);  // feature-edge neighbours per node
  for (const auto& kv : edges) {
    const auto& ts = kv.second;
    const bool feature = ts.size() != 2 || tp.patch_of[ts[0]] != tp.patch_of[ts[1]];
    if (!feature) continue;
    fadj[kv.first.first].push_back(kv.first.second);
    fadj[kv.first.second].push_back(kv.first.first);
    std::set<int> ps;
    for (int t : ts) ps.insert(tp.patch_of[t]);
    for (int p : ps) tp.patch_edges[p].push_back(kv.first);
  }
  std::vector<char> corner(nn, 0);
  for (size_t v = 0; v < nn; ++v) {
    if (fadj[v].empty()) continue;
    if (fadj[v].size() != 2) {
      corner[v] = 1;
      continue;
    }
    const V3 a = unit(m.pos[fadj[v][0]] - m.pos[v]), b = unit(m.pos[fadj[v][1]] - m.pos[v]);
    if (dot(a, b) > -(1.0 - kFlat)) corner[v] = 1;
  }
  // straight runs from each corner; feature loops without a corner get one
  std::set<std::pair<int, int>> used;
  auto walk = [&](int c0) {
    for (int nb : fadj[c0]) {
      if (used.count({std::min(c0, nb), std::max(c0, nb)})) continue;
      int prev = c0, cur = nb;
      used.insert({std::min(prev, cur), std::max(prev, cur)});
      while (!corner[cur]) {
        const int nxt = fadj[cur][0] == prev ? fadj[cur][1] : fadj[cur][0];
        prev = cur;
        cur = nxt;
        used.insert({std::min(prev, cur), std::max(prev, cur)});
      }
      tp.runs.push_back({c0, cur});
    }
  };
  std::vector<int> by_tag(nn);
  for (size_t v = 0; v < nn; ++v) by_tag[v] = (int)v;
  std::sort(by_tag.begin(), by_tag.end(), [&](int a, int b) { return m.tag[a] < m.tag[b]; });
  for (int v : by_tag)
    if (corner[v]) walk(v);
  for (int v : by_tag)
    for (int nb : fadj[v])
      if (!used.count({std::min(v, nb), std::max(v, nb)})) {
        corner[v] = 1;
        walk(v);
      }
  for (int v : by_tag)
    if (corner[v]) tp.corners.push_back(v);
  // lattice origin: the patch's corner with the smallest tag, else its first node
  tp.anchor.resize(tp.patches.size());
  for (size_t p = 0; p < tp.patches.size(); ++p) {
    int best = -1;
    for (int t : tp.patches[p])
      for (int v : m.tris[t])
        if (corner[v] && (best < 0 || m.tag[v] < m.tag[best])) best = v;
    tp.anchor[p] = best >= 0 ? best : m.tris[tp.patches[p][0]][0];
  }
  // layer displacement per node: n_f . d = 1 over the distinct adjacent planes
  std::vector<std::vector<V3>> planes(nn);
  for (size_t t = 0; t < nt; ++t)
    for (int v : m.tris[t]) {
      bool seen = false;
      for (const V3& q : planes[v]) seen = seen || dot(q, fn[t]) >= 1.0 - kFlat;
      if (!seen) planes[v].push_back(fn[t]);
    }
  tp.offset.assign(nn, V3{0.0, 0.0, 0.0});
  for (size_t v = 0; v < nn; ++v) tp.offset[v] = plane_offset(planes[v]);
  return tp;
}

// Steps of a straight run of length len: len / dx when that is an integer to
// 1e-9 (so a side of k*dx gives k steps despite rounding), else ceil.
long long steps_for(double len, double dx) {
  const double q = len / dx, r = std::round(q);
  const long long s = std::fabs(q - r) <= 1e-9 * std::fmax(1.0, q) ? (long long)r : (long long)std::ceil(q);
  return std::max(1ll, s);
}

void sample_layer(const Mesh& m, const Topology& tp, const std::vector<V3>& P, double dx, std::vector<V3>& out) {
  for (int v : tp.corners) out.push_back(P[v]);
  for (const auto& r : tp.runs) {
    const V3 a = P[r.first], b = P[r.second];
    const long long ns = steps_for(norm(b - a), dx);
    for (long long k = 1; k < ns; ++k) out.push_back(a + ((double)k / (double)ns) * (b - a));
  }
  const double tol = 1e-9 * dx;
  for (size_t p = 0; p < tp.patches.size(); ++p) {
    const auto& tri0 = m.tris[tp.patches[p][0]];
    const V3 n = unit(cross(P[tri0[1]] - P[tri0[0]], P[tri0[2]] - P[tri0[0]]));
    int ax = 0;
    const double an[3] = {std::fabs(n.x), std::fabs(n.y), std::fabs(n.z)};
    for (int k = 1; k < 3; ++k)
      if (an[k] < an[ax]) ax = k;
    const V3 axis{ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
    const V3 e1 = unit(axis - dot(axis, n) * n), e2 = cross(n, e1);
    const V3 o = P[tp.anchor[p]];
    auto uv = [&](V3 q) { return std::make_pair(dot(q - o, e1), dot(q - o, e2)); };
    std::set<std::pair<long long, long long>> pts;
    for (int t : tp.patches[p]) {
      std::pair<double, double> c[3];
      for (int k = 0; k < 3; ++k) c[k] = uv(P[m.tris[t][k]]);
      const double area2 = (c[1].first - c[0].first) * (c[2].second - c[0].second) -
                           (c[2].first - c[0].first) * (c[1].second - c[0].second);
      const double sg = area2 >= 0.0 ? 1.0 : -1.0;
      double u0 = c[0].first, u1 = u0, w0 = c[0].second, w1 = w0;
      for (int k = 1; k < 3; ++k) {
        u0 = std::fmin(u0, c[k].first), u1 = std::fmax(u1, c[k].first);
        w0 = std::fmin(w0, c[k].second), w1 = std::fmax(w1, c[k].second);
      }
      for (long long i = (long long)std::ceil(u0 / dx - 1e-9); (double)i <= u1 / dx + 1e-9; ++i)
        for (long long j = (long long)std::ceil(w0 / dx - 1e-9); (double)j <= w1 / dx + 1e-9; ++j) {
          const double pu = (double)i * dx, pw = (double)j * dx;
          bool in = true;
          for (int k = 0; k < 3 && in; ++k) {
            const auto& s = c[k];
            const auto& e = c[(k + 1) % 3];
            const double eu = e.first - s.first, ew = e.second - s.second;
            const double d = sg * (eu * (pw - s.second) - ew * (pu - s.first)) / std::sqrt(eu * eu + ew * ew);
            in = d >= -tol;
          }
          if (in) pts.insert({i, j});
        }
    }
    const double keep = 0.5 * dx - tol;
    for (const auto& ij : pts) {
      const double pu = (double)ij.first * dx, pw = (double)ij.second * dx;
      bool ok = true;
      for (const auto& e : tp.patch_edges[p]) {
        const auto a = uv(P[e.first]), b = uv(P[e.second]);
        const double eu = b.first - a.first, ew = b.second - a.second, l2 = eu * eu + ew * ew;
        double s = l2 > 0.0 ? ((pu - a.first) * eu + (pw - a.second) * ew) / l2 : 0.0;
        s = std::fmin(1.0, std::fmax(0.0, s));
        const double du = pu - (a.first + s * eu), dw = pw - (a.second + s * ew);
        if (du * du + dw * dw < keep * keep) {
          ok = false;
          break;
        }
      }
      if (ok) out.push_back(o + pu * e1 + pw * e2);
    }
  }
}

}  // namespace

extern "C" prestige_status prestige_mesh_boundary(const char* path, double dx, uint32_t layers, int32_t side,
                                                  double* x, double* y, double* z, uint64_t cap, uint64_t* n) {
  if (!path || !n || !(dx > 0.0) || !std::isfinite(dx) || layers == 0 || (side != 1 && side != -1)) {
    prestige::set_context_free_error("mesh_boundary: null path or count, dx must be finite and > 0, "
                                     "layers >= 1, side +1 or -1");
    return PRESTIGE_E_INVALID;
  }
  try {
    const Mesh m = read_msh(path);
    const Topology tp = topology(m);
    std::vector<V3> pts, P(m.pos.size());
    for (uint32_t l = 0; l < layers; ++l) {
      const double s = (double)side * (double)l * dx;
      for (size_t v = 0; v < P.size(); ++v) P[v] = m.pos[v] + s * tp.offset[v];
      sample_layer(m, tp, P, dx, pts);
    }
    *n = pts.size();
    if (!x && !y && !z) return PRESTIGE_OK;
    if (!x || !y || !z || cap < pts.size()) {
      prestige::set_context_free_error("mesh_boundary: output arrays missing or smaller than the sample count");
      return PRESTIGE_E_INVALID;
    }
    for (size_t k = 0; k < pts.size(); ++k) {
      x[k] = pts[k].x;
      y[k] = pts[k].y;
      z[k] = pts[k].z;
    }
    return PRESTIGE_OK;
  } catch (const MeshError& e) {
    prestige::set_context_free_error("mesh_boundary: " + e.msg);
    return e.code;
  } catch (const std::exception& e) {
    prestige::set_context_free_error(std::string("mesh_boundary: ") + e.what());
    return PRESTIGE_E_INVALID;
  }
}
