// Mesh reader/generator and face connectivity.
#include "geometry.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cctype>
#include <fstream>
#include <memory>
#include <sstream>
#include <unordered_map>

namespace hifr_host {

namespace {

// Gambit brick / quad node order -> internal shape-node position
// (src/geometry.cpp:1800-1855)
const int kHexPos[8] = {0, 2, 4, 6, 1, 3, 5, 7};
const int kQuadPos[4] = {0, 1, 3, 2};
const int kTetPos[4] = {0, 1, 2, 3};

}  // namespace

int bc_number(std::string name) {
  std::transform(name.begin(), name.end(), name.begin(), ::tolower);
  static const char* names[] = {"sub_in_simp", "sub_out_simp", "sub_in_char", "sub_out_char", "sup_in",
                                "sup_out", "slip_wall", "", "cyclic", "", "isotherm_fix", "adiabat_fix",
                                "isotherm_move", "adiabat_move", "char", "slip_wall_dual"};
  for (int i = 0; i < 16; i++)
    if (names[i][0] && name == names[i]) return i + 1;
  if (name == "ad_wall") return 50;
  return -1;
}

std::string read_gambit(const std::string& path, mesh& m) {
  std::ifstream f(path);
  if (!f) return "Unable to open mesh file " + path;
  std::string line;
  for (int i = 0; i < 6; i++) std::getline(f, line);
  int nn = 0, ne = 0, ngrp = 0, nbc = 0, nd = 0;
  f >> nn >> ne >> ngrp >> nbc >> nd;
  if (nd != 2 && nd != 3) return "Invalid mesh dimensionality";
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, line);
  m.n_dims = nd;
  m.n_verts = nn;
  m.xv.assign((size_t)nn * nd, 0.0);
  for (int i = 0; i < nn; i++) {
    std::getline(f, line);
    std::istringstream ss(line);
    int id;
    ss >> id;
    for (int k = 0; k < nd; k++) ss >> m.xv[(size_t)(id - 1) * nd + k];
  }
  std::getline(f, line);  // ENDOFSECTION
  std::getline(f, line);  // ELEMENTS/CELLS
  m.n_eles = ne;
  m.ele_type = 0;
  for (int e = 0; e < ne; e++) {
    int id, gtype, nv;
    f >> id >> gtype >> nv;
    // Gambit element types: 2 quadrilateral, 4 brick, 6 tetrahedron
    int et = gtype == 4 ? 4 : (gtype == 2 ? 1 : (gtype == 6 ? 2 : -1));
    if (et < 0 || (et == 4 && nv != 8) || (et != 4 && nv != 4))
      return "only linear quads, linear tets and 8-node hexes are supported";
    if (m.ele_type == 0) {
      m.ele_type = et;
      m.n_spts = nv;
      m.c2v.assign((size_t)ne * nv, 0);
    } else if (m.ele_type != et) {
      return "mixed element types are not supported";
    }
    const int* pos = et == 4 ? kHexPos : (et == 1 ? kQuadPos : kTetPos);
    for (int k = 0; k < nv; k++) {
      int v;
      f >> v;
      m.c2v[(size_t)(id - 1) * nv + pos[k]] = v - 1;
    }
  }
  if (!f) return "error reading Gambit elements";
  // element groups, then boundary groups (src/geometry.cpp:1186-1290)
  // Gambit face number -> internal face (hexes, tets; quads: face - 1)
  static const int kHexFace[7] = {-1, 0, 3, 5, 1, 4, 2};
  static const int kTetFace[5] = {-1, 3, 2, 0, 1};
  m.bf_ele.clear();
  m.bf_face.clear();
  m.bf_type.clear();
  while (std::getline(f, line)) {
    if (line.find("BOUNDARY CONDITIONS") == std::string::npos) continue;
    std::getline(f, line);
    std::istringstream hs(line);
    std::string name;
    int itype = 0, nent = 0;
    hs >> name >> itype >> nent;
    if (!hs) return "error reading a Gambit boundary group header";
    const int type = bc_number(name);
    if (type < 0) return "boundary condition '" + name + "' not recognized";
    for (int q = 0; q < nent; q++) {
      int ge, gt, gf;
      f >> ge >> gt >> gf;
      if (!f || ge < 1 || ge > ne) return "error reading Gambit boundary faces of group " + name;
      if (type == 9) continue;  // cyclic: matched by translation
      int lf = gf - 1;
      if (m.ele_type == 4) lf = gf >= 1 && gf <= 6 ? kHexFace[gf] : -1;
      if (m.ele_type == 2) lf = gf >= 1 && gf <= 4 ? kTetFace[gf] : -1;
      if (lf < 0 || lf >= (m.ele_type == 4 ? 6 : 4)) return "invalid Gambit face number in group " + name;
      m.bf_ele.push_back(ge - 1);
      m.bf_face.push_back(lf);
      m.bf_type.push_back(type);
    }
    std::getline(f, line);
  }
  return "";
}

namespace {

struct FaceKey {
  std::array<int, 4> v;
  bool operator==(const FaceKey& o) const { return v == o.v; }
};
struct FaceKeyHash {
  size_t operator()(const FaceKey& k) const {
    size_t h = 1469598103934665603ull;
    for (int x : k.v) h = (h ^ (size_t)(unsigned)x) * 1099511628211ull;
    return h;
  }
};

}  // namespace

namespace {

// Gmsh node order of a linear element -> internal shape-node position
// (src/geometry.cpp:1987-2050: quads and hexes swap the last two nodes of each
// quadrilateral, tets keep theirs)
const int kGmshQuadPos[4] = {0, 1, 3, 2};
const int kGmshHexPos[8] = {0, 1, 3, 2, 4, 5, 7, 6};

// nodes of a linear Gmsh element type (2.x numbering), 0 if not read here
int gmsh_nodes(int etype) {
  switch (etype) {
    case 1: return 2;  // line
    case 2: return 3;  // triangle
    case 3: return 4;  // quadrangle
    case 4: return 4;  // tetrahedron
    case 5: return 8;  // hexahedron
    default: return 0;
  }
}

// Physical name of a $PhysicalNames line as the reference reads it: the first
// word after the dimension and tag, quotes removed
// (src/geometry.cpp:1340-1346, 1377-1381: at most 14 characters)
std::string gmsh_name(const std::string& word) {
  std::string n = word;
  if (!n.empty() && n[0] == '"') n.erase(0, 1);
  n = n.substr(0, 14);
  while (!n.empty() && (n.back() == '"' || std::isspace((unsigned char)n.back()))) n.pop_back();
  return n;
}

}  // namespace

std::string read_gmsh(const std::string& path, mesh& m) {
  std::ifstream f(path);
  if (!f) return "Unable to open mesh file " + path;
  std::string line;
  auto seek = [&](const char* tag) {
    f.clear();
    f.seekg(0);
    while (std::getline(f, line))
      if (line.find(tag) != std::string::npos) return true;
    return false;
  };
  if (seek("$MeshFormat")) {
    double version = 0;
    int file_type = -1, dsize = 0;
    f >> version >> file_type >> dsize;
    if (!f || version >= 3.0 || file_type != 0) return "Gmsh: only the ASCII 2.x mesh format is read";
  }
  if (!seek("$PhysicalNames")) return "$PhysicalNames tag not found!";
  int nphys = 0;
  f >> nphys;
  std::getline(f, line);
  std::unordered_map<int, std::string> names;
  std::unordered_map<int, bool> fluid;
  int nd = 0;
  for (int i = 0; i < nphys; i++) {
    std::getline(f, line);
    std::istringstream ss(line);
    int dim = 0, tag = 0;
    std::string word;
    ss >> dim >> tag >> word;
    if (!ss) return "error reading $PhysicalNames";
    // the reference takes every name that contains FLUID as the domain
    // (strstr, src/geometry.cpp:1918-1921) and its dimension as the mesh's
    fluid[tag] = word.find("FLUID") != std::string::npos;
    names[tag] = gmsh_name(word);
    if (fluid[tag]) nd = dim;
  }
  if (nd != 2 && nd != 3) return "Invalid mesh dimensionality. Expected 2D or 3D.";
  if (!seek("$Nodes")) return "$Nodes tag not found!";
  int nn = 0;
  f >> nn;
  if (!f || nn <= 0) return "error reading $Nodes";
  // vertex index = rank of the node tag (tag - 1 for tags 1..nn, as the reference
  // indexes them), whatever order the lines list them in
  std::vector<std::pair<long, std::array<double, 3>>> nodes((size_t)nn);
  for (int i = 0; i < nn; i++) {
    f >> nodes[i].first >> nodes[i].second[0] >> nodes[i].second[1] >> nodes[i].second[2];
    if (!f) return "error reading $Nodes";
  }
  std::sort(nodes.begin(), nodes.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
  std::unordered_map<long, int> node_index;
  node_index.reserve(nn);
  m.n_dims = nd;
  m.n_verts = nn;
  m.xv.assign((size_t)nn * nd, 0.0);
  for (int i = 0; i < nn; i++) {
    if (i > 0 && nodes[i].first == nodes[i - 1].first) return "Gmsh node tag listed twice";
    node_index[nodes[i].first] = i;
    for (int k = 0; k < nd; k++) m.xv[(size_t)i * nd + k] = nodes[i].second[k];
  }
  if (!seek("$Elements")) return "$Elements tag not found!";
  int nent = 0;
  f >> nent;
  if (!f || nent <= 0) return "error reading $Elements";
  m.ele_type = 0;
  m.n_eles = 0;
  m.c2v.clear();
  struct BFace {
    std::array<int, 4> v;
    int type;
  };
  std::vector<BFace> bfaces;
  for (int i = 0; i < nent; i++) {
    long id;
    int etype = 0, ntags = 0;
    f >> id >> etype >> ntags;
    if (!f || ntags < 1) return "error reading $Elements (each element needs its physical tag)";
    int phys = 0, dummy;
    f >> phys;
    for (int t = 1; t < ntags; t++) f >> dummy;
    const int nv = gmsh_nodes(etype);
    if (nv == 0) return "Gmsh element type " + std::to_string(etype) + " is not read (linear elements only)";
    int v[8];
    for (int q = 0; q < nv; q++) {
      long nid;
      f >> nid;
      auto it = node_index.find(nid);
      if (!f || it == node_index.end()) return "Gmsh element " + std::to_string(id) + " names an unknown node";
      v[q] = it->second;
    }
    if (!names.count(phys)) return "Gmsh element " + std::to_string(id) + " has no physical name";
    if (fluid[phys]) {
      const int et = etype == 3 ? 1 : (etype == 5 ? 4 : (etype == 4 ? 2 : -1));
      if (et < 0) return "only linear quads, linear tets and 8-node hexes are supported";
      if (m.ele_type == 0) {
        m.ele_type = et;
        m.n_spts = nv;
      } else if (m.ele_type != et) {
        return "mixed element types are not supported";
      }
      const int* pos = et == 4 ? kGmshHexPos : (et == 1 ? kGmshQuadPos : kTetPos);
      const size_t base = m.c2v.size();
      m.c2v.resize(base + nv);
      for (int q = 0; q < nv; q++) m.c2v[base + pos[q]] = v[q];
      m.n_eles++;
    } else {
      const int type = bc_number(names[phys]);
      if (type < 0) return "boundary condition '" + names[phys] + "' not recognized";
      if (type == 9) continue;  // cyclic: matched by translation
      if (nv > 4) return "Gmsh boundary element type " + std::to_string(etype) + " not recognized";
      BFace b;
      b.v = {-1, -1, -1, -1};
      for (int q = 0; q < nv; q++) b.v[q] = v[q];
      std::sort(b.v.begin(), b.v.begin() + nv);
      b.type = type;
      bfaces.push_back(b);
    }
  }
  if (m.n_eles == 0) return "the mesh has no FLUID elements";
  // boundary faces -> (element, local face) by their vertex sets
  // (read_boundary_gmsh, src/geometry.cpp:1480-1530)
  std::unique_ptr<eles> shape;
  if (m.ele_type == 4) shape.reset(new eles_hexas());
  else if (m.ele_type == 1) shape.reset(new eles_quads());
  else shape.reset(new eles_tets());
  const int nfaces = m.ele_type == 4 ? 6 : (m.ele_type == 1 ? 4 : 4);
  std::unordered_map<FaceKey, long, FaceKeyHash> where;
  where.reserve((size_t)m.n_eles * nfaces);
  for (int e = 0; e < m.n_eles; e++)
    for (int lf = 0; lf < nfaces; lf++) {
      int nv;
      const int* fn = shape->face_nodes(lf, nv);
      FaceKey k;
      k.v = {-1, -1, -1, -1};
      for (int q = 0; q < nv; q++) k.v[q] = m.c2v[(size_t)e * m.n_spts + fn[q]];
      std::sort(k.v.begin(), k.v.begin() + nv);
      where.emplace(k, (long)e * nfaces + lf);
    }
  m.bf_ele.clear();
  m.bf_face.clear();
  m.bf_type.clear();
  for (const BFace& b : bfaces) {
    FaceKey k;
    k.v = b.v;
    auto it = where.find(k);
    if (it == where.end()) return "a Gmsh boundary face matches no element face";
    m.bf_ele.push_back((int)(it->second / nfaces));
    m.bf_face.push_back((int)(it->second % nfaces));
    m.bf_type.push_back(b.type);
  }
  return "";
}

std::string read_mesh(const std::string& path, mesh& m) {
  auto ends = [&](const char* ext) {
    const size_t n = std::strlen(ext);
    return path.size() >= n && path.compare(path.size() - n, n, ext) == 0;
  };
  if (ends("neu")) return read_gambit(path, m);
  if (ends("msh")) return read_gmsh(path, m);
  return "Mesh format not recognized";
}

void make_box(int ele_type, const int* nxyz, const double* l, const double* o, int digits, mesh& m) {
  const int nd = ele_type == 4 ? 3 : 2;
  m.n_dims = nd;
  m.ele_type = ele_type;
  m.n_spts = ele_type == 4 ? 8 : 4;
  int nv[3] = {1, 1, 1};
  for (int d = 0; d < nd; d++) nv[d] = nxyz[d] + 1;
  std::vector<double> c[3];
  for (int d = 0; d < nd; d++) {
    c[d].resize(nv[d]);
    for (int i = 0; i < nv[d]; i++) {
      double x = o[d] + l[d] * i / nxyz[d];
      if (digits > 0) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.*e", digits - 1, x);
        x = std::strtod(buf, nullptr);
      }
      c[d][i] = x;
    }
  }
  m.n_verts = nv[0] * nv[1] * nv[2];
  m.xv.assign((size_t)m.n_verts * nd, 0.0);
  for (int k = 0; k < nv[2]; k++)
    for (int j = 0; j < nv[1]; j++)
      for (int i = 0; i < nv[0]; i++) {
        const size_t v = i + (size_t)nv[0] * (j + (size_t)nv[1] * k);
        m.xv[v * nd + 0] = c[0][i];
        m.xv[v * nd + 1] = c[1][j];
        if (nd == 3) m.xv[v * nd + 2] = c[2][k];
      }
  auto vid = [&](int i, int j, int k) { return i + nv[0] * (j + nv[1] * k); };
  const int nz = nd == 3 ? nxyz[2] : 1;
  m.n_eles = nxyz[0] * nxyz[1] * nz;
  m.c2v.assign((size_t)m.n_eles * m.n_spts, 0);
  size_t e = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < nxyz[1]; j++)
      for (int i = 0; i < nxyz[0]; i++, e++) {
        int* c2v = &m.c2v[e * m.n_spts];
        if (nd == 3) {
          // Gambit brick order (x fastest), then the reader's permutation
          const int g[8] = {vid(i, j, k),     vid(i + 1, j, k),     vid(i, j + 1, k),     vid(i + 1, j + 1, k),
                            vid(i, j, k + 1), vid(i + 1, j, k + 1), vid(i, j + 1, k + 1), vid(i + 1, j + 1, k + 1)};
          for (int q = 0; q < 8; q++) c2v[kHexPos[q]] = g[q];
        } else {
          const int g[4] = {vid(i, j, 0), vid(i + 1, j, 0), vid(i + 1, j + 1, 0), vid(i, j + 1, 0)};
          for (int q = 0; q < 4; q++) c2v[kQuadPos[q]] = g[q];
        }
      }
}

std::string build_interfaces(const eles& el, const mesh& m, const double* delta, int_inters& out) {
  const int N = el.n_eles, nfaces = el.n_inters_per_ele, nd = el.n_dims, nfp = el.n_fpts_per_ele,
            nfpf = el.n_fpts_per_inter, ns = m.n_spts;
  const long nrec = (long)N * nfaces;
  std::vector<long> partner(nrec, -1);
  // boundary faces: partner -2, never matched
  out.bdy_fpt.clear();
  out.bdy_type.clear();
  for (size_t q = 0; q < m.bf_ele.size(); q++) {
    const int e = m.bf_ele[q], f = m.bf_face[q];
    if (e < 0 || e >= N || f < 0 || f >= nfaces) return "boundary face out of range";
    if (partner[(long)e * nfaces + f] == -2) return "boundary face listed twice";
    partner[(long)e * nfaces + f] = -2;
    for (int j = 0; j < nfpf; j++) {
      out.bdy_fpt.push_back(f * nfpf + j + nfp * e);
      out.bdy_type.push_back(m.bf_type[q]);
    }
  }
  {
    std::unordered_map<FaceKey, long, FaceKeyHash> first;
    first.reserve(nrec);
    for (int e = 0; e < N; e++)
      for (int f = 0; f < nfaces; f++) {
        int nv;
        const int* fn = el.face_nodes(f, nv);
        FaceKey k;
        k.v = {-1, -1, -1, -1};
        for (int q = 0; q < nv; q++) k.v[q] = m.c2v[(size_t)e * ns + fn[q]];
        std::sort(k.v.begin(), k.v.begin() + nv);
        const long id = (long)e * nfaces + f;
        if (partner[id] == -2) continue;
        auto it = first.find(k);
        if (it == first.end()) {
          first.emplace(k, id);
        } else {
          if (partner[it->second] >= 0) return "a face is shared by more than two elements";
          partner[it->second] = id;
          partner[id] = it->second;
        }
      }
  }
  // centroids of the unmatched faces
  std::vector<long> un;
  for (long i = 0; i < nrec; i++)
    if (partner[i] == -1) un.push_back(i);
  auto centroid = [&](long id, double* c) {
    const int e = (int)(id / nfaces), f = (int)(id % nfaces);
    int nv;
    const int* fn = el.face_nodes(f, nv);
    for (int d = 0; d < nd; d++) c[d] = 0.0;
    for (int q = 0; q < nv; q++)
      for (int d = 0; d < nd; d++) c[d] += m.xv[(size_t)m.c2v[(size_t)e * ns + fn[q]] * nd + d] / nv;
  };
  const double tol = 1.e-6;
  const double bin = 1.e-3;
  std::vector<double> cen(un.size() * 3, 0.0);
  std::unordered_map<long long, std::vector<int>> grid;
  auto key = [&](const double* c) {
    long long k = 0;
    for (int d = 0; d < 3; d++) k = k * 2000003LL + (long long)std::floor((d < nd ? c[d] : 0.0) / bin);
    return k;
  };
  for (size_t i = 0; i < un.size(); i++) {
    centroid(un[i], &cen[3 * i]);
    grid[key(&cen[3 * i])].push_back((int)i);
  }
  for (size_t i = 0; i < un.size(); i++) {
    if (partner[un[i]] >= 0) continue;
    bool found = false;
    for (int m2 = 0; m2 < nd && !found; m2++) {
      if (!std::isfinite(delta[m2])) continue;
      for (int sgn = -1; sgn <= 1 && !found; sgn += 2) {
        double t[3] = {cen[3 * i], cen[3 * i + 1], cen[3 * i + 2]};
        t[m2] += sgn * delta[m2];
        // search the neighbouring bins of the translated centroid
        int lo[3], hi[3];
        for (int d = 0; d < 3; d++) {
          lo[d] = d < nd ? -1 : 0;
          hi[d] = d < nd ? 1 : 0;
        }
        for (int a = lo[0]; a <= hi[0] && !found; a++)
          for (int b = lo[1]; b <= hi[1] && !found; b++)
            for (int cc = lo[2]; cc <= hi[2] && !found; cc++) {
              double q[3] = {t[0] + a * bin, t[1] + b * bin, t[2] + cc * bin};
              auto it = grid.find(key(q));
              if (it == grid.end()) continue;
              for (int j : it->second) {
                if ((size_t)j == i || partner[un[j]] >= 0) continue;
                bool ok = true;
                for (int d = 0; d < nd; d++) {
                  const double dd = cen[3 * j + d] - cen[3 * i + d];
                  if (d == m2) ok = ok && std::fabs(std::fabs(dd) - delta[d]) < tol;
                  else ok = ok && std::fabs(dd) < tol;
                }
                if (ok) {
                  partner[un[i]] = un[j];
                  partner[un[j]] = un[i];
                  found = true;
                  break;
                }
              }
            }
      }
    }
  }
  long n_bdy_faces = 0;
  for (long i = 0; i < nrec; i++) {
    if (partner[i] == -2) n_bdy_faces++;
    if (partner[i] == -1) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "face %ld of element %ld has no neighbour and no boundary condition",
                    i % nfaces, i / nfaces);
      return buf;
    }
  }
  // pair flux points by position
  out.n_inters = (int)((nrec - n_bdy_faces) / 2);
  out.left.assign((size_t)out.n_inters * nfpf, 0);
  out.right.assign((size_t)out.n_inters * nfpf, 0);
  const size_t NN = (size_t)N;
  int it = 0;
  for (long a = 0; a < nrec; a++) {
    const long b = partner[a];
    if (b < a) continue;
    const int eL = (int)(a / nfaces), fL = (int)(a % nfaces), eR = (int)(b / nfaces), fR = (int)(b % nfaces);
    double cl[3], cr[3], shift[3] = {0, 0, 0};
    centroid(a, cl);
    centroid(b, cr);
    for (int d = 0; d < nd; d++) {
      const double dd = cr[d] - cl[d];
      if (std::isfinite(delta[d]) && std::fabs(std::fabs(dd) - delta[d]) < tol) shift[d] = dd > 0 ? delta[d] : -delta[d];
    }
    for (int j = 0; j < nfpf; j++) {
      const int fl = fL * nfpf + j;
      double pl[3];
      for (int d = 0; d < nd; d++) pl[d] = el.pos_fpts[fl + nfp * (eL + NN * d)] + shift[d];
      int best = -1;
      double bd = 1e300;
      for (int q = 0; q < nfpf; q++) {
        const int fr = fR * nfpf + q;
        double dd = 0.0;
        for (int d = 0; d < nd; d++) {
          const double t = el.pos_fpts[fr + nfp * (eR + NN * d)] - pl[d];
          dd += t * t;
        }
        if (dd < bd) {
          bd = dd;
          best = fr;
        }
      }
      out.left[(size_t)it * nfpf + j] = fl + nfp * eL;
      out.right[(size_t)it * nfpf + j] = best + nfp * eR;
    }
    it++;
  }
  return "";
}

}  // namespace hifr_host
