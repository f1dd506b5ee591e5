// Mesh input and face connectivity of the host solver (the part of
// src/geometry.cpp that the residual needs): Gambit neutral-file reader, a
// structured box generator for synthetic meshes, interior/periodic face matching
// and the pairing of flux points across each interface.
#pragma once
#include <string>
#include <vector>

#include "eles.h"

namespace hifr_host {

struct mesh {
  int n_dims = 0, ele_type = 0, n_verts = 0, n_eles = 0, n_spts = 0;
  std::vector<double> xv;  // (n_verts, d) row per vertex
  std::vector<int> c2v;    // (n_eles, n_spts) in the internal shape-node order
  // physical boundary faces: element, local face (internal numbering), bctype
  std::vector<int> bf_ele, bf_face, bf_type;
};

// Boundary-condition name of a Gambit group -> bctype (src/geometry.cpp:121-149);
// -1 for an unknown name.
int bc_number(std::string name);

// Gambit .neu reader (src/geometry.cpp:1122-1881): one linear element type (quads
// or 8-node hexes); node order permuted as in read_connectivity_gambit; boundary
// groups other than "cyclic" (matched geometrically) become boundary faces.
std::string read_gambit(const std::string& path, mesh& m);

// Gmsh 2.x ASCII reader (src/geometry.cpp:1309-1530, 1594-1640, 1881-2088): the
// elements of the physical group(s) whose name contains FLUID (one linear element
// type: 4-node quads, 8-node hexes or 4-node tets; node order permuted as the
// reference does); elements of the other physical groups are boundary faces, given
// the bctype of their name and matched to element faces by their vertices, except
// "cyclic" faces, which are matched by translation.
std::string read_gmsh(const std::string& path, mesh& m);

// The reader the mesh file's extension names: .neu Gambit, .msh Gmsh
// (src/input.cpp:355-360).
std::string read_mesh(const std::string& path, mesh& m);

// Structured box of nx*ny(*nz) elements on [o, o+l], optionally with vertex
// coordinates rounded to `digits` significant digits (0 = exact).
void make_box(int ele_type, const int* nxyz, const double* l, const double* o, int digits, mesh& m);

// Interior + periodic interfaces as flux-point pairs (fpt + Nfp*ele) left/right,
// and the flux points of the boundary faces with their bctype.
struct int_inters {
  int n_inters = 0;
  std::vector<int> left, right;
  std::vector<int> bdy_fpt, bdy_type;
};

// Faces listed in m.bf_* are boundary faces and are not matched.  Other faces are
// matched by shared vertices; faces left over are matched by
// translation over the finite entries of delta (dx/dy/dz_cyclic) with the
// reference's tolerance 1e-6 (src/geometry.cpp:489-563, 3634-3664).  Flux points
// are paired by position, which is what the rot_tag lookup of
// src/inters.cpp:187-296 encodes.
std::string build_interfaces(const eles& el, const mesh& m, const double* delta, int_inters& out);

}  // namespace hifr_host
