// Element classes of the host solver: the reference's eles / eles_quads / eles_hexas /
// eles_tets surface (include/eles.h, src/eles.cpp, src/eles_quads.cpp,
// src/eles_hexas.cpp, src/eles_tets.cpp)
// reduced to what the residual needs.  They build the solution/flux point sets,
// the dense element operators opp_0, opp_3, opp_4 in the reference layouts and the
// static-mesh metrics, hand them to the device context (include/hifr.h) and then
// drive the residual and RK update through it.
#pragma once
#include <string>
#include <vector>

#include "../../include/hifr.h"
#include "input.h"

namespace hifr_host {

class eles {
 public:
  virtual ~eles();

  // src/eles.cpp:84-405 (storage sizes, operators) for n_eles elements
  std::string setup(const input& in, int n_eles);
  void set_shape_node(int spt, int ele, const double* pos);
  // src/eles.cpp:4557-4847
  void set_transforms();
  // src/eles.cpp:5744-5757
  void calc_pos(const double* loc, int ele, double* pos) const;

  // device side (replaces mv_all_cpu_gpu, src/eles.cpp:775)
  std::string mv_all_cpu_gpu(const input& in);
  hifr_ctx* ctx = nullptr;

  int ele_type = 0, n_dims = 0, n_fields = 0, order = 0;
  int n_upts_per_ele = 0, n_fpts_per_ele = 0, n_inters_per_ele = 0, n_fpts_per_inter = 0;
  int n_eles = 0, n_spts_per_ele = 0;
  std::vector<double> loc_1d_upts;
  std::vector<double> loc_upts, tloc_fpts, tnorm_fpts;  // (d, Nu), (d, Nfp), (d, Nfp)
  std::vector<double> opp_0, opp_3, opp_4;              // (Nfp,Nu), (Nu,Nfp), (Nu,Nu,d)
  std::vector<double> shape;                            // (d, nspts, N)
  std::vector<double> detjac_upts, JGinv_upts, tdA_fpts, norm_fpts, pos_upts, pos_fpts;
  double eta = 0.0;

  // local faces as shape-node lists (for connectivity)
  virtual const int* face_nodes(int face, int& n) const = 0;

  // Volume cubature of the integral diagnostics (src/eles_hexas.cpp:418-440,
  // src/eles_quads.cpp set_volume_cubpts, src/eles.cpp:4111-4130, 5318-5352):
  // the reference's tabulated rule with `rule` points per direction, the nodal
  // basis at its points and detJ there.  Quads and hexes.
  std::string set_volume_cubature(int rule);
  int n_cubpts_per_ele = 0;
  std::vector<double> loc_volume_cubpts;     // (d, ncub)
  std::vector<double> weight_volume_cubpts;  // (ncub)
  std::vector<double> opp_volume_cubpts;     // (ncub, Nu)
  std::vector<double> vol_detjac_cubpts;     // (ncub, N)

 protected:
  virtual void set_points() = 0;
  virtual double eval_nodal_basis(int idx, const double* loc) const = 0;
  virtual double eval_d_nodal_basis(int idx, int cpnt, const double* loc) const = 0;
  virtual double eval_div_vcjh_basis(int idx, const double* loc) const = 0;
  virtual void eval_shape(const double* loc, double* N, double* dN) const = 0;
  void set_opp_0();
  virtual void set_opp_3();
  void set_opp_4();
};

class eles_quads : public eles {
 public:
  const int* face_nodes(int face, int& n) const override;

 protected:
  void set_points() override;
  double eval_nodal_basis(int idx, const double* loc) const override;
  double eval_d_nodal_basis(int idx, int cpnt, const double* loc) const override;
  double eval_div_vcjh_basis(int idx, const double* loc) const override;
  void eval_shape(const double* loc, double* N, double* dN) const override;
};

class eles_hexas : public eles {
 public:
  const int* face_nodes(int face, int& n) const override;

 protected:
  void set_points() override;
  double eval_nodal_basis(int idx, const double* loc) const override;
  double eval_d_nodal_basis(int idx, int cpnt, const double* loc) const override;
  double eval_div_vcjh_basis(int idx, const double* loc) const override;
  void eval_shape(const double* loc, double* N, double* dN) const override;
};

// Tetrahedra with the DG correction (vcjh_scheme_tet 1, or 0 with c_tet 0), orders
// 1-3, linear (4-node) shape.  The nodal basis is expressed through the orthonormal
// (Dubiner) basis and the inverse of its Vandermonde matrix at the solution points.
class eles_tets : public eles {
 public:
  const int* face_nodes(int face, int& n) const override;

 protected:
  void set_points() override;
  double eval_nodal_basis(int idx, const double* loc) const override;
  double eval_d_nodal_basis(int idx, int cpnt, const double* loc) const override;
  double eval_div_vcjh_basis(int idx, const double* loc) const override;
  void eval_shape(const double* loc, double* N, double* dN) const override;
  void set_opp_3() override;
  std::vector<double> inv_vdm;  // (Nu, Nu): nodal coefficients of each orthonormal mode
};

// 1-D bases, src/funcs.cpp:57-250, 1503-1546
double eval_lagrange(double r, int mode, const std::vector<double>& pts);
double eval_d_lagrange(double r, int mode, const std::vector<double>& pts);
double eval_legendre(double r, int mode);
double eval_d_legendre(double r, int mode);
double eval_d_vcjh_1d(double r, int mode, int order, double eta);
double compute_eta(int scheme, int order);
// 1-D Gauss points as tabulated by data/loc_1d_gauss_pts.dat (orders 0..6)
std::vector<double> gauss_points(int order);
// The reference's tensor-product volume cubature rules with n points per direction
// (data/cubature_hexa.dat, n = 2..12; data/cubature_quad.dat, n = 2..20): Gauss-
// Legendre points and products of weights, rounded as the tables print them: hexes
// 12 decimals; quads points 15 decimals, weights 12 significant digits.  Point
// order: hexes x slowest, z fastest; quads x fastest.  Returns false for a rule the
// tables do not hold.
bool volume_cubature_rule(int nd, int n, std::vector<double>& loc, std::vector<double>& w);

}  // namespace hifr_host
