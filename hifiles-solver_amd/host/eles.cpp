// Host element classes: point sets, dense element operators, metrics.
#include "eles.h"

#include <cmath>
#include <cstdio>
#include <stdexcept>

namespace hifr_host {

// ----------------------------------------------------------------------------
// 1-D bases

// Gauss-Legendre points with the 15 decimals the reference tabulates
// (data/loc_1d_gauss_pts.dat); parity of the operators at 1e-15 needs these exact
// values, including their asymmetric last digits at P = 5, 6.
std::vector<double> gauss_points(int order) {
  static const double tab[7][7] = {
      {0.0},
      {-0.577350269189626, 0.577350269189626},
      {-0.774596669241483, -0.000000000000000, 0.774596669241483},
      {-0.861136311594053, -0.339981043584856, 0.339981043584856, 0.861136311594053},
      {-0.906179845938664, -0.538469310105683, -0.000000000000000, 0.538469310105683, 0.906179845938664},
      {-0.932469514203152, -0.661209386466264, -0.238619186083197, 0.238619186083197, 0.661209386466265,
       0.932469514203152},
      {-0.949107912342758, -0.741531185599395, -0.405845151377397, 0.000000000000000, 0.405845151377397,
       0.741531185599394, 0.949107912342758}};
  if (order < 0 || order > 6) throw std::runtime_error("Gauss points tabulated for orders 0..6");
  return std::vector<double>(tab[order], tab[order] + order + 1);
}

// l_mode(r) = prod_{i != mode} (r - x_i) / (x_mode - x_i)
double eval_lagrange(double r, int mode, const std::vector<double>& x) {
  double t = 1.0;
  for (size_t i = 0; i < x.size(); i++)
    if ((int)i != mode) t = t * ((r - x[i]) / (x[mode] - x[i]));
  return t;
}

// l'_mode(r) = sum_{i != mode} prod_{j != mode, i} (r - x_j) / prod_{j != mode} (x_mode - x_j)
double eval_d_lagrange(double r, int mode, const std::vector<double>& x) {
  const int n = (int)x.size();
  double s = 0.0;
  for (int i = 0; i < n; i++) {
    if (i == mode) continue;
    double num = 1.0, den = 1.0;
    for (int j = 0; j < n; j++) {
      if (j != mode && j != i) num = num * (r - x[j]);
      if (j != mode) den = den * (x[mode] - x[j]);
    }
    s = s + num / den;
  }
  return s;
}

// Bonnet recursion
double eval_legendre(double r, int mode) {
  if (mode == 0) return 1.0;
  if (mode == 1) return r;
  return ((2 * mode - 1) * r * eval_legendre(r, mode - 1) - (mode - 1) * eval_legendre(r, mode - 2)) / mode;
}

double eval_d_legendre(double r, int mode) {
  if (mode == 0) return 0.0;
  if (r > -1.0 && r < 1.0)
    return (mode * ((r * eval_legendre(r, mode)) - eval_legendre(r, mode - 1))) / ((r * r) - 1.0);
  if (r == -1.0) return std::pow(-1.0, mode - 1.0) * 0.5 * mode * (mode + 1.0);
  return 0.5 * mode * (mode + 1.0);
}

// Derivative of the VCJH correction functions g_L (mode 0) and g_R (mode 1):
// g_{L,R} = (-1)^P/2 [P_P -+ (eta P_{P-1} + P_{P+1})/(1+eta)] (Vincent, Castonguay &
// Jameson 2011), as in src/funcs.cpp:216-250.
double eval_d_vcjh_1d(double r, int mode, int order, double eta) {
  if (mode == 0) {
    if (order == 0)
      return 0.5 * std::pow(-1.0, order) * (eval_d_legendre(r, order) - (eval_d_legendre(r, order + 1) / (1.0 + eta)));
    return 0.5 * std::pow(-1.0, order) *
           (eval_d_legendre(r, order) -
            (((eta * eval_d_legendre(r, order - 1)) + eval_d_legendre(r, order + 1)) / (1.0 + eta)));
  }
  if (order == 0) return 0.5 * (eval_d_legendre(r, order) + (eval_d_legendre(r, order + 1) / (1.0 + eta)));
  return 0.5 * (eval_d_legendre(r, order) +
                (((eta * eval_d_legendre(r, order - 1)) + eval_d_legendre(r, order + 1)) / (1.0 + eta)));
}

// eta of the energy-stable family: 1 DG, 2 SD, 3 Huynh (src/funcs.cpp:1503-1546)
double compute_eta(int scheme, int order) {
  if (scheme == 1) return 0.0;
  if (scheme == 2) return (1.0 * order) / (1.0 * (order + 1));
  if (scheme == 3) return (1.0 * (order + 1)) / (1.0 * order);
  throw std::runtime_error("VCJH scheme not supported (use 0 with eta, 1 DG, 2 SD or 3 Hu)");
}

// ----------------------------------------------------------------------------
// eles

eles::~eles() {
  if (ctx) hifr_destroy(ctx);
}

std::string eles::setup(const input& in, int n) {
  n_eles = n;
  order = in.order;
  const int scheme = ele_type == 4 ? in.vcjh_scheme_hexa : in.vcjh_scheme_quad;
  const int upts_type = ele_type == 4 ? in.upts_type_hexa : in.upts_type_quad;
  if (upts_type != 0) return "only Gauss solution points (upts_type 0) are supported";
  try {
    eta = scheme == 0 ? (ele_type == 4 ? in.eta_hexa : in.eta_quad) : compute_eta(scheme, order);
    loc_1d_upts = gauss_points(order);
  } catch (const std::exception& e) {
    return e.what();
  }
  const int n1 = order + 1;
  if (ele_type == 4) {
    n_dims = 3; n_inters_per_ele = 6; n_upts_per_ele = n1 * n1 * n1; n_fpts_per_inter = n1 * n1; n_spts_per_ele = 8;
  } else {
    n_dims = 2; n_inters_per_ele = 4; n_upts_per_ele = n1 * n1; n_fpts_per_inter = n1; n_spts_per_ele = 4;
  }
  n_fields = n_dims + 2;
  n_fpts_per_ele = n_inters_per_ele * n_fpts_per_inter;
  set_points();
  set_opp_0();
  set_opp_3();
  set_opp_4();
  shape.assign((size_t)n_dims * n_spts_per_ele * n_eles, 0.0);
  return "";
}

void eles::set_shape_node(int spt, int ele, const double* pos) {
  for (int m = 0; m < n_dims; m++) shape[m + n_dims * (spt + (size_t)n_spts_per_ele * ele)] = pos[m];
}

// opp_0(j,i) = l_i(fpt j)   (src/eles.cpp:3620-3660)
void eles::set_opp_0() {
  const int nu = n_upts_per_ele, nfp = n_fpts_per_ele;
  opp_0.assign((size_t)nfp * nu, 0.0);
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < nfp; j++) opp_0[j + (size_t)nfp * i] = eval_nodal_basis(i, &tloc_fpts[n_dims * j]);
}

// opp_3(j,i) = div of the correction function of flux point i at solution point j
// (fill_opp_3, src/eles_hexas.cpp:1148-1162)
void eles::set_opp_3() {
  const int nu = n_upts_per_ele, nfp = n_fpts_per_ele;
  opp_3.assign((size_t)nu * nfp, 0.0);
  for (int i = 0; i < nfp; i++)
    for (int j = 0; j < nu; j++) opp_3[j + (size_t)nu * i] = eval_div_vcjh_basis(i, &loc_upts[n_dims * j]);
}

// opp_4(d)(k,j) = d l_j / d r_d at solution point k   (src/eles.cpp:3860-3900)
void eles::set_opp_4() {
  const int nu = n_upts_per_ele;
  opp_4.assign((size_t)nu * nu * n_dims, 0.0);
  for (int d = 0; d < n_dims; d++)
    for (int j = 0; j < nu; j++)
      for (int k = 0; k < nu; k++)
        opp_4[k + (size_t)nu * (j + (size_t)nu * d)] = eval_d_nodal_basis(j, d, &loc_upts[n_dims * k]);
}

void eles::calc_pos(const double* loc, int e, double* pos) const {
  double N[8], dN[24];
  eval_shape(loc, N, dN);
  for (int i = 0; i < n_dims; i++) {
    pos[i] = 0.0;
    for (int j = 0; j < n_spts_per_ele; j++) pos[i] += N[j] * shape[i + n_dims * (j + (size_t)n_spts_per_ele * e)];
  }
}

void eles::set_transforms() {
  const int nd = n_dims, nu = n_upts_per_ele, nfp = n_fpts_per_ele, N = n_eles, ns = n_spts_per_ele;
  detjac_upts.assign((size_t)nu * N, 0.0);
  JGinv_upts.assign((size_t)nd * nd * nu * N, 0.0);
  pos_upts.assign((size_t)nu * N * nd, 0.0);
  tdA_fpts.assign((size_t)nfp * N, 0.0);
  norm_fpts.assign((size_t)nfp * N * nd, 0.0);
  pos_fpts.assign((size_t)nfp * N * nd, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double Nb[8], dN[24], d[3][3], pos[3];
    auto dpos = [&](const double* loc) {
      eval_shape(loc, Nb, dN);
      for (int j = 0; j < nd; j++) {
        pos[j] = 0.0;
        for (int s = 0; s < ns; s++) pos[j] += Nb[s] * shape[j + nd * (s + (size_t)ns * i)];
        for (int k = 0; k < nd; k++) {
          d[j][k] = 0.0;
          for (int s = 0; s < ns; s++) d[j][k] += dN[s + ns * k] * shape[j + nd * (s + (size_t)ns * i)];
        }
      }
    };
    for (int j = 0; j < nu; j++) {
      dpos(&loc_upts[nd * j]);
      const size_t pe = j + (size_t)nu * i;
      for (int k = 0; k < nd; k++) pos_upts[pe + (size_t)nu * N * k] = pos[k];
      double* J = &JGinv_upts[(size_t)nd * nd * pe];  // J[l + nd*m] = JGinv(l,m)
      if (nd == 2) {
        const double xr = d[0][0], xs = d[0][1], yr = d[1][0], ys = d[1][1];
        detjac_upts[pe] = xr * ys - xs * yr;
        J[0] = ys; J[2] = -xs; J[1] = -yr; J[3] = xr;
      } else {
        const double xr = d[0][0], xs = d[0][1], xt = d[0][2], yr = d[1][0], ys = d[1][1], yt = d[1][2],
                     zr = d[2][0], zs = d[2][1], zt = d[2][2];
        detjac_upts[pe] = xr * (ys * zt - yt * zs) - xs * (yr * zt - yt * zr) + xt * (yr * zs - ys * zr);
        J[0 + 3 * 0] = ys * zt - yt * zs; J[0 + 3 * 1] = xt * zs - xs * zt; J[0 + 3 * 2] = xs * yt - xt * ys;
        J[1 + 3 * 0] = yt * zr - yr * zt; J[1 + 3 * 1] = xr * zt - xt * zr; J[1 + 3 * 2] = xt * yr - xr * yt;
        J[2 + 3 * 0] = yr * zs - ys * zr; J[2 + 3 * 1] = xs * zr - xr * zs; J[2 + 3 * 2] = xr * ys - xs * yr;
      }
    }
    for (int j = 0; j < nfp; j++) {
      dpos(&tloc_fpts[nd * j]);
      const double* tn = &tnorm_fpts[nd * j];
      const size_t fe = j + (size_t)nfp * i;
      for (int k = 0; k < nd; k++) pos_fpts[fe + (size_t)nfp * N * k] = pos[k];
      double t[3] = {0, 0, 0};
      if (nd == 2) {
        t[0] = (tn[0] * d[1][1]) - (tn[1] * d[1][0]);
        t[1] = -(tn[0] * d[0][1]) + (tn[1] * d[0][0]);
      } else {
        t[0] = tn[0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) + tn[1] * (d[1][2] * d[2][0] - d[1][0] * d[2][2]) +
               tn[2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
        t[1] = tn[0] * (d[0][2] * d[2][1] - d[0][1] * d[2][2]) + tn[1] * (d[0][0] * d[2][2] - d[0][2] * d[2][0]) +
               tn[2] * (d[0][1] * d[2][0] - d[0][0] * d[2][1]);
        t[2] = tn[0] * (d[0][1] * d[1][2] - d[0][2] * d[1][1]) + tn[1] * (d[0][2] * d[1][0] - d[0][0] * d[1][2]) +
               tn[2] * (d[0][0] * d[1][1] - d[0][1] * d[1][0]);
      }
      double s = 0.0;
      for (int k = 0; k < nd; k++) s += t[k] * t[k];
      const double tdA = std::sqrt(s);
      tdA_fpts[fe] = tdA;
      for (int k = 0; k < nd; k++) norm_fpts[fe + (size_t)nfp * N * k] = t[k] / tdA;
    }
  }
}

std::string eles::mv_all_cpu_gpu(const input& in) {
  hifr_config c = {};
  c.ele_type = ele_type;
  c.order = order;
  c.n_eles = n_eles;
  c.viscous = in.viscous;
  c.riemann_solve_type = in.riemann_solve_type;
  c.vis_riemann_solve_type = in.vis_riemann_solve_type;
  c.precision = in.precision == 1 ? HIFR_MIXED : HIFR_FP64;
  c.gamma = in.gamma;
  c.prandtl = in.prandtl;
  c.mu_inf = in.mu_inf;
  c.rt_inf = in.rt_inf;
  c.c_sth = in.c_sth;
  c.fix_vis = in.fix_vis;
  c.tau = in.tau;
  c.pen_fact = in.pen_fact;
  c.const_src = in.const_src;
  c.metric_storage = in.metric_storage;
  c.les = in.LES;
  c.sgs_model = in.SGS_model;
  c.filter_ratio = in.filter_ratio;
  c.R_ref = in.R_ref;
  c.T_wall = in.T_wall;
  if (ctx) { hifr_destroy(ctx); ctx = nullptr; }
  hifr_status s = hifr_create(&c, &ctx);
  if (s != HIFR_OK) return std::string("hifr_create: ") + hifr_status_string(s);
  s = hifr_set_operators(ctx, opp_0.data(), opp_3.data(), opp_4.data());
  if (s != HIFR_OK) return std::string("hifr_set_operators: ") + hifr_status_string(s);
  s = hifr_set_metrics(ctx, detjac_upts.data(), JGinv_upts.data(), tdA_fpts.data(), norm_fpts.data());
  if (s != HIFR_OK) return std::string("hifr_set_metrics: ") + hifr_status_string(s);
  return "";
}

// ----------------------------------------------------------------------------
// quads (src/eles_quads.cpp)

void eles_quads::set_points() {
  const int P = order, n = P + 1;
  const std::vector<double>& g = loc_1d_upts;
  loc_upts.assign(2 * n_upts_per_ele, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const int u = j + n * i;
      loc_upts[2 * u] = g[j];
      loc_upts[2 * u + 1] = g[i];
    }
  tloc_fpts.assign(2 * n_fpts_per_ele, 0.0);
  tnorm_fpts.assign(2 * n_fpts_per_ele, 0.0);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < n; j++) {
      const int f = j + n * i;
      double* t = &tloc_fpts[2 * f];
      double* tn = &tnorm_fpts[2 * f];
      switch (i) {
        case 0: t[0] = g[j]; t[1] = -1.0; tn[1] = -1.0; break;
        case 1: t[0] = 1.0; t[1] = g[j]; tn[0] = 1.0; break;
        case 2: t[0] = g[P - j]; t[1] = 1.0; tn[1] = 1.0; break;
        default: t[0] = -1.0; t[1] = g[P - j]; tn[0] = -1.0; break;
      }
    }
}

double eles_quads::eval_nodal_basis(int idx, const double* loc) const {
  const int n = order + 1, i = idx / n, j = idx - n * i;
  return eval_lagrange(loc[0], j, loc_1d_upts) * eval_lagrange(loc[1], i, loc_1d_upts);
}

double eles_quads::eval_d_nodal_basis(int idx, int c, const double* loc) const {
  const int n = order + 1, i = idx / n, j = idx - n * i;
  if (c == 0) return eval_d_lagrange(loc[0], j, loc_1d_upts) * eval_lagrange(loc[1], i, loc_1d_upts);
  return eval_lagrange(loc[0], j, loc_1d_upts) * eval_d_lagrange(loc[1], i, loc_1d_upts);
}

double eles_quads::eval_div_vcjh_basis(int idx, const double* loc) const {
  const int P = order, i = idx / n_fpts_per_inter, j = idx - n_fpts_per_inter * i;
  const std::vector<double>& g = loc_1d_upts;
  switch (i) {
    case 0: return -eval_lagrange(loc[0], j, g) * eval_d_vcjh_1d(loc[1], 0, P, eta);
    case 1: return eval_lagrange(loc[1], j, g) * eval_d_vcjh_1d(loc[0], 1, P, eta);
    case 2: return eval_lagrange(loc[0], P - j, g) * eval_d_vcjh_1d(loc[1], 1, P, eta);
    default: return -eval_lagrange(loc[1], P - j, g) * eval_d_vcjh_1d(loc[0], 0, P, eta);
  }
}

void eles_quads::eval_shape(const double* loc, double* N, double* dN) const {
  static const std::vector<double> sp = {-1.0, 1.0};
  for (int m = 0; m < 4; m++) {
    const int i = m / 2, j = m - 2 * i;
    N[m] = eval_lagrange(loc[0], j, sp) * eval_lagrange(loc[1], i, sp);
    dN[m + 4 * 0] = eval_d_lagrange(loc[0], j, sp) * eval_lagrange(loc[1], i, sp);
    dN[m + 4 * 1] = eval_lagrange(loc[0], j, sp) * eval_d_lagrange(loc[1], i, sp);
  }
}

const int* eles_quads::face_nodes(int face, int& n) const {
  static const int fn[4][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}};
  n = 2;
  return fn[face];
}

// ----------------------------------------------------------------------------
// hexas (src/eles_hexas.cpp)

void eles_hexas::set_points() {
  const int P = order, n = P + 1;
  const std::vector<double>& g = loc_1d_upts;
  loc_upts.assign(3 * n_upts_per_ele, 0.0);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        const int u = k + n * j + n * n * i;
        loc_upts[3 * u] = g[k];
        loc_upts[3 * u + 1] = g[j];
        loc_upts[3 * u + 2] = g[i];
      }
  tloc_fpts.assign(3 * n_fpts_per_ele, 0.0);
  tnorm_fpts.assign(3 * n_fpts_per_ele, 0.0);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        const int f = k + n * j + n * n * i;
        double* t = &tloc_fpts[3 * f];
        double* tn = &tnorm_fpts[3 * f];
        switch (i) {
          case 0: t[0] = g[P - k]; t[1] = g[j]; t[2] = -1.0; tn[2] = -1.0; break;
          case 1: t[0] = g[k]; t[1] = -1.0; t[2] = g[j]; tn[1] = -1.0; break;
          case 2: t[0] = 1.0; t[1] = g[k]; t[2] = g[j]; tn[0] = 1.0; break;
          case 3: t[0] = g[P - k]; t[1] = 1.0; t[2] = g[j]; tn[1] = 1.0; break;
          case 4: t[0] = -1.0; t[1] = g[P - k]; t[2] = g[j]; tn[0] = -1.0; break;
          default: t[0] = g[k]; t[1] = g[j]; t[2] = 1.0; tn[2] = 1.0; break;
        }
      }
}

double eles_hexas::eval_nodal_basis(int idx, const double* loc) const {
  const int n = order + 1, i = idx / (n * n), j = (idx - n * n * i) / n, k = idx - n * j - n * n * i;
  return eval_lagrange(loc[0], k, loc_1d_upts) * eval_lagrange(loc[1], j, loc_1d_upts) *
         eval_lagrange(loc[2], i, loc_1d_upts);
}

double eles_hexas::eval_d_nodal_basis(int idx, int c, const double* loc) const {
  const int n = order + 1, i = idx / (n * n), j = (idx - n * n * i) / n, k = idx - n * j - n * n * i;
  const std::vector<double>& g = loc_1d_upts;
  if (c == 0) return eval_d_lagrange(loc[0], k, g) * eval_lagrange(loc[1], j, g) * eval_lagrange(loc[2], i, g);
  if (c == 1) return eval_lagrange(loc[0], k, g) * eval_d_lagrange(loc[1], j, g) * eval_lagrange(loc[2], i, g);
  return eval_lagrange(loc[0], k, g) * eval_lagrange(loc[1], j, g) * eval_d_lagrange(loc[2], i, g);
}

double eles_hexas::eval_div_vcjh_basis(int idx, const double* loc) const {
  const int P = order, n = P + 1, nf = n_fpts_per_inter;
  const int i = idx / nf, j = (idx - nf * i) / n, k = idx - nf * i - n * j;
  const std::vector<double>& g = loc_1d_upts;
  switch (i) {
    case 0: return -eval_lagrange(loc[0], P - k, g) * eval_lagrange(loc[1], j, g) * eval_d_vcjh_1d(loc[2], 0, P, eta);
    case 1: return -eval_lagrange(loc[0], k, g) * eval_lagrange(loc[2], j, g) * eval_d_vcjh_1d(loc[1], 0, P, eta);
    case 2: return eval_lagrange(loc[1], k, g) * eval_lagrange(loc[2], j, g) * eval_d_vcjh_1d(loc[0], 1, P, eta);
    case 3: return eval_lagrange(loc[0], P - k, g) * eval_lagrange(loc[2], j, g) * eval_d_vcjh_1d(loc[1], 1, P, eta);
    case 4: return -eval_lagrange(loc[1], P - k, g) * eval_lagrange(loc[2], j, g) * eval_d_vcjh_1d(loc[0], 0, P, eta);
    default: return eval_lagrange(loc[0], k, g) * eval_lagrange(loc[1], j, g) * eval_d_vcjh_1d(loc[2], 1, P, eta);
  }
}

void eles_hexas::eval_shape(const double* loc, double* N, double* dN) const {
  static const std::vector<double> sp = {-1.0, 1.0};
  for (int m = 0; m < 8; m++) {
    const int i = m / 4, j = (m - 4 * i) / 2, k = m - 2 * j - 4 * i;
    const double lk = eval_lagrange(loc[0], k, sp), lj = eval_lagrange(loc[1], j, sp), li = eval_lagrange(loc[2], i, sp);
    N[m] = lk * lj * li;
    dN[m + 8 * 0] = eval_d_lagrange(loc[0], k, sp) * lj * li;
    dN[m + 8 * 1] = lk * eval_d_lagrange(loc[1], j, sp) * li;
    dN[m + 8 * 2] = lk * lj * eval_d_lagrange(loc[2], i, sp);
  }
}

const int* eles_hexas::face_nodes(int face, int& n) const {
  static const int fn[6][4] = {{0, 1, 2, 3}, {0, 1, 4, 5}, {1, 3, 5, 7}, {2, 3, 6, 7}, {0, 2, 4, 6}, {4, 5, 6, 7}};
  n = 4;
  return fn[face];
}

}  // namespace hifr_host
