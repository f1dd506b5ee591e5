// Host element classes: point sets, dense element operators, metrics.
#include "eles.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <utility>

#include "tet_tables.h"

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
  const int n1 = order + 1;
  if (ele_type == 2) {
    // src/eles_tets.cpp:68-141
    if (in.upts_type_tet != 0 || in.fpts_type_tet != 0)
      return "tets are built with the tabulated interior points only (upts_type_tet 0, fpts_type_tet 0)";
    if (!(in.vcjh_scheme_tet == 1 || (in.vcjh_scheme_tet == 0 && in.c_tet == 0.0)))
      return "tets are built with the DG correction only (vcjh_scheme_tet 1, or 0 with c_tet 0)";
    if (order < 1 || order > 3) return "tets are built for orders 1-3";
    n_dims = 3;
    n_inters_per_ele = 4;
    n_upts_per_ele = n1 * (n1 + 1) * (n1 + 2) / 6;
    n_fpts_per_inter = n1 * (n1 + 1) / 2;
    n_spts_per_ele = 4;
    n_fields = 5;
    n_fpts_per_ele = n_inters_per_ele * n_fpts_per_inter;
    eta = 0.0;
    set_points();
    set_opp_0();
    set_opp_3();
    set_opp_4();
    shape.assign((size_t)n_dims * n_spts_per_ele * n_eles, 0.0);
    return "";
  }
  const int scheme = ele_type == 4 ? in.vcjh_scheme_hexa : in.vcjh_scheme_quad;
  const int upts_type = ele_type == 4 ? in.upts_type_hexa : in.upts_type_quad;
  if (upts_type != 0) return "only Gauss solution points (upts_type 0) are supported";
  try {
    eta = scheme == 0 ? (ele_type == 4 ? in.eta_hexa : in.eta_quad) : compute_eta(scheme, order);
    loc_1d_upts = gauss_points(order);
  } catch (const std::exception& e) {
    return e.what();
  }
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

namespace {

// Gauss-Legendre nodes and weights on [-1, 1] (Newton's iteration on P_n from the
// Chebyshev guesses; weights 2 / ((1 - x^2) P_n'(x)^2)), ascending nodes
void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  const double pi = 3.14159265358979323846;
  for (int i = 0; i < n; i++) {
    double z = -std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; k++) {
      const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

// x rounded to `dec` decimals (or `dec` significant digits), as the tables print it
double round_decimals(double x, int dec) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*f", dec, x);
  return std::strtod(buf, nullptr);
}
double round_significant(double x, int digits) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.*g", digits, x);
  return std::strtod(buf, nullptr);
}

}  // namespace

bool volume_cubature_rule(int nd, int n, std::vector<double>& loc, std::vector<double>& w) {
  const int dec = nd == 3 ? 12 : 15;
  if (n < 2 || n > (nd == 3 ? 12 : 20)) return false;
  std::vector<double> x, wx;
  gauss_legendre(n, x, wx);
  const int ncub = nd == 3 ? n * n * n : n * n;
  loc.assign((size_t)nd * ncub, 0.0);
  w.assign(ncub, 0.0);
  for (int q = 0; q < ncub; q++) {
    int idx[3];
    if (nd == 3) {
      idx[0] = q / (n * n);
      idx[1] = (q / n) % n;
      idx[2] = q % n;
    } else {
      idx[0] = q % n;
      idx[1] = q / n;
    }
    double prod = 1.0;
    for (int d = 0; d < nd; d++) {
      loc[d + (size_t)nd * q] = round_decimals(x[idx[d]], dec);
      prod *= wx[idx[d]];
    }
    w[q] = nd == 3 ? round_decimals(prod, dec) : round_significant(prod, 12);
  }
  return true;
}

std::string eles::set_volume_cubature(int rule) {
  if (ele_type != 1 && ele_type != 4) return "integral diagnostics are built for quads and hexes";
  const int nd = n_dims, nu = n_upts_per_ele, N = n_eles, ns = n_spts_per_ele;
  if (!volume_cubature_rule(nd, rule, loc_volume_cubpts, weight_volume_cubpts))
    return "cubature rule not implemented.";
  const int ncub = (int)weight_volume_cubpts.size();
  n_cubpts_per_ele = ncub;
  opp_volume_cubpts.assign((size_t)ncub * nu, 0.0);
  for (int i = 0; i < nu; i++)
    for (int j = 0; j < ncub; j++)
      opp_volume_cubpts[j + (size_t)ncub * i] = eval_nodal_basis(i, &loc_volume_cubpts[(size_t)nd * j]);
  vol_detjac_cubpts.assign((size_t)ncub * N, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < N; i++) {
    double Nb[8], dN[24], d[3][3];
    for (int j = 0; j < ncub; j++) {
      eval_shape(&loc_volume_cubpts[(size_t)nd * j], Nb, dN);
      for (int a = 0; a < nd; a++)
        for (int k = 0; k < nd; k++) {
          d[a][k] = 0.0;
          for (int sn = 0; sn < ns; sn++) d[a][k] += dN[sn + ns * k] * shape[a + nd * (sn + (size_t)ns * i)];
        }
      double dj;
      if (nd == 2) {
        dj = d[0][0] * d[1][1] - d[0][1] * d[1][0];
      } else {
        dj = d[0][0] * (d[1][1] * d[2][2] - d[1][2] * d[2][1]) - d[0][1] * (d[1][0] * d[2][2] - d[1][2] * d[2][0]) +
             d[0][2] * (d[1][0] * d[2][1] - d[1][1] * d[2][0]);
      }
      vol_detjac_cubpts[j + (size_t)ncub * i] = dj;
    }
  }
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
  c.sgs_model = in.learned_sgs ? HIFR_SGS_LEARNED : in.SGS_model;
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

// ----------------------------------------------------------------------------
// tets (src/eles_tets.cpp)

namespace {

// Orthonormal Jacobi polynomial P_n^(a,b)(x) (unit norm on [-1,1] with weight
// (1-x)^a (1+x)^b), by the normalised three-term recurrence (Hesthaven & Warburton,
// Nodal Discontinuous Galerkin Methods, 2008, JacobiP).
double jacobi(double x, double a, double b, int n) {
  const double g0 = std::pow(2.0, a + b + 1.0) / (a + b + 1.0) * std::tgamma(a + 1.0) * std::tgamma(b + 1.0) /
                    std::tgamma(a + b + 1.0);
  double pm = 1.0 / std::sqrt(g0);
  if (n == 0) return pm;
  const double g1 = (a + 1.0) * (b + 1.0) / (a + b + 3.0) * g0;
  double p = ((a + b + 2.0) * x / 2.0 + (a - b) / 2.0) / std::sqrt(g1);
  double aold = 2.0 / (2.0 + a + b) * std::sqrt((a + 1.0) * (b + 1.0) / (a + b + 3.0));
  for (int i = 1; i < n; i++) {
    const double h1 = 2.0 * i + a + b;
    const double anew =
        2.0 / (h1 + 2.0) * std::sqrt((i + 1.0) * (i + 1.0 + a + b) * (i + 1.0 + a) * (i + 1.0 + b) / (h1 + 1.0) / (h1 + 3.0));
    const double bnew = -(a * a - b * b) / h1 / (h1 + 2.0);
    const double pn = (-aold * pm + (x - bnew) * p) / anew;
    pm = p;
    p = pn;
    aold = anew;
  }
  return p;
}

double d_jacobi(double x, double a, double b, int n) {
  return n == 0 ? 0.0 : std::sqrt(n * (n + a + b + 1.0)) * jacobi(x, a + 1.0, b + 1.0, n - 1);
}

// Mode numbering of the orthonormal bases as the reference orders them
// (src/funcs.cpp:1190-1378): triangles by total degree, then j; tetrahedra by total
// degree m, then n = j + k, then k.
void tri_ij(int mode, int& i, int& j) {
  int q = 0;
  for (int deg = 0;; deg++)
    for (int jj = 0; jj <= deg; jj++, q++)
      if (q == mode) { i = deg - jj; j = jj; return; }
}
void tet_ijk(int mode, int& i, int& j, int& k) {
  int q = 0;
  for (int m = 0;; m++)
    for (int n = 0; n <= m; n++)
      for (int kk = 0; kk <= n; kk++, q++)
        if (q == mode) { i = m - n; j = n - kk; k = kk; return; }
}

// Orthonormal triangle basis on {r, s >= -1, r + s <= 0} through the collapsed
// coordinates a = 2(1+r)/(1-s) - 1, b = s.
double tri_mode(double r, double s, int mode) {
  int i, j;
  tri_ij(mode, i, j);
  const double a = s == 1.0 ? -1.0 : 2.0 * (1.0 + r) / (1.0 - s) - 1.0;
  return std::sqrt(2.0) * jacobi(a, 0, 0, i) * jacobi(s, 2 * i + 1, 0, j) * std::pow(1.0 - s, i);
}

void collapse3(double r, double s, double t, double& a, double& b, double& c) {
  a = (s + t == 0.0) ? -1.0 : -2.0 * (1.0 + r) / (s + t) - 1.0;
  b = t == 1.0 ? -1.0 : 2.0 * (1.0 + s) / (1.0 - t) - 1.0;
  c = t;
}

// Orthonormal tetrahedron basis (Hesthaven & Warburton, Simplex3DP).
double tet_mode(double r, double s, double t, int mode) {
  int i, j, k;
  tet_ijk(mode, i, j, k);
  double a, b, c;
  collapse3(r, s, t, a, b, c);
  return 2.0 * std::sqrt(2.0) * jacobi(a, 0, 0, i) * jacobi(b, 2 * i + 1, 0, j) * std::pow(1.0 - b, i) *
         jacobi(c, 2 * (i + j) + 2, 0, k) * std::pow(1.0 - c, i + j);
}

// Gradient of the tetrahedron basis (Hesthaven & Warburton, GradSimplex3DP).
void tet_mode_grad(double r, double s, double t, int mode, double* g) {
  int i, j, k;
  tet_ijk(mode, i, j, k);
  double a, b, c;
  collapse3(r, s, t, a, b, c);
  const double fa = jacobi(a, 0, 0, i), dfa = d_jacobi(a, 0, 0, i);
  const double gb = jacobi(b, 2 * i + 1, 0, j), dgb = d_jacobi(b, 2 * i + 1, 0, j);
  const double hc = jacobi(c, 2 * (i + j) + 2, 0, k), dhc = d_jacobi(c, 2 * (i + j) + 2, 0, k);
  const double hb = 0.5 * (1.0 - b), hh = 0.5 * (1.0 - c);
  double vr = dfa * gb * hc;
  if (i > 0) vr *= std::pow(hb, i - 1);
  if (i + j > 0) vr *= std::pow(hh, i + j - 1);
  double tmp = dgb * std::pow(hb, i);
  if (i > 0) tmp += -0.5 * i * gb * std::pow(hb, i - 1);
  if (i + j > 0) tmp *= std::pow(hh, i + j - 1);
  tmp = fa * tmp * hc;
  const double vs = 0.5 * (1.0 + a) * vr + tmp;
  double vt = 0.5 * (1.0 + a) * vr + 0.5 * (1.0 + b) * tmp;
  double tc = dhc * std::pow(hh, i + j);
  if (i + j > 0) tc -= 0.5 * (i + j) * hc * std::pow(hh, i + j - 1);
  vt += fa * gb * tc * std::pow(hb, i);
  const double sc = std::pow(2.0, 2 * i + j + 1.5);
  g[0] = vr * sc;
  g[1] = vs * sc;
  g[2] = vt * sc;
}

// Inverse of an n x n column-major matrix by Gauss-Jordan elimination with partial
// pivoting.
std::vector<double> inverse(std::vector<double> A, int n) {
  std::vector<double> X((size_t)n * n, 0.0);
  for (int i = 0; i < n; i++) X[i + (size_t)n * i] = 1.0;
  for (int k = 0; k < n; k++) {
    int piv = k;
    for (int i = k + 1; i < n; i++)
      if (std::fabs(A[i + (size_t)n * k]) > std::fabs(A[piv + (size_t)n * k])) piv = i;
    if (A[piv + (size_t)n * k] == 0.0) throw std::runtime_error("singular Vandermonde matrix");
    if (piv != k)
      for (int j = 0; j < n; j++) {
        std::swap(A[k + (size_t)n * j], A[piv + (size_t)n * j]);
        std::swap(X[k + (size_t)n * j], X[piv + (size_t)n * j]);
      }
    const double d = A[k + (size_t)n * k];
    for (int j = 0; j < n; j++) {
      A[k + (size_t)n * j] /= d;
      X[k + (size_t)n * j] /= d;
    }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      const double f = A[i + (size_t)n * k];
      if (f == 0.0) continue;
      for (int j = 0; j < n; j++) {
        A[i + (size_t)n * j] -= f * A[k + (size_t)n * j];
        X[i + (size_t)n * j] -= f * X[k + (size_t)n * j];
      }
    }
  }
  return X;
}

// Face coordinates (r_f, s_f) of a point on face f of the reference tet, and the
// inverse map; the slanted face 0 is parametrised by (r, t) (src/eles_tets.cpp:1018-1131).
void tet_face_coords(int f, const double* x, double& rf, double& sf) {
  switch (f) {
    case 0: rf = x[0]; sf = x[2]; break;
    case 1: rf = x[2]; sf = x[1]; break;
    case 2: rf = x[0]; sf = x[2]; break;
    default: rf = x[1]; sf = x[0]; break;
  }
}
void tet_face_point(int f, double rf, double sf, double* x) {
  switch (f) {
    case 0: x[0] = rf; x[2] = sf; x[1] = -1.0 - sf - rf; break;
    case 1: x[0] = -1.0; x[1] = sf; x[2] = rf; break;
    case 2: x[0] = rf; x[1] = -1.0; x[2] = sf; break;
    default: x[0] = sf; x[1] = rf; x[2] = -1.0; break;
  }
}

}  // namespace

// Solution points: the tabulated interior set (src/eles_tets.cpp:233-270).  Flux
// points: the triangle set mapped onto each face (set_tloc_fpts, :273-329), with the
// first coordinate of face 0 taken from the mirrored triangle point.  Normals: face 0
// (1,1,1)/sqrt(3), faces 1-3 -e_r, -e_s, -e_t (:589-620).
void eles_tets::set_points() {
  const int P = order, nu = n_upts_per_ele, nf1 = n_fpts_per_inter;
  const double* up = tet_tables::upts(P);
  const double* tp = tet_tables::tri(P);
  loc_upts.assign(up, up + 3 * nu);
  tloc_fpts.assign(3 * n_fpts_per_ele, 0.0);
  tnorm_fpts.assign(3 * n_fpts_per_ele, 0.0);
  for (int j = 0; j <= P; j++)
    for (int i = 0; i <= P - j; i++) {
      const int it = j * (P + 1) - (j - 1) * j / 2 + i;
      const int im = j * (P + 1) - (j - 1) * j / 2 + (P - j - i);
      const double r = tp[2 * it], s = tp[2 * it + 1];
      double* t0 = &tloc_fpts[3 * it];
      t0[0] = tp[2 * im]; t0[1] = r; t0[2] = s;
      double* t1 = &tloc_fpts[3 * (nf1 + it)];
      t1[0] = -1.0; t1[1] = s; t1[2] = r;
      double* t2 = &tloc_fpts[3 * (2 * nf1 + it)];
      t2[0] = r; t2[1] = -1.0; t2[2] = s;
      double* t3 = &tloc_fpts[3 * (3 * nf1 + it)];
      t3[0] = s; t3[1] = r; t3[2] = -1.0;
    }
  for (int f = 0; f < 4; f++)
    for (int q = 0; q < nf1; q++) {
      double* tn = &tnorm_fpts[3 * (f * nf1 + q)];
      if (f == 0) tn[0] = tn[1] = tn[2] = 1.0 / std::sqrt(3.0);
      else tn[f - 1] = -1.0;
    }
  // Vandermonde V(i, m) = psi_m(x_i) and its inverse (set_vandermonde, :722-733)
  std::vector<double> V((size_t)nu * nu);
  for (int i = 0; i < nu; i++)
    for (int m = 0; m < nu; m++) V[i + (size_t)nu * m] = tet_mode(loc_upts[3 * i], loc_upts[3 * i + 1], loc_upts[3 * i + 2], m);
  inv_vdm = inverse(V, nu);
}

// l_idx(x) = sum_m Vinv(m, idx) psi_m(x)   (src/eles_tets.cpp:805-822)
double eles_tets::eval_nodal_basis(int idx, const double* loc) const {
  const int nu = n_upts_per_ele;
  double v = 0.0;
  for (int m = 0; m < nu; m++) v += inv_vdm[m + (size_t)nu * idx] * tet_mode(loc[0], loc[1], loc[2], m);
  return v;
}

// (src/eles_tets.cpp:843-860)
double eles_tets::eval_d_nodal_basis(int idx, int c, const double* loc) const {
  const int nu = n_upts_per_ele;
  double v = 0.0, g[3];
  for (int m = 0; m < nu; m++) {
    tet_mode_grad(loc[0], loc[1], loc[2], m, g);
    v += inv_vdm[m + (size_t)nu * idx] * g[c];
  }
  return v;
}

double eles_tets::eval_div_vcjh_basis(int, const double*) const {
  throw std::logic_error("tets build their correction operator in set_opp_3");
}

// DG correction (c_tet = 0, so the filter of src/eles_tets.cpp:972-1016 is the
// identity): the correction field of flux point q is the L2 projection onto P_P(tet)
// of its face Lagrange polynomial, lifted from the face (eval_div_dg_tet,
// :1018-1131).  With the orthonormal basis the projection's coefficients are the
// face integrals int_face psi_m l_q dA, taken with the 36-point triangle rule and the
// face Jacobian (sqrt(3) on the slanted face).
void eles_tets::set_opp_3() {
  const int nu = n_upts_per_ele, nf1 = n_fpts_per_inter, nfp = n_fpts_per_ele;
  opp_3.assign((size_t)nu * nfp, 0.0);
  std::vector<double> coef(nu);
  for (int f = 0; f < 4; f++) {
    // face Vandermonde W(q, m) = phi_m(face point q); column q of W^-1 gives the
    // orthonormal coefficients of the face Lagrange polynomial l_q
    std::vector<double> W((size_t)nf1 * nf1);
    for (int q = 0; q < nf1; q++) {
      double rf, sf;
      tet_face_coords(f, &tloc_fpts[3 * (f * nf1 + q)], rf, sf);
      for (int m = 0; m < nf1; m++) W[q + (size_t)nf1 * m] = tri_mode(rf, sf, m);
    }
    const std::vector<double> Wi = inverse(W, nf1);
    const double jac = f == 0 ? std::sqrt(3.0) : 1.0;
    // quadrature values of psi_m on the face and of phi_k in face coordinates
    std::vector<double> psi((size_t)36 * nu), phi((size_t)36 * nf1);
    for (int g = 0; g < 36; g++) {
      const double rf = tet_tables::tri_rule12[g][0], sf = tet_tables::tri_rule12[g][1];
      double x[3];
      tet_face_point(f, rf, sf, x);
      for (int m = 0; m < nu; m++) psi[g + 36 * (size_t)m] = tet_mode(x[0], x[1], x[2], m);
      for (int k = 0; k < nf1; k++) phi[g + 36 * (size_t)k] = tri_mode(rf, sf, k);
    }
    for (int q = 0; q < nf1; q++) {
      for (int m = 0; m < nu; m++) {
        double acc = 0.0;
        for (int g = 0; g < 36; g++) {
          double l = 0.0;
          for (int k = 0; k < nf1; k++) l += Wi[k + (size_t)nf1 * q] * phi[g + 36 * (size_t)k];
          acc += tet_tables::tri_rule12[g][2] * psi[g + 36 * (size_t)m] * l;
        }
        coef[m] = acc * jac;
      }
      const int fp = f * nf1 + q;
      for (int j = 0; j < nu; j++) {
        double v = 0.0;
        for (int m = 0; m < nu; m++) v += coef[m] * tet_mode(loc_upts[3 * j], loc_upts[3 * j + 1], loc_upts[3 * j + 2], m);
        opp_3[j + (size_t)nu * fp] = v;
      }
    }
  }
}

// Linear tet shape functions (src/eles_tets.cpp:862-880, 909-935)
void eles_tets::eval_shape(const double* loc, double* N, double* dN) const {
  N[0] = -0.5 * (loc[0] + loc[1] + loc[2] + 1.0);
  N[1] = 0.5 * (loc[0] + 1.0);
  N[2] = 0.5 * (loc[1] + 1.0);
  N[3] = 0.5 * (loc[2] + 1.0);
  for (int m = 0; m < 12; m++) dN[m] = 0.0;
  for (int k = 0; k < 3; k++) {
    dN[0 + 4 * k] = -0.5;
    dN[k + 1 + 4 * k] = 0.5;
  }
}

// face f is the face opposite vertex f for f = 0 (the slanted face) and the
// coordinate planes r, s, t = -1 for f = 1, 2, 3 (src/geometry.cpp:3156-3185)
const int* eles_tets::face_nodes(int face, int& n) const {
  static const int fn[4][3] = {{1, 2, 3}, {0, 3, 2}, {0, 1, 3}, {0, 2, 1}};
  n = 3;
  return fn[face];
}

}  // namespace hifr_host
