// Input deck of the solver: the reference's `name value` file format, key names and
// defaults (src/input.cpp:89-331) and its non-dimensionalisation (:345-515).
#pragma once
#include <limits>
#include <string>
#include <vector>

namespace hifr_host {

struct input {
  // basic
  int equation = 0, viscous = 0, ic_form = 1, test_case = 0, order = 3, n_steps = 0;
  int restart_flag = 0, restart_iter = 0, n_restart_files = 1;
  std::string mesh_file;
  // output / monitoring
  int plot_freq = 500, restart_dump_freq = 0, monitor_res_freq = 100, res_norm_type = 2;
  // integral diagnostics (kineticenergy, vorticity, pressuredilatation,
  // straincolonproduct, devstraincolonproduct; src/input.cpp:133-143) and the
  // points per direction of their volume cubature rule
  std::vector<std::string> integral_quantities;
  int volume_cub_order = 3;
  std::string data_file_name = "Mesh";
  // solver
  int riemann_solve_type = 0, vis_riemann_solve_type = 0, adv_type = 3, dt_type = 0;
  double dt = 0.0, CFL = 0.0;
  int turb_model = 0, LES = 0, SGS_model = 0, wall_model = 0, motion = 0;
  double filter_ratio = 1.0;
  // gas
  double gamma = 1.4, prandtl = 0.72, S_gas = 120., T_gas = 291.15, R_gas = 286.9, mu_gas = 1.827E-5;
  // boundary conditions / reference state
  double dx_cyclic = std::numeric_limits<double>::infinity(), dy_cyclic = std::numeric_limits<double>::infinity(),
         dz_cyclic = std::numeric_limits<double>::infinity();
  double Mach_free_stream = 0., nx_free_stream = 0., ny_free_stream = 0., nz_free_stream = 0.;
  double Re_free_stream = 0., L_free_stream = 1., T_free_stream = 300.;
  double rho_bound = 1., v_bound[3] = {0, 0, 0}, p_bound = 1.;
  double Mach_wall = 0., nx_wall = 0., ny_wall = 0., nz_wall = 0., T_wall = 300.;
  double fix_vis = 1., tau = 0., pen_fact = 0.5;
  double Mach_c_ic = 0., nx_c_ic = 0., ny_c_ic = 0., nz_c_ic = 0., Re_c_ic = 0., T_c_ic = 300.;
  double rho_c_ic = 1., u_c_ic = 0., v_c_ic = 0., w_c_ic = 0., p_c_ic = 1.;
  // elements
  int upts_type_quad = 0, vcjh_scheme_quad = 0, upts_type_hexa = 0, vcjh_scheme_hexa = 0;
  double eta_quad = 0., eta_hexa = 0.;
  int upts_type_tet = 0, fpts_type_tet = 0, vcjh_scheme_tet = 0, sparse_tet = 0;
  double c_tet = 0., eta_tet = 0.;
  double const_src = 0.;
  // Not a reference key: hifr_host_only 1 builds the host-side element data
  // (operators, metrics, interfaces) without creating the device context.
  int host_only = 0;
  // Not a reference key: hifr_metrics 1 keeps per-point metrics on the device even
  // when every element is affine (HIFR_METRICS_PER_POINT).
  int metric_storage = 0;
  // Not a reference key: hifr_precision 1 selects the mixed-precision kernels
  // (HIFR_MIXED); 0 is FP64.
  int precision = 0;
  // Not a reference key: hifr_learned_sgs 1 (with LES 1, SGS_model 1) selects the
  // learned eddy viscosity with the WALE fallback (HIFR_SGS_LEARNED).
  int learned_sgs = 0;
  // derived (setup_params)
  double mu_inf = 0., rt_inf = 0., c_sth = 0., R_ref = 0., prandtl_t = 0., T_ref = 0., L_ref = 0.;
  double uvw_ref = 0., rho_ref = 0., p_ref = 0., mu_ref = 0., time_ref = 0.;
  double v_wall[3] = {0, 0, 0};

  input();
  // Reads a deck; returns an empty string or an error message.
  std::string read_input_file(const std::string& path);
  // src/input.cpp:345-515
  void setup_params();
};

}  // namespace hifr_host
