// Deck reader and non-dimensionalisation of the host solver.
// Key names and defaults follow src/input.cpp:89-331; setup_params restates the
// arithmetic of src/input.cpp:345-515 in the same order so the derived constants
// (mu_inf, rt_inf, ...) are the same doubles as the reference's.
#include "input.h"

#include <cctype>
#include <cmath>
#include <fstream>
#include <map>
#include <cstdlib>
#include <sstream>

namespace hifr_host {

input::input() {
  dx_cyclic = dy_cyclic = dz_cyclic = INFINITY;
}

namespace {

struct deck {
  std::map<std::string, std::string> kv;
  bool has(const char* k) const { return kv.count(k) != 0; }
  template <typename T>
  void get(const char* k, T& v) const {
    auto it = kv.find(k);
    if (it == kv.end()) return;
    std::istringstream ss(it->second);
    T t;
    if (ss >> t) v = t;
  }
  // doubles through strtod, which also reads inf / nan
  void get(const char* k, double& v) const {
    auto it = kv.find(k);
    if (it == kv.end()) return;
    char* end = nullptr;
    const double t = std::strtod(it->second.c_str(), &end);
    if (end != it->second.c_str()) v = t;
  }
};

}  // namespace

std::string input::read_input_file(const std::string& path) {
  std::ifstream f(path);
  if (!f) return "cannot open input file " + path;
  deck d;
  std::string line;
  // first occurrence of a key wins, like fileReader::getScalarValue
  while (std::getline(f, line)) {
    std::istringstream ss(line);
    std::string key;
    if (!(ss >> key)) continue;
    std::string rest;
    std::getline(ss, rest);
    if (!d.kv.count(key)) d.kv[key] = rest;
  }
  const char* required[] = {"equation", "order", "mesh_file", "n_steps", "riemann_solve_type",
                            "vis_riemann_solve_type", "adv_type", "dt_type"};
  for (const char* r : required)
    if (!d.has(r)) return std::string("Required option not found: ") + r;
  d.get("equation", equation);
  d.get("order", order);
  d.get("viscous", viscous);
  d.get("mesh_file", mesh_file);
  d.get("ic_form", ic_form);
  d.get("test_case", test_case);
  d.get("n_steps", n_steps);
  d.get("restart_flag", restart_flag);
  if (restart_flag == 1) {
    d.get("restart_iter", restart_iter);
    d.get("n_restart_files", n_restart_files);
  }
  d.get("plot_freq", plot_freq);
  d.get("data_file_name", data_file_name);
  d.get("restart_dump_freq", restart_dump_freq);
  d.get("monitor_res_freq", monitor_res_freq);
  d.get("res_norm_type", res_norm_type);
  d.get("volume_cub_order", volume_cub_order);
  integral_quantities.clear();
  if (d.has("integral_quantities")) {
    // "integral_quantities n name_1 .. name_n", names case-insensitive
    std::istringstream ss(d.kv.at("integral_quantities"));
    int n = 0;
    if (!(ss >> n) || n < 0) return "Unable to read number of entries for vector option integral_quantities";
    for (int i = 0; i < n; i++) {
      std::string q;
      if (!(ss >> q)) return "Unable to assign all values to vector option integral_quantities";
      for (char& ch : q) ch = (char)std::tolower((unsigned char)ch);
      integral_quantities.push_back(q);
    }
  }
  d.get("riemann_solve_type", riemann_solve_type);
  d.get("vis_riemann_solve_type", vis_riemann_solve_type);
  d.get("adv_type", adv_type);
  d.get("dt_type", dt_type);
  if (dt_type == 0) d.get("dt", dt);
  else d.get("CFL", CFL);
  d.get("turb_model", turb_model);
  d.get("LES", LES);
  if (LES) {
    d.get("filter_ratio", filter_ratio);
    d.get("SGS_model", SGS_model);
    d.get("wall_model", wall_model);
  }
  d.get("motion_flag", motion);
  d.get("gamma", gamma);
  d.get("prandtl", prandtl);
  d.get("S_gas", S_gas);
  d.get("T_gas", T_gas);
  d.get("R_gas", R_gas);
  d.get("mu_gas", mu_gas);
  d.get("dx_cyclic", dx_cyclic);
  d.get("dy_cyclic", dy_cyclic);
  d.get("dz_cyclic", dz_cyclic);
  d.get("Mach_free_stream", Mach_free_stream);
  d.get("nx_free_stream", nx_free_stream);
  d.get("ny_free_stream", ny_free_stream);
  d.get("nz_free_stream", nz_free_stream);
  d.get("Re_free_stream", Re_free_stream);
  d.get("L_free_stream", L_free_stream);
  d.get("T_free_stream", T_free_stream);
  d.get("rho_bound", rho_bound);
  d.get("u_bound", v_bound[0]);
  d.get("v_bound", v_bound[1]);
  d.get("w_bound", v_bound[2]);
  d.get("p_bound", p_bound);
  d.get("Mach_wall", Mach_wall);
  d.get("nx_wall", nx_wall);
  d.get("ny_wall", ny_wall);
  d.get("nz_wall", nz_wall);
  d.get("T_wall", T_wall);
  d.get("fix_vis", fix_vis);
  d.get("tau", tau);
  d.get("pen_fact", pen_fact);
  Mach_c_ic = Mach_free_stream; nx_c_ic = nx_free_stream; ny_c_ic = ny_free_stream; nz_c_ic = nz_free_stream;
  Re_c_ic = Re_free_stream; T_c_ic = T_free_stream; rho_c_ic = rho_bound;
  u_c_ic = v_bound[0]; v_c_ic = v_bound[1]; w_c_ic = v_bound[2]; p_c_ic = p_bound;
  d.get("Mach_c_ic", Mach_c_ic);
  d.get("nx_c_ic", nx_c_ic);
  d.get("ny_c_ic", ny_c_ic);
  d.get("nz_c_ic", nz_c_ic);
  d.get("Re_c_ic", Re_c_ic);
  d.get("T_c_ic", T_c_ic);
  d.get("rho_c_ic", rho_c_ic);
  d.get("u_c_ic", u_c_ic);
  d.get("v_c_ic", v_c_ic);
  d.get("w_c_ic", w_c_ic);
  d.get("p_c_ic", p_c_ic);
  d.get("upts_type_quad", upts_type_quad);
  d.get("vcjh_scheme_quad", vcjh_scheme_quad);
  d.get("eta_quad", eta_quad);
  d.get("upts_type_hexa", upts_type_hexa);
  d.get("vcjh_scheme_hexa", vcjh_scheme_hexa);
  d.get("eta_hexa", eta_hexa);
  d.get("upts_type_tet", upts_type_tet);
  d.get("fpts_type_tet", fpts_type_tet);
  d.get("vcjh_scheme_tet", vcjh_scheme_tet);
  d.get("c_tet", c_tet);
  d.get("eta_tet", eta_tet);
  d.get("sparse_tet", sparse_tet);
  d.get("const_src", const_src);
  d.get("hifr_host_only", host_only);
  d.get("hifr_metrics", metric_storage);
  d.get("hifr_precision", precision);
  d.get("hifr_learned_sgs", learned_sgs);
  return "";
}

void input::setup_params() {
  if (!viscous) return;
  if (ic_form == 0) {  // isentropic vortex (src/input.cpp:392-399)
    fix_vis = 1.;
    R_ref = 1.;
    c_sth = 1.;
    rt_inf = 1.;
    mu_inf = 0.1;
    return;
  }
  T_ref = T_free_stream;
  L_ref = L_free_stream;
  uvw_ref = Mach_free_stream * std::sqrt(gamma * R_gas * T_free_stream);
  double mu_free_stream;
  if (fix_vis) mu_free_stream = mu_gas;
  else mu_free_stream = mu_gas * std::pow(T_free_stream / T_gas, 1.5) * ((T_gas + S_gas) / (T_free_stream + S_gas));
  const double rho_free_stream = (mu_free_stream * Re_free_stream) / (uvw_ref * L_free_stream);
  const double p_free_stream = rho_free_stream * R_gas * T_free_stream;
  rho_ref = rho_free_stream;
  p_ref = rho_ref * uvw_ref * uvw_ref;
  mu_ref = rho_ref * uvw_ref * L_ref;
  time_ref = L_ref / uvw_ref;
  R_ref = (R_gas * T_ref) / (uvw_ref * uvw_ref);
  c_sth = S_gas / T_gas;
  mu_inf = mu_gas / mu_ref;
  rt_inf = T_gas * R_gas / (uvw_ref * uvw_ref);
  rho_bound = 1.;
  v_bound[0] = uvw_ref * nx_free_stream / uvw_ref;
  v_bound[1] = uvw_ref * ny_free_stream / uvw_ref;
  v_bound[2] = uvw_ref * nz_free_stream / uvw_ref;
  p_bound = p_free_stream / p_ref;
  const double uvw_wall = Mach_wall * std::sqrt(gamma * R_gas * T_wall);
  v_wall[0] = (uvw_wall * nx_wall) / uvw_ref;
  v_wall[1] = (uvw_wall * ny_wall) / uvw_ref;
  v_wall[2] = (uvw_wall * nz_wall) / uvw_ref;
  T_wall = T_wall / T_ref;
  const double uvw_c_ic = Mach_c_ic * std::sqrt(gamma * R_gas * T_c_ic);
  u_c_ic = (uvw_c_ic * nx_c_ic) / uvw_ref;
  v_c_ic = (uvw_c_ic * ny_c_ic) / uvw_ref;
  w_c_ic = (uvw_c_ic * nz_c_ic) / uvw_ref;
  double mu_c_ic;
  if (fix_vis) mu_c_ic = mu_gas;
  else mu_c_ic = mu_gas * std::pow(T_c_ic / T_gas, 1.5) * ((T_gas + S_gas) / (T_c_ic + S_gas));
  rho_c_ic = (mu_c_ic * Re_c_ic) / (uvw_c_ic * L_ref);
  p_c_ic = rho_c_ic * R_gas * T_c_ic;
  rho_c_ic = rho_c_ic / rho_ref;
  p_c_ic = p_c_ic / p_ref;
  T_c_ic = T_c_ic / T_ref;
  prandtl_t = 0.9;
}

}  // namespace hifr_host
