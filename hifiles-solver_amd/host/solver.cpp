// Host driver of the solver: GeoPreprocess, InitSolution, CalcResidual,
// AdvanceSolution, CalcNormResidual and restart I/O (src/solver.cpp,
// src/HiFiLES.cpp, src/output.cpp:1057-1107, 1452-1616), over the device ABI.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>

#include "../../include/hifr_solver.h"
#include "eles.h"
#include "geometry.h"
#include "input.h"

using namespace hifr_host;

struct hifr_solver {
  input in;
  mesh msh;
  std::unique_ptr<eles> el;
  int_inters ii;
  double time = 0.0;
  std::vector<double> u_host;  // the state of a host-only solver (no device context)
  std::vector<int> integral_kinds;  // the deck's integral_quantities as HIFR_* kinds
};

namespace {

int fail(const std::string& msg, char* err, int errlen) {
  if (err && errlen > 0) {
    std::snprintf(err, errlen, "%s", msg.c_str());
  }
  return 1;
}

// dimension and linear shape-node count of quads (1), tets (2) and hexes (4)
bool mesh_shape(int ele_type, int& nd, int& nspts) {
  if (ele_type == 1) { nd = 2; nspts = 4; return true; }
  if (ele_type == 2) { nd = 3; nspts = 4; return true; }
  if (ele_type == 4) { nd = 3; nspts = 8; return true; }
  return false;
}

// section tag of an element type in restart files (src/eles_*.cpp write_restart_info)
const char* ele_tag(int ele_type) { return ele_type == 4 ? "HEXAS" : ele_type == 2 ? "TETS" : "QUADS"; }

std::string check(hifr_status s, const char* what) {
  if (s == HIFR_OK) return "";
  return std::string(what) + ": " + hifr_status_string(s);
}

// src/funcs.cpp:1603-1659 (isentropic vortex advected at (1,1), eps = 5)
void eval_isentropic_vortex(const double* pos, double t, double& rho, double& vx, double& vy, double& p) {
  const double pi = 4.0 * std::atan(1.0), gamma = 1.4, eps = 5.0;
  const double x = pos[0] - t, y = pos[1] - t;
  const double f = 1.0 - (x * x + y * y);
  rho = std::pow(1.0 - eps * eps * (gamma - 1.0) / (8.0 * gamma * pi * pi) * std::exp(f), 1.0 / (gamma - 1.0));
  vx = 1. - eps * y / (2.0 * pi) * std::exp(f / 2.0);
  vy = 1. + eps * x / (2.0 * pi) * std::exp(f / 2.0);
  p = std::pow(rho, gamma);
}

// eles::set_ics (src/eles.cpp:434-620) for ic_form 0, 1, 7
std::string set_ics(hifr_solver* s, std::vector<double>& u) {
  eles& e = *s->el;
  const input& in = s->in;
  const int nu = e.n_upts_per_ele, N = e.n_eles, nf = e.n_fields, nd = e.n_dims;
  u.assign((size_t)nu * N * nf, 0.0);
  const double gamma = in.gamma;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < nu; j++) {
      double pos[3] = {0, 0, 0};
      for (int k = 0; k < nd; k++) pos[k] = e.pos_upts[j + (size_t)nu * (i + (size_t)N * k)];
      double ics[5] = {0, 0, 0, 0, 0};
      if (in.ic_form == 0) {
        double rho, vx, vy, p;
        eval_isentropic_vortex(pos, 0.0, rho, vx, vy, p);
        ics[0] = rho; ics[1] = rho * vx; ics[2] = rho * vy;
        if (nd == 2) ics[3] = (p / (gamma - 1.0)) + (0.5 * rho * ((vx * vx) + (vy * vy)));
        else { ics[3] = 0.0; ics[4] = (p / (gamma - 1.0)) + (0.5 * rho * ((vx * vx) + (vy * vy))); }
      } else if (in.ic_form == 1) {
        const double rho = in.rho_c_ic, vx = in.u_c_ic, vy = in.v_c_ic, vz = in.w_c_ic, p = in.p_c_ic;
        ics[0] = rho; ics[1] = rho * vx; ics[2] = rho * vy;
        if (nd == 2) ics[3] = (p / (gamma - 1.0)) + (0.5 * rho * ((vx * vx) + (vy * vy)));
        else { ics[3] = rho * vz; ics[4] = (p / (gamma - 1.0)) + (0.5 * rho * ((vx * vx) + (vy * vy) + (vz * vz))); }
      } else if (in.ic_form == 7) {
        const double rho = in.rho_c_ic;
        ics[0] = rho;
        if (nd == 2) {
          const double p = 100 + rho / 16.0 * (std::cos(2.0 * pos[0]) + std::cos(2.0 * pos[1])) * (std::cos(2.0 * pos[2]) + 2.0);
          ics[1] = std::sin(pos[0] / 2.) * std::cos(pos[1] / 2.);
          ics[2] = -1.0 * std::cos(pos[0] / 2.) * std::sin(pos[1] / 2.);
          ics[3] = p / (gamma - 1.0) + 0.5 * rho * (ics[1] * ics[1] + ics[2] * ics[2]);
        } else {
          const double p = 100 + rho / 16.0 * (std::cos(2.0 * pos[0]) + std::cos(2.0 * pos[1])) * (std::cos(2.0 * pos[2]) + 2.0);
          ics[1] = std::sin(pos[0]) * std::cos(pos[1]) * std::cos(pos[2]);
          ics[2] = -1.0 * std::cos(pos[0]) * std::sin(pos[1]) * std::cos(pos[2]);
          ics[3] = 0.0;
          ics[4] = p / (gamma - 1.0) + 0.5 * rho * (ics[1] * ics[1] + ics[2] * ics[2] + ics[3] * ics[3]);
        }
      } else {
        return "initial condition form not supported (ic_form 0, 1, 7)";
      }
      for (int k = 0; k < nf; k++) u[j + (size_t)nu * (i + (size_t)N * k)] = ics[k];
    }
  return "";
}

// Reads a restart file (src/solver.cpp:390-444, src/eles.cpp:623-716); the file's
// solution points must be the solver's (same order and 1-D points).
std::string read_restart_file(hifr_solver* s, const char* name, std::vector<double>& u) {
  std::ifstream f(name);
  if (!f) return std::string("Could not open restart file ") + name;
  eles& e = *s->el;
  const int nu = e.n_upts_per_ele, N = e.n_eles, nf = e.n_fields;
  f >> s->time;
  std::string line;
  const char* tag = ele_tag(e.ele_type);
  while (std::getline(f, line) && line != tag) {}
  int order_rest = -1, nu_rest = -1;
  std::getline(f, line);
  f >> order_rest;
  std::getline(f, line);
  std::getline(f, line);
  f >> nu_rest;
  std::getline(f, line);
  std::getline(f, line);
  if (order_rest != e.order || nu_rest != nu)
    return "restart file has other solution points (interpolating restarts are not supported)";
  // quads/hexes list the 1-D points, tets every solution point (src/eles_tets.cpp:749-780)
  const std::vector<double>& pts = e.ele_type == 2 ? e.loc_upts : e.loc_1d_upts;
  for (size_t i = 0; i < pts.size(); i++) {
    double x;
    f >> x;
    if (std::fabs(x - pts[i]) > 1e-14)
      return "restart file has other solution points (interpolating restarts are not supported)";
  }
  while (std::getline(f, line) && line != "n_eles") {}
  int nread = 0;
  f >> nread;
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, line);
  std::getline(f, line);
  u.assign((size_t)nu * N * nf, 0.0);
  for (int i = 0; i < nread; i++) {
    int ele;
    f >> ele;
    if (ele < 0 || ele >= N) return "restart element index out of range";
    for (int j = 0; j < nu; j++)
      for (int k = 0; k < nf; k++) f >> u[j + (size_t)nu * (ele + (size_t)N * k)];
  }
  if (!f) return "error reading restart data";
  return "";
}

// Rest_<restart_iter>_p0000.dat in the working directory, as the reference reads it
std::string read_restart(hifr_solver* s, std::vector<double>& u) {
  char name[64];
  std::snprintf(name, sizeof name, "Rest_%09d_p%04d.dat", s->in.restart_iter, 0);
  return read_restart_file(s, name, u);
}

// GeoPreprocess + InitSolution on an already filled mesh
std::string setup_solver(hifr_solver* s) {
  input& in = s->in;
  if (in.equation != 0) return "only the Euler / Navier-Stokes equations are built";
  if (in.motion != 0) return "moving meshes are not built";
  if (in.turb_model != 0) return "turbulence models are not built";
  if (in.LES != 0) {
    if (!in.viscous) return "LES needs viscous 1";
    if (in.SGS_model != 0 && in.SGS_model != 1) return "only the Smagorinsky (0) and WALE (1) SGS models are built";
    if (in.wall_model != 0) return "LES wall models are not built";
    if (in.fix_vis != 1.0) return "LES is built for constant viscosity (fix_vis 1) only";
    if (in.learned_sgs && in.SGS_model != 1) return "the learned SGS closure falls back to WALE: set SGS_model 1";
  } else if (in.learned_sgs) {
    return "hifr_learned_sgs needs LES 1";
  }
  if (in.dt_type != 0) return "only user-supplied time steps (dt_type 0) are built";
  in.setup_params();
  const mesh& m = s->msh;
  if (m.ele_type == 4) s->el.reset(new eles_hexas());
  else if (m.ele_type == 1) s->el.reset(new eles_quads());
  else if (m.ele_type == 2) s->el.reset(new eles_tets());
  else return "mesh element type not supported";
  eles& e = *s->el;
  e.ele_type = m.ele_type;
  if (in.learned_sgs && !(m.ele_type == 4 && (in.order == 2 || in.order == 3)))
    return "the learned SGS closure is built for hexes of order 2 and 3";
  std::string err = e.setup(in, m.n_eles);
  if (!err.empty()) return err;
  for (int i = 0; i < m.n_eles; i++)
    for (int j = 0; j < m.n_spts; j++) e.set_shape_node(j, i, &m.xv[(size_t)m.c2v[(size_t)i * m.n_spts + j] * m.n_dims]);
  e.set_transforms();
  const double delta[3] = {in.dx_cyclic, in.dy_cyclic, in.dz_cyclic};
  err = build_interfaces(e, m, delta, s->ii);
  if (!err.empty()) return err;
  for (int t : s->ii.bdy_type)
    if (t != 11) return "boundary condition " + std::to_string(t) + " is not built (isothermal walls, bctype 11)";
  // integral diagnostics (src/eles.cpp:6503-6643): names -> kinds, volume cubature
  s->integral_kinds.clear();
  for (const std::string& q : in.integral_quantities) {
    static const char* names[] = {"kineticenergy", "vorticity", "pressuredilatation", "straincolonproduct",
                                  "devstraincolonproduct"};
    int k = -1;
    for (int i = 0; i < 5; i++)
      if (q == names[i]) k = i;
    if (k < 0) return "integral diagnostic quantity not recognized";
    s->integral_kinds.push_back(k);
  }
  if (s->integral_kinds.size() > 5) return "at most 5 integral quantities";
  if (!s->integral_kinds.empty()) {
    err = e.set_volume_cubature(in.volume_cub_order);
    if (!err.empty()) return err;
  }
  if (in.host_only) {
    // the initial state stays on the host
    err = in.restart_flag ? read_restart(s, s->u_host) : set_ics(s, s->u_host);
    return err;
  }
  err = e.mv_all_cpu_gpu(in);
  if (!err.empty()) return err;
  if (!s->ii.bdy_fpt.empty()) {
    err = check(hifr_set_boundaries(e.ctx, (int)s->ii.bdy_fpt.size(), s->ii.bdy_fpt.data(), s->ii.bdy_type.data()),
                "hifr_set_boundaries");
    if (!err.empty()) return err;
  }
  err = check(hifr_set_interfaces(e.ctx, (int)s->ii.left.size(), s->ii.left.data(), s->ii.right.data()),
              "hifr_set_interfaces");
  if (!err.empty()) return err;
  if (!s->integral_kinds.empty()) {
    err = check(hifr_set_cubature(e.ctx, e.n_cubpts_per_ele, e.opp_volume_cubpts.data(),
                                  e.weight_volume_cubpts.data(), e.vol_detjac_cubpts.data()),
                "hifr_set_cubature");
    if (!err.empty()) return err;
  }
  std::vector<double> u;
  err = in.restart_flag ? read_restart(s, u) : set_ics(s, u);
  if (!err.empty()) return err;
  return check(hifr_upload_state(e.ctx, u.data()), "hifr_upload_state");
}

std::string read_text_deck(const char* text, input& in) {
  char tmpl[] = "/tmp/hifr_deck_XXXXXX";
  const int fd = mkstemp(tmpl);
  if (fd < 0) return "cannot create a temporary deck file";
  FILE* f = fdopen(fd, "w");
  std::fputs(text, f);
  std::fclose(f);
  std::string err = in.read_input_file(tmpl);
  std::remove(tmpl);
  return err;
}

}  // namespace

extern "C" {

int hifr_solver_from_deck(const char* path, hifr_solver** out, char* err, int errlen) {
  *out = nullptr;
  std::unique_ptr<hifr_solver> s(new hifr_solver());
  std::string e = s->in.read_input_file(path);
  if (!e.empty()) return fail(e, err, errlen);
  std::string mpath = s->in.mesh_file;
  const std::string dp(path);
  const size_t slash = dp.find_last_of('/');
  if (!mpath.empty() && mpath[0] != '/' && slash != std::string::npos) mpath = dp.substr(0, slash + 1) + mpath;
  e = read_mesh(mpath, s->msh);
  if (!e.empty()) return fail(e, err, errlen);
  e = setup_solver(s.get());
  if (!e.empty()) return fail(e, err, errlen);
  *out = s.release();
  return 0;
}

int hifr_solver_from_arrays(const char* deck_text, int ele_type, int n_verts, const double* xv, int n_eles,
                            const int* c2v, hifr_solver** out, char* err, int errlen) {
  *out = nullptr;
  std::unique_ptr<hifr_solver> s(new hifr_solver());
  std::string e = read_text_deck(deck_text, s->in);
  if (!e.empty()) return fail(e, err, errlen);
  mesh& m = s->msh;
  m.ele_type = ele_type;
  if (!mesh_shape(ele_type, m.n_dims, m.n_spts)) return fail("element type not supported", err, errlen);
  m.n_verts = n_verts;
  m.n_eles = n_eles;
  m.xv.assign(xv, xv + (size_t)n_verts * m.n_dims);
  m.c2v.assign(c2v, c2v + (size_t)n_eles * m.n_spts);
  e = setup_solver(s.get());
  if (!e.empty()) return fail(e, err, errlen);
  *out = s.release();
  return 0;
}

int hifr_solver_from_arrays_bc(const char* deck_text, int ele_type, int n_verts, const double* xv, int n_eles,
                               const int* c2v, int n_bfaces, const int* bfaces, hifr_solver** out, char* err,
                               int errlen) {
  *out = nullptr;
  std::unique_ptr<hifr_solver> s(new hifr_solver());
  std::string e = read_text_deck(deck_text, s->in);
  if (!e.empty()) return fail(e, err, errlen);
  mesh& m = s->msh;
  m.ele_type = ele_type;
  if (!mesh_shape(ele_type, m.n_dims, m.n_spts)) return fail("element type not supported", err, errlen);
  m.n_verts = n_verts;
  m.n_eles = n_eles;
  m.xv.assign(xv, xv + (size_t)n_verts * m.n_dims);
  m.c2v.assign(c2v, c2v + (size_t)n_eles * m.n_spts);
  for (int q = 0; q < n_bfaces; q++) {
    m.bf_ele.push_back(bfaces[3 * q]);
    m.bf_face.push_back(bfaces[3 * q + 1]);
    m.bf_type.push_back(bfaces[3 * q + 2]);
  }
  e = setup_solver(s.get());
  if (!e.empty()) return fail(e, err, errlen);
  *out = s.release();
  return 0;
}

int hifr_solver_box(const char* deck_text, int ele_type, const int* n, const double* l, const double* o, int digits,
                    hifr_solver** out, char* err, int errlen) {
  *out = nullptr;
  std::unique_ptr<hifr_solver> s(new hifr_solver());
  std::string e = read_text_deck(deck_text, s->in);
  if (!e.empty()) return fail(e, err, errlen);
  if (ele_type != 1 && ele_type != 4) return fail("the box generator builds quads and hexes", err, errlen);
  make_box(ele_type, n, l, o, digits, s->msh);
  e = setup_solver(s.get());
  if (!e.empty()) return fail(e, err, errlen);
  *out = s.release();
  return 0;
}

void hifr_solver_destroy(hifr_solver* s) { delete s; }

void hifr_solver_sizes(const hifr_solver* s, int* o) {
  const eles& e = *s->el;
  o[0] = e.n_upts_per_ele;
  o[1] = e.n_fpts_per_ele;
  o[2] = e.n_fields;
  o[3] = e.n_dims;
  o[4] = e.n_eles;
  o[5] = s->ii.n_inters;
}

long hifr_solver_get(hifr_solver* s, int what, double* out) {
  eles& e = *s->el;
  const long nu = e.n_upts_per_ele, nfp = e.n_fpts_per_ele, nf = e.n_fields, nd = e.n_dims, N = e.n_eles;
  const std::vector<double>* v = nullptr;
  long n = 0;
  int dev = -1;
  switch (what) {
    case 0: dev = HIFR_DISU_UPTS; n = nu * N * nf; break;
    case 1: dev = HIFR_DIV_TCONF_UPTS; n = nu * N * nf; break;
    case 2: dev = HIFR_DISU_FPTS; n = nfp * N * nf; break;
    case 3: dev = HIFR_GRAD_DISU_FPTS; n = s->in.viscous ? nfp * N * nf * nd : 0; break;
    case 4: v = &e.detjac_upts; break;
    case 5: v = &e.pos_upts; break;
    case 6: v = &e.opp_0; break;
    case 7: v = &e.opp_3; break;
    case 8: v = &e.opp_4; break;
    case 9: v = &e.JGinv_upts; break;
    case 10: v = &e.tdA_fpts; break;
    case 11: v = &e.norm_fpts; break;
    case 12: v = &s->msh.xv; break;
    case 13:
      n = (long)s->msh.c2v.size();
      if (out)
        for (long i = 0; i < n; i++) out[i] = s->msh.c2v[i];
      return n;
    case 14: dev = HIFR_RK_REGISTER; n = nu * N * nf; break;
    case 15: v = &e.loc_volume_cubpts; break;
    case 16: v = &e.weight_volume_cubpts; break;
    case 17: v = &e.opp_volume_cubpts; break;
    case 18: v = &e.vol_detjac_cubpts; break;
    default: return -1;
  }
  if (v) {
    n = (long)v->size();
    if (out) std::memcpy(out, v->data(), sizeof(double) * n);
    return n;
  }
  if (!e.ctx) {
    // host-only solver: the state it holds
    if (dev != HIFR_DISU_UPTS || s->u_host.size() != (size_t)n) return -1;
    if (out) std::memcpy(out, s->u_host.data(), sizeof(double) * n);
    return n;
  }
  if (out && n > 0 && hifr_download(e.ctx, dev, out) != HIFR_OK) return -1;
  return n;
}

int hifr_solver_set_state(hifr_solver* s, const double* u) {
  if (!s->el->ctx) {
    const eles& e = *s->el;
    s->u_host.assign(u, u + (size_t)e.n_upts_per_ele * e.n_eles * e.n_fields);
    return 0;
  }
  return hifr_upload_state(s->el->ctx, u) == HIFR_OK ? 0 : 1;
}

int hifr_solver_read_restart(hifr_solver* s, const char* path, char* err, int errlen) {
  std::vector<double> u;
  const double t0 = s->time;
  const std::string e = read_restart_file(s, path, u);
  if (!e.empty()) {
    s->time = t0;
    return fail(e, err, errlen);
  }
  if (hifr_solver_set_state(s, u.data()) != 0) {
    s->time = t0;
    return fail("hifr_upload_state failed", err, errlen);
  }
  return 0;
}

double hifr_solver_time(const hifr_solver* s) { return s->time; }

int hifr_solver_n_integrals(const hifr_solver* s) { return (int)s->integral_kinds.size(); }

long hifr_solver_n_boundary_points(const hifr_solver* s) { return (long)s->ii.bdy_fpt.size(); }

int hifr_solver_integrals(hifr_solver* s, double* out) {
  if (!s->el->ctx) return 1;
  if (s->integral_kinds.empty()) return 0;
  const hifr_status st =
      hifr_integral_quantities(s->el->ctx, (int)s->integral_kinds.size(), s->integral_kinds.data(), out);
  return st == HIFR_OK ? 0 : (int)st;
}

long hifr_solver_interfaces(const hifr_solver* s, int* left, int* right) {
  const long n = (long)s->ii.left.size();
  if (left) std::memcpy(left, s->ii.left.data(), sizeof(int) * n);
  if (right) std::memcpy(right, s->ii.right.data(), sizeof(int) * n);
  return n;
}

int hifr_solver_residual(hifr_solver* s) {
  if (!s->el->ctx) return 1;
  return hifr_residual(s->el->ctx) == HIFR_OK ? 0 : 1; }

int hifr_solver_advance(hifr_solver* s, int stage) {
  if (!s->el->ctx) return 1;
  return hifr_advance(s->el->ctx, stage, s->in.adv_type, s->in.dt) == HIFR_OK ? 0 : 1;
}

int hifr_solver_run(hifr_solver* s, int n_steps) {
  if (!s->el->ctx) return 1;
  if (hifr_run(s->el->ctx, n_steps, s->in.adv_type, s->in.dt) != HIFR_OK) return 1;
  s->time += n_steps * s->in.dt;
  return 0;
}

int hifr_solver_norms(hifr_solver* s, int norm_type, double* out) {
  if (!s->el->ctx) return 1;
  const hifr_status st = hifr_norm_residual(s->el->ctx, norm_type, out);
  return st == HIFR_OK ? 0 : (int)st;
}

int hifr_solver_write_restart(hifr_solver* s, const char* path) {
  eles& e = *s->el;
  const int nu = e.n_upts_per_ele, N = e.n_eles, nf = e.n_fields;
  std::vector<double> u((size_t)nu * N * nf);
  if (!e.ctx) {
    if (s->u_host.size() != u.size()) return 1;
    u = s->u_host;
  } else if (hifr_download(e.ctx, HIFR_DISU_UPTS, u.data()) != HIFR_OK) {
    return 1;
  }
  std::ofstream f(path);
  if (!f) return 1;
  f.precision(15);
  f << s->time << "\n";
  f << ele_tag(e.ele_type) << "\nOrder\n" << e.order << "\n";
  if (e.ele_type == 2) {
    // src/eles_tets.cpp:782-800
    f << "Number of solution points per element\n" << nu << "\nLocation of solution points in tetrahedral elements\n";
    for (int j = 0; j < nu; j++) {
      for (int k = 0; k < 3; k++) f << e.loc_upts[3 * j + k] << " ";
      f << "\n";
    }
  } else {
    f << "Number of solution points per " << (e.ele_type == 4 ? "hexahedral" : "quadrilateral") << " element\n"
      << nu << "\nLocation of solution points in 1D\n";
    for (double x : e.loc_1d_upts) f << x << " ";
    f << "\n";
  }
  f << "n_eles\n" << N << "\nele2global_ele array\n";
  for (int i = 0; i < N; i++) f << i << " ";
  f << "\ndata\n";
  for (int i = 0; i < N; i++) {
    f << i << "\n";
    for (int j = 0; j < nu; j++) {
      for (int k = 0; k < nf; k++) f << u[j + (size_t)nu * (i + (size_t)N * k)] << " ";
      f << "\n";
    }
  }
  f << "\n";
  return f ? 0 : 1;
}

hifr_ctx* hifr_solver_ctx(hifr_solver* s) { return s->el->ctx; }
double hifr_solver_dt(const hifr_solver* s) { return s->in.dt; }
int hifr_solver_adv_type(const hifr_solver* s) { return s->in.adv_type; }

}  // extern "C"
