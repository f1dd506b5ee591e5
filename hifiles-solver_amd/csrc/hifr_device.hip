// C ABI of the gfx950 FR residual / RK update (include/hifr.h): context, device
// buffers, operator factorisation, interface tables and kernel dispatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/hifr.h"
#include "fr_tensor_kernels.hpp"

using hifr::Ops1D;
using hifr::StageArgs;

struct hifr_ctx {
  hifr_config cfg;
  int nd, nf, n1, nu, nfp, nfpf, N;
  hipStream_t stream;
  Ops1D ops;
  bool have_ops, have_metrics, have_faces, have_state;
  // device buffers
  double *u, *res, *div, *uf[2], *fv, *gf, *detj, *jg, *fgeo;
  int* nbr;
  double* red;  // reduction scratch
  int cur;      // which uf buffer holds the traces of the current state
  // host copies needed to build the interface tables
  std::vector<double> h_tdA, h_norm;
  // timing
  bool timing;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
  size_t ev_used;
};

namespace {

const int EV_POOL = 8192;

hifr_status dev_err(hipError_t e, const char* what) {
  if (e == hipSuccess) return HIFR_OK;
  std::fprintf(stderr, "hifr: %s: %s\n", what, hipGetErrorString(e));
  return HIFR_ERR_DEVICE;
}

#define HIP_TRY(x)                                     \
  do {                                                 \
    hifr_status _s = dev_err((x), #x);                 \
    if (_s != HIFR_OK) return _s;                      \
  } while (0)

// ---------------------------------------------------------------------------
// launch helpers

template <int P, int ND>
struct Cfg {
  // elements per workgroup: about 128 threads per block
  static constexpr int EPB = hifr::TP<P, ND, 1>::NU >= 128 ? 1 : 128 / hifr::TP<P, ND, 1>::NU;
};

template <int P, int ND, int VIS, int MODE>
hipError_t launch_one(const StageArgs& a, hipStream_t s) {
  constexpr int EPB = Cfg<P, ND>::EPB;
  using T = hifr::TP<P, ND, EPB>;
  const int grid = (a.N + EPB - 1) / EPB;
  hipLaunchKernelGGL((hifr::fr_tensor_kernel<P, ND, EPB, VIS, MODE>), dim3(grid), dim3(T::BLOCK), 0, s, a);
  return hipGetLastError();
}

// vis: 0 Euler, 1 Navier-Stokes with constant viscosity, 2 Sutherland viscosity
template <int P, int ND>
hipError_t launch_mode(int mode, int vis, const StageArgs& a, hipStream_t s) {
  using namespace hifr;
  switch (mode) {
    case MODE_TRACE_U: return launch_one<P, ND, 0, MODE_TRACE_U>(a, s);
    case MODE_TRACE_GRAD:
      return vis == 2 ? launch_one<P, ND, 2, MODE_TRACE_GRAD>(a, s) : launch_one<P, ND, 1, MODE_TRACE_GRAD>(a, s);
    case MODE_RESID:
      return vis == 2 ? launch_one<P, ND, 2, MODE_RESID>(a, s)
           : vis == 1 ? launch_one<P, ND, 1, MODE_RESID>(a, s)
                      : launch_one<P, ND, 0, MODE_RESID>(a, s);
    default:
      return vis == 2 ? launch_one<P, ND, 2, MODE_ADVANCE>(a, s)
           : vis == 1 ? launch_one<P, ND, 1, MODE_ADVANCE>(a, s)
                      : launch_one<P, ND, 0, MODE_ADVANCE>(a, s);
  }
}

bool order_supported(int ele_type, int P) {
  if (ele_type == HIFR_HEX) return P >= 1 && P <= 5;
  if (ele_type == HIFR_QUAD) return P >= 1 && P <= 6;
  return false;
}

hipError_t launch(hifr_ctx* c, int mode, const StageArgs& a) {
  const int v = c->cfg.viscous == 0 ? 0 : (c->cfg.fix_vis == 1.0 ? 1 : 2);
  const int P = c->cfg.order;
  if (c->nd == 3) {
    switch (P) {
      case 1: return launch_mode<1, 3>(mode, v, a, c->stream);
      case 2: return launch_mode<2, 3>(mode, v, a, c->stream);
      case 3: return launch_mode<3, 3>(mode, v, a, c->stream);
      case 4: return launch_mode<4, 3>(mode, v, a, c->stream);
      default: return launch_mode<5, 3>(mode, v, a, c->stream);
    }
  }
  switch (P) {
    case 1: return launch_mode<1, 2>(mode, v, a, c->stream);
    case 2: return launch_mode<2, 2>(mode, v, a, c->stream);
    case 3: return launch_mode<3, 2>(mode, v, a, c->stream);
    case 4: return launch_mode<4, 2>(mode, v, a, c->stream);
    case 5: return launch_mode<5, 2>(mode, v, a, c->stream);
    default: return launch_mode<6, 2>(mode, v, a, c->stream);
  }
}

StageArgs base_args(hifr_ctx* c) {
  StageArgs a;
  std::memset(&a, 0, sizeof(a));
  a.N = c->N;
  a.u = c->u;
  a.res = c->res;
  a.detj = c->detj;
  a.jg = c->jg;
  a.nbr = c->nbr;
  a.fgeo = c->fgeo;
  a.ops = c->ops;
  a.ph.gamma = c->cfg.gamma;
  a.ph.prandtl = c->cfg.prandtl;
  a.ph.mu_inf = c->cfg.mu_inf;
  a.ph.rt_inf = c->cfg.rt_inf;
  a.ph.c_sth = c->cfg.c_sth;
  a.ph.fix_vis = c->cfg.fix_vis;
  a.ph.tau = c->cfg.tau;
  a.ph.pen_fact = c->cfg.pen_fact;
  a.ph.const_src = c->cfg.const_src;
  return a;
}

hifr_status timed_launch(hifr_ctx* c, int mode, int kind, const StageArgs& a) {
  if (c->timing) {
    if (c->ev_used + 2 > c->ev.size()) return HIFR_ERR_ARG;
    HIP_TRY(hipEventRecord(c->ev[c->ev_used], c->stream));
    c->ev_kind[c->ev_used / 2] = kind;
  }
  HIP_TRY(launch(c, mode, a));
  if (c->timing) {
    HIP_TRY(hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    c->ev_used += 2;
  }
  return HIFR_OK;
}

hifr_status trace_u(hifr_ctx* c) {
  StageArgs a = base_args(c);
  a.uf_out = c->uf[c->cur];
  return timed_launch(c, hifr::MODE_TRACE_U, 0, a);
}

// write_grad: also store the corrected gradient at the flux points (for
// hifr_download; the residual itself only needs the viscous normal fluxes)
hifr_status trace_grad(hifr_ctx* c, bool write_grad) {
  StageArgs a = base_args(c);
  a.uf_in = c->uf[c->cur];
  a.fv_out = c->fv;
  a.gf_out = write_grad ? c->gf : nullptr;
  return timed_launch(c, hifr::MODE_TRACE_GRAD, 1, a);
}

hifr_status check_ready(hifr_ctx* c) {
  if (!c) return HIFR_ERR_ARG;
  if (!c->have_ops || !c->have_metrics || !c->have_faces || !c->have_state) return HIFR_ERR_STATE;
  return HIFR_OK;
}

// ---------------------------------------------------------------------------
// operator factorisation

// Flux point -> (dir, side, first point, stride); host copy of face_line.
void host_face_line(int nd, int P, int fp, int& dir, int& side, int& base, int& st) {
  const int n = P + 1;
  if (nd == 3) {
    const int f = fp / (n * n), q = fp - f * n * n, k = q % n, j = q / n;
    switch (f) {
      case 0: dir = 2; side = 0; base = (P - k) + n * j; break;
      case 1: dir = 1; side = 0; base = k + n * n * j; break;
      case 2: dir = 0; side = 1; base = n * k + n * n * j; break;
      case 3: dir = 1; side = 1; base = (P - k) + n * n * j; break;
      case 4: dir = 0; side = 0; base = n * (P - k) + n * n * j; break;
      default: dir = 2; side = 1; base = k + n * j; break;
    }
  } else {
    const int f = fp / n, j = fp - f * n;
    switch (f) {
      case 0: dir = 1; side = 0; base = j; break;
      case 1: dir = 0; side = 1; base = n * j; break;
      case 2: dir = 1; side = 1; base = P - j; break;
      default: dir = 0; side = 0; base = n * (P - j); break;
    }
  }
  st = dir == 0 ? 1 : (dir == 1 ? n : n * n);
}

}  // namespace

extern "C" {

const char* hifr_status_string(hifr_status s) {
  switch (s) {
    case HIFR_OK: return "ok";
    case HIFR_ERR_ARG: return "invalid argument";
    case HIFR_ERR_UNSUPPORTED: return "unsupported configuration";
    case HIFR_ERR_DEVICE: return "HIP device error";
    case HIFR_ERR_NOT_TENSOR: return "operators are not tensor-product";
    case HIFR_ERR_STATE: return "inputs not set";
    case HIFR_ERR_NAN: return "NaN in residual";
  }
  return "unknown";
}

hifr_status hifr_create(const hifr_config* cfg, hifr_ctx** out) {
  if (!cfg || !out) return HIFR_ERR_ARG;
  *out = nullptr;
  if (cfg->ele_type != HIFR_HEX && cfg->ele_type != HIFR_QUAD) return HIFR_ERR_UNSUPPORTED;
  if (!order_supported(cfg->ele_type, cfg->order)) return HIFR_ERR_UNSUPPORTED;
  if (cfg->n_eles <= 0) return HIFR_ERR_ARG;
  if (cfg->riemann_solve_type != 0 || cfg->vis_riemann_solve_type != 0) return HIFR_ERR_UNSUPPORTED;
  if (cfg->precision != HIFR_FP64) return HIFR_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HIFR_ERR_DEVICE;
  hifr_ctx* c = new hifr_ctx();
  c->cfg = *cfg;
  c->nd = cfg->ele_type == HIFR_HEX ? 3 : 2;
  c->nf = c->nd + 2;
  c->n1 = cfg->order + 1;
  c->nu = c->nd == 3 ? c->n1 * c->n1 * c->n1 : c->n1 * c->n1;
  c->nfpf = c->nd == 3 ? c->n1 * c->n1 : c->n1;
  c->nfp = 2 * c->nd * c->nfpf;
  c->N = cfg->n_eles;
  c->stream = nullptr;
  c->cur = 0;
  c->timing = false;
  c->ev_used = 0;
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nf = c->nf, nd = c->nd;
  struct { void** p; size_t bytes; } allocs[] = {
      {(void**)&c->u, sizeof(double) * nu * N * nf},
      {(void**)&c->res, sizeof(double) * nu * N * nf},
      {(void**)&c->div, sizeof(double) * nu * N * nf},
      {(void**)&c->uf[0], sizeof(double) * nfp * N * nf},
      {(void**)&c->uf[1], sizeof(double) * nfp * N * nf},
      {(void**)&c->fv, cfg->viscous ? sizeof(double) * nfp * N * nf : 8},
      {(void**)&c->gf, cfg->viscous ? sizeof(double) * nfp * N * nf * nd : 8},
      {(void**)&c->detj, sizeof(double) * nu * N},
      {(void**)&c->jg, sizeof(double) * nd * nd * nu * N},
      {(void**)&c->fgeo, sizeof(double) * (nd + 1) * nfp * N},
      {(void**)&c->nbr, sizeof(int) * nfp * N},
      {(void**)&c->red, sizeof(double) * 4096},
  };
  for (auto& al : allocs) *al.p = nullptr;
  for (auto& al : allocs) {
    if (hipMalloc(al.p, al.bytes) != hipSuccess) {
      hifr_destroy(c);
      return HIFR_ERR_DEVICE;
    }
  }
  if (hipMemset(c->res, 0, sizeof(double) * nu * N * nf) != hipSuccess) {
    hifr_destroy(c);
    return HIFR_ERR_DEVICE;
  }
  *out = c;
  return HIFR_OK;
}

hifr_status hifr_destroy(hifr_ctx* c) {
  if (!c) return HIFR_OK;
  void* ps[] = {c->u, c->res, c->div, c->uf[0], c->uf[1], c->fv, c->gf, c->detj, c->jg, c->fgeo, c->nbr, c->red};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  delete c;
  return HIFR_OK;
}

hifr_status hifr_set_stream(hifr_ctx* c, void* s) {
  if (!c) return HIFR_ERR_ARG;
  c->stream = (hipStream_t)s;
  return HIFR_OK;
}

hifr_status hifr_sizes(const hifr_ctx* c, int* nu, int* nfp, int* nf, int* nd) {
  if (!c) return HIFR_ERR_ARG;
  if (nu) *nu = c->nu;
  if (nfp) *nfp = c->nfp;
  if (nf) *nf = c->nf;
  if (nd) *nd = c->nd;
  return HIFR_OK;
}

hifr_status hifr_set_operators(hifr_ctx* c, const double* opp_0, const double* opp_3, const double* opp_4) {
  if (!c || !opp_0 || !opp_3 || !opp_4) return HIFR_ERR_ARG;
  const int n = c->n1, P = c->cfg.order, nu = c->nu, nfp = c->nfp, nd = c->nd;
  Ops1D o;
  std::memset(&o, 0, sizeof(o));
  // flux points at the r=+1 and r=-1 ends of the line through solution point 0
  int fpR, fpL;
  if (nd == 3) { fpR = 2 * n * n; fpL = 4 * n * n + P; }
  else { fpR = n; fpL = 3 * n + P; }
  for (int a = 0; a < n; a++) {
    o.lR[a] = opp_0[fpR + (size_t)nfp * a];
    o.lL[a] = opp_0[fpL + (size_t)nfp * a];
    o.gR[a] = opp_3[a + (size_t)nu * fpR];
    o.gL[a] = -opp_3[a + (size_t)nu * fpL];
    for (int m = 0; m < n; m++) o.D[a][m] = opp_4[a + (size_t)nu * m];
  }
  // verify the dense operators against their tensor-product reconstruction
  double mx = 0.0, err = 0.0;
  auto chk = [&](double dense, double rec) {
    mx = std::fmax(mx, std::fabs(dense));
    err = std::fmax(err, std::fabs(dense - rec));
  };
  for (int fp = 0; fp < nfp; fp++) {
    int dir, side, base, st;
    host_face_line(nd, P, fp, dir, side, base, st);
    std::vector<double> row(nu, 0.0), col(nu, 0.0);
    for (int m = 0; m < n; m++) {
      row[base + m * st] = side ? o.lR[m] : o.lL[m];
      col[base + m * st] = side ? o.gR[m] : -o.gL[m];
    }
    for (int p = 0; p < nu; p++) {
      chk(opp_0[fp + (size_t)nfp * p], row[p]);
      chk(opp_3[p + (size_t)nu * fp], col[p]);
    }
  }
  for (int d = 0; d < nd; d++) {
    const int st = d == 0 ? 1 : (d == 1 ? n : n * n);
    for (int p = 0; p < nu; p++) {
      const int pos = (p / st) % n, lb = p - pos * st;
      for (int q = 0; q < nu; q++) {
        double rec = 0.0;
        const int qpos = (q / st) % n;
        if (q - qpos * st == lb) rec = o.D[pos][qpos];
        chk(opp_4[p + (size_t)nu * (q + (size_t)nu * d)], rec);
      }
    }
  }
  if (!(err <= 1e-13 * mx)) {
    std::fprintf(stderr, "hifr: operators are not tensor-product (max deviation %.3e of %.3e)\n", err, mx);
    return HIFR_ERR_NOT_TENSOR;
  }
  c->ops = o;
  c->have_ops = true;
  return HIFR_OK;
}

hifr_status hifr_set_metrics(hifr_ctx* c, const double* detjac, const double* JGinv, const double* tdA,
                             const double* norm) {
  if (!c || !detjac || !JGinv || !tdA || !norm) return HIFR_ERR_ARG;
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nd = c->nd;
  HIP_TRY(hipMemcpy(c->detj, detjac, sizeof(double) * nu * N, hipMemcpyHostToDevice));
  // JGinv(l,m,j,i) (first index fastest) -> component-major jg[(l*nd+m)][j + nu*i]
  std::vector<double> jg(nd * nd * nu * N);
  for (size_t pe = 0; pe < nu * N; pe++)
    for (size_t l = 0; l < nd; l++)
      for (size_t m = 0; m < nd; m++) jg[(l * nd + m) * nu * N + pe] = JGinv[l + nd * m + nd * nd * pe];
  HIP_TRY(hipMemcpy(c->jg, jg.data(), sizeof(double) * jg.size(), hipMemcpyHostToDevice));
  c->h_tdA.assign(tdA, tdA + nfp * N);
  c->h_norm.assign(norm, norm + nfp * N * nd);
  c->have_metrics = true;
  c->have_faces = false;
  return HIFR_OK;
}

hifr_status hifr_set_interfaces(hifr_ctx* c, int n_pts, const int32_t* left, const int32_t* right) {
  if (!c || n_pts < 0 || (n_pts > 0 && (!left || !right))) return HIFR_ERR_ARG;
  if (!c->have_metrics) return HIFR_ERR_STATE;
  const size_t N = c->N, nfp = c->nfp, nd = c->nd, tot = nfp * N;
  std::vector<int> nbr(tot, -1);
  std::vector<double> geo((nd + 1) * tot, 0.0);
  std::vector<char> seen(tot, 0);
  for (int q = 0; q < n_pts; q++) {
    const long L = left[q], R = right[q];
    if (L < 0 || R < 0 || (size_t)L >= tot || (size_t)R >= tot || seen[L] || seen[R] || L == R) return HIFR_ERR_ARG;
    seen[L] = seen[R] = 1;
    double nl[3] = {0, 0, 0};
    for (size_t m = 0; m < nd; m++) nl[m] = c->h_norm[L + tot * m];
    // LDG switch direction of the left normal (src/inters.cpp:650-660)
    const double sw = nd == 3 ? (nl[0] + nl[1] + std::sqrt(2.) * nl[2]) : (nl[0] + nl[1]);
    const bool neg_l = sw < 0.;
    nbr[L] = (int)R | (neg_l ? (int)0x80000000u : 0);
    nbr[R] = (int)L | (!neg_l ? (int)0x80000000u : 0);
    for (size_t m = 0; m < nd; m++) {
      geo[m * tot + L] = nl[m];
      geo[m * tot + R] = -nl[m];
    }
    geo[nd * tot + L] = c->h_tdA[L];
    geo[nd * tot + R] = c->h_tdA[R];
  }
  for (size_t i = 0; i < tot; i++)
    if (!seen[i]) {
      std::fprintf(stderr, "hifr: flux point %zu of element %zu has no interface (boundaries not built)\n", i % nfp,
                   i / nfp);
      return HIFR_ERR_UNSUPPORTED;
    }
  HIP_TRY(hipMemcpy(c->nbr, nbr.data(), sizeof(int) * tot, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->fgeo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice));
  c->have_faces = true;
  return HIFR_OK;
}

hifr_status hifr_upload_state(hifr_ctx* c, const double* u) {
  if (!c || !u) return HIFR_ERR_ARG;
  if (!c->have_ops) return HIFR_ERR_STATE;
  const size_t n = (size_t)c->nu * c->N * c->nf;
  HIP_TRY(hipMemcpyAsync(c->u, u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->res, 0, sizeof(double) * n, c->stream));
  c->cur = 0;
  c->have_state = true;
  hifr_status s = trace_u(c);
  if (s != HIFR_OK) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return HIFR_OK;
}

hifr_status hifr_download(hifr_ctx* c, int what, double* host) {
  if (!c || !host) return HIFR_ERR_ARG;
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nf = c->nf, nd = c->nd;
  const double* src = nullptr;
  size_t n = 0;
  switch (what) {
    case HIFR_DISU_UPTS: src = c->u; n = nu * N * nf; break;
    case HIFR_RK_REGISTER: src = c->res; n = nu * N * nf; break;
    case HIFR_DIV_TCONF_UPTS: src = c->div; n = nu * N * nf; break;
    case HIFR_DISU_FPTS: src = c->uf[c->cur]; n = nfp * N * nf; break;
    case HIFR_GRAD_DISU_FPTS:
      if (!c->cfg.viscous) return HIFR_ERR_ARG;
      src = c->gf; n = nfp * N * nf * nd; break;
    default: return HIFR_ERR_ARG;
  }
  HIP_TRY(hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return HIFR_OK;
}

hifr_status hifr_residual(hifr_ctx* c) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (c->cfg.viscous) {
    s = trace_grad(c, true);
    if (s != HIFR_OK) return s;
  }
  StageArgs a = base_args(c);
  a.uf_in = c->uf[c->cur];
  a.fv_in = c->fv;
  a.div = c->div;
  return timed_launch(c, hifr::MODE_RESID, 2, a);
}

static hifr_status advance_impl(hifr_ctx* c, int stage, int adv_type, double dt, bool write_div) {
  static const double rk4a[5] = {0.0, -0.417890474499852, -1.192151694642677, -1.697784692471528,
                                  -1.514183444257156};
  static const double rk4b[5] = {0.149659021999229, 0.379210312999627, 0.822955029386982, 0.699450455949122,
                                 0.153057247968152};
  if (adv_type != 0 && adv_type != 3) return HIFR_ERR_UNSUPPORTED;
  if (adv_type == 3 && (stage < 0 || stage > 4)) return HIFR_ERR_ARG;
  if (c->cfg.viscous) {
    hifr_status s = trace_grad(c, false);
    if (s != HIFR_OK) return s;
  }
  StageArgs a = base_args(c);
  a.uf_in = c->uf[c->cur];
  a.uf_out = c->uf[1 - c->cur];
  a.fv_in = c->fv;
  a.div = write_div ? c->div : nullptr;
  a.adv_type = adv_type;
  a.dt = dt;
  a.rk_a = adv_type == 3 ? rk4a[stage] : 0.0;
  a.rk_b = adv_type == 3 ? rk4b[stage] : 0.0;
  hifr_status s = timed_launch(c, hifr::MODE_ADVANCE, 2, a);
  if (s != HIFR_OK) return s;
  c->cur = 1 - c->cur;
  return HIFR_OK;
}

hifr_status hifr_advance(hifr_ctx* c, int stage, int adv_type, double dt) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  return advance_impl(c, stage, adv_type, dt, true);
}

hifr_status hifr_run(hifr_ctx* c, int n_steps, int adv_type, double dt) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (n_steps < 0) return HIFR_ERR_ARG;
  const int nst = adv_type == 3 ? 5 : 1;
  for (int it = 0; it < n_steps; it++)
    for (int st = 0; st < nst; st++) {
      const bool last = (it == n_steps - 1) && (st == nst - 1);
      s = advance_impl(c, st, adv_type, dt, last);
      if (s != HIFR_OK) return s;
    }
  return HIFR_OK;
}

hifr_status hifr_synchronize(hifr_ctx* c) {
  if (!c) return HIFR_ERR_ARG;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return HIFR_OK;
}

hifr_status hifr_run_timed(hifr_ctx* c, int n_steps, int adv_type, double dt, double* total_ms, double* kernel_ms,
                           int* kernel_launches) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  const int nst = adv_type == 3 ? 5 : 1;
  const size_t need = (size_t)n_steps * nst * 2 * 2 + 2;
  if (need > (size_t)EV_POOL) return HIFR_ERR_ARG;
  while (c->ev.size() < need) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c->ev.push_back(e);
  }
  c->ev_kind.assign(c->ev.size() / 2 + 1, -1);
  hipEvent_t t0, t1;
  HIP_TRY(hipEventCreate(&t0));
  HIP_TRY(hipEventCreate(&t1));
  HIP_TRY(hipEventRecord(t0, c->stream));
  c->timing = true;
  c->ev_used = 0;
  s = hifr_run(c, n_steps, adv_type, dt);
  c->timing = false;
  HIP_TRY(hipEventRecord(t1, c->stream));
  HIP_TRY(hipEventSynchronize(t1));
  if (s != HIFR_OK) return s;
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, t0, t1));
  if (total_ms) *total_ms = ms;
  double km[3] = {0, 0, 0};
  int kl[3] = {0, 0, 0};
  for (size_t i = 0; i + 1 < c->ev_used; i += 2) {
    float d = 0.f;
    HIP_TRY(hipEventElapsedTime(&d, c->ev[i], c->ev[i + 1]));
    const int k = c->ev_kind[i / 2];
    if (k >= 0 && k < 3) {
      km[k] += d;
      kl[k] += 1;
    }
  }
  for (int k = 0; k < 3; k++) {
    if (kernel_ms) kernel_ms[k] = km[k];
    if (kernel_launches) kernel_launches[k] = kl[k];
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  return HIFR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// residual norms (CalcNormResidual): two-pass deterministic reduction

namespace {

template <int NF>
__global__ void __launch_bounds__(256) norm_partial_kernel(const double* __restrict__ div, const double* __restrict__ detj,
                                                           long npts, int norm_type, double const_src, double* out) {
  __shared__ double sm[NF][256];
  double acc[NF];
  for (int k = 0; k < NF; k++) acc[k] = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < npts; i += (long)gridDim.x * 256L) {
    const double inv = 1.0 / detj[i];
    for (int k = 0; k < NF; k++) {
      const double r = div[i + npts * k] * inv - const_src;
      if (norm_type == 0) acc[k] = fmax(acc[k], fabs(r));
      else if (norm_type == 1) acc[k] += fabs(r);
      else acc[k] += r * r;
    }
  }
  for (int k = 0; k < NF; k++) sm[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < NF; k++) {
        const double o = sm[k][threadIdx.x + w];
        sm[k][threadIdx.x] = norm_type == 0 ? fmax(sm[k][threadIdx.x], o) : sm[k][threadIdx.x] + o;
      }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < NF; k++) out[blockIdx.x * NF + k] = sm[k][0];
}

}  // namespace

extern "C" hifr_status hifr_norm_residual(hifr_ctx* c, int norm_type, double* out) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (!out || norm_type < 0 || norm_type > 2) return HIFR_ERR_ARG;
  const long npts = (long)c->nu * c->N;
  const int nb = 512;
  if (c->nf == 5)
    hipLaunchKernelGGL(norm_partial_kernel<5>, dim3(nb), dim3(256), 0, c->stream, c->div, c->detj, npts, norm_type,
                       c->cfg.const_src, c->red);
  else
    hipLaunchKernelGGL(norm_partial_kernel<4>, dim3(nb), dim3(256), 0, c->stream, c->div, c->detj, npts, norm_type,
                       c->cfg.const_src, c->red);
  HIP_TRY(hipGetLastError());
  std::vector<double> part((size_t)nb * c->nf);
  HIP_TRY(hipMemcpyAsync(part.data(), c->red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  bool nan = false;
  for (int k = 0; k < c->nf; k++) {
    double v = 0.0;
    for (int b = 0; b < nb; b++) {
      const double x = part[(size_t)b * c->nf + k];
      v = norm_type == 0 ? std::fmax(v, x) : v + x;
      if (std::isnan(x)) nan = true;
    }
    if (norm_type == 1) v = v / (double)npts;
    else if (norm_type == 2) v = std::sqrt(v) / (double)npts;
    out[k] = v;
  }
  return nan ? HIFR_ERR_NAN : HIFR_OK;
}

// ---------------------------------------------------------------------------
// streaming copy used to calibrate byte counters

namespace {

__global__ void __launch_bounds__(256) stream_copy_kernel(double* __restrict__ dst, const double* __restrict__ src, long n) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256L) dst[i] = src[i];
}

}  // namespace

extern "C" hifr_status hifr_calibrate_copy(long n, int reps, double* ms) {
  if (n <= 0 || reps <= 0 || !ms) return HIFR_ERR_ARG;
  double *a = nullptr, *b = nullptr;
  HIP_TRY(hipMalloc(&a, sizeof(double) * n));
  if (hipMalloc(&b, sizeof(double) * n) != hipSuccess) {
    (void)hipFree(a);
    return HIFR_ERR_DEVICE;
  }
  hifr_status st = HIFR_OK;
  hipEvent_t e0, e1;
  if (hipMemset(a, 0, sizeof(double) * n) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    st = HIFR_ERR_DEVICE;
  } else {
    const int grid = 256 * 8 * 4;
    hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, 0, b, a, n);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < reps; r++) hipLaunchKernelGGL(stream_copy_kernel, dim3(grid), dim3(256), 0, 0, b, a, n);
    (void)hipEventRecord(e1, 0);
    if (hipEventSynchronize(e1) != hipSuccess) st = HIFR_ERR_DEVICE;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  (void)hipFree(a);
  (void)hipFree(b);
  return st;
}
