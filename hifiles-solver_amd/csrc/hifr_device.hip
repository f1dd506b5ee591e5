// C ABI of the gfx950 FR residual / RK update (include/hifr.h): context, device
// buffers, operator factorisation, interface tables and kernel dispatch.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/hifr.h"
#include "../../include/hifr_sgs.h"
#include "fr_launch.hpp"
#include "fr_launch_check.hpp"
#include "hifr_ctx_view.hpp"
#include "sgs_mlp_state.hpp"

using hifr::Ops1D;
using hifr::StageArgsT;

struct hifr_ctx {
  hifr_config cfg;
  int nd, nf, n1, nu, nfp, nfpf, nfaces, N;
  hipStream_t stream;
  Ops1D ops;
  hifr::Ops1DT<float> opsf;
  bool have_ops, have_metrics, have_faces, have_state;
  bool affine;  // metrics kept per element / per face
  bool mixed;   // HIFR_MIXED: FP32 traces / JGinv / arithmetic on deviations from ref
  size_t tsz;   // bytes of a trace / JGinv value (8, or 4 when mixed)
  hifr::RefState<double> ref;
  // device buffers
  double *u, *res, *div, *detj, *fgeo;
  void *uf[2], *fv, *gf, *jg;  // element type double, or float when mixed
  void *tI, *tG, *tD;          // tets: combined dense operators as MFMA A-operand tiles
                               // (double, or float when mixed)
  // learned SGS closure (sgs_model HIFR_SGS_LEARNED): the context's own copy of the
  // attached network's weights (mlp_buf, zero while none is attached), which the
  // LES kernels read through mlpw
  void* mlp_buf;
  hifr::mlp::Weights mlpw;
  hifr::mlp::Norm mlpn;
  int mlp_on;
  int* nbr;
  double* red;  // reduction scratch
  // integral diagnostics (hifr_set_cubature): corrected gradient at the solution
  // points, volume cubature operator / weights / detJ, per-element partial sums
  double *gu, *cub_opp, *cub_w, *cub_detj, *cub_part;
  int ncub;
  // learned closure: eddy viscosity per solution point (Nu, N), element type of
  // the traces, written by the gradient-trace kernel, read by the flux kernels
  void* nut;
  int cur;      // which uf buffer holds the traces of the current state
  // host copies needed to build the interface tables
  std::vector<double> h_tdA, h_norm;
  std::vector<int> h_bdy_fpt, h_bdy_type;  // physical boundary points (hifr_set_boundaries)
  // timing
  bool timing;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;
  size_t ev_used;
  // hifr_run: one time step captured as a HIP graph per trace-buffer parity, for
  // one (adv_type, dt); dropped by every call that changes what the kernels read
  hipStream_t gstream;
  hipGraphExec_t gexec[2];
  int g_adv_type;
  double g_dt;
};

namespace {

const int EV_POOL = 8192;

void drop_graphs(hifr_ctx* c) {
  for (hipGraphExec_t& g : c->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

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

// The context's copy of a network's weights, one buffer in the layouts of
// hifr::mlp::Weights: BF16 W1 [64][16], W2 and W3 [64][64], FP32 w4 [64], then the
// FP32 parameter vector (the kernels read the biases from it).
constexpr size_t kMlpW1 = 0, kMlpW2 = kMlpW1 + 2 * hifr::mlp::H * hifr::mlp::NPAD,
                 kMlpW3 = kMlpW2 + 2 * hifr::mlp::H * hifr::mlp::H, kMlpW4 = kMlpW3 + 2 * hifr::mlp::H * hifr::mlp::H,
                 kMlpP = kMlpW4 + 4 * hifr::mlp::H, kMlpBufBytes = kMlpP + 4 * hifr::mlp::NPARAM;

hifr::mlp::Weights mlp_weights_in(void* buf) {
  char* b = (char*)buf;
  return hifr::mlp::Weights{(const unsigned short*)(b + kMlpW1), (const unsigned short*)(b + kMlpW2),
                            (const unsigned short*)(b + kMlpW3), (const float*)(b + kMlpW4),
                            (const float*)(b + kMlpP)};
}

bool order_supported(int ele_type, int P) {
  if (ele_type == HIFR_HEX) return P >= 1 && P <= 5;
  if (ele_type == HIFR_QUAD) return P >= 1 && P <= 6;
  if (ele_type == HIFR_TET) return P >= 1 && P <= 3;
  return false;
}

int vis_kind(const hifr_ctx* c) { return c->cfg.viscous == 0 ? 0 : (c->cfg.fix_vis == 1.0 ? 1 : 2); }

// 0 no LES, 1 Smagorinsky, 2 WALE
int les_kind(const hifr_ctx* c) {
  if (!c->cfg.les) return 0;
  return c->cfg.sgs_model == HIFR_SGS_LEARNED ? 3 : c->cfg.sgs_model + 1;
}

template <class R>
StageArgsT<R> base_args(hifr_ctx* c) {
  StageArgsT<R> a;
  std::memset(&a, 0, sizeof(a));
  a.N = c->N;
  a.u = c->u;
  a.res = c->res;
  a.detj = c->detj;
  a.jg = (const R*)c->jg;
  a.nbr = c->nbr;
  a.fgeo = c->fgeo;
  if constexpr (std::is_same_v<R, float>) a.ops = c->opsf;
  else a.ops = c->ops;
  a.ph.gamma = c->cfg.gamma;
  a.ph.prandtl = c->cfg.prandtl;
  a.ph.mu_inf = c->cfg.mu_inf;
  a.ph.rt_inf = c->cfg.rt_inf;
  a.ph.c_sth = c->cfg.c_sth;
  a.ph.fix_vis = c->cfg.fix_vis;
  a.ph.tau = c->cfg.tau;
  a.ph.pen_fact = c->cfg.pen_fact;
  a.ph.const_src = c->cfg.const_src;
  a.ph.R_ref = c->cfg.R_ref;
  a.ph.T_wall = c->cfg.T_wall;
  a.ref = c->ref;
  a.filter_ratio = c->cfg.filter_ratio;
  a.tI = (const R*)c->tI;
  a.tG = (const R*)c->tG;
  a.tD = (const R*)c->tD;
  a.mlpw = c->mlpw;
  for (int f = 0; f < hifr::mlp::NIN; f++) {
    a.mlp_mean[f] = c->mlpn.mean[f];
    a.mlp_scale[f] = c->mlpn.scale[f];
    a.mlp_lo[f] = c->mlpn.lo[f];
    a.mlp_hi[f] = c->mlpn.hi[f];
  }
  a.mlp_out_scale = c->mlpn.out_scale;
  a.mlp_on = c->mlp_on;
  a.nut = (R*)c->nut;
  return a;
}

// Buffers of one stage launch, filled by the callers below.
struct StageIO {
  int uf_in = -1, uf_out = -1;  // trace buffer indices
  bool fv_in = false, fv_out = false, gf_out = false, gu_out = false, div = false;
  int adv_type = 0;
  double rk_a = 0, rk_b = 0, dt = 0;
};

template <class R>
hipError_t launch_args(hifr_ctx* c, int mode, const StageArgsT<R>& a) {
  if constexpr (std::is_same_v<R, double>) {
    // the Roe flux changes only the kernels that evaluate the inviscid interface flux
    if (c->cfg.riemann_solve_type == 2 && (mode == hifr::MODE_RESID || mode == hifr::MODE_ADVANCE)) {
      if (c->cfg.ele_type == HIFR_TET)
        return hifr::launch_tet_roe_fp64(c->cfg.order, mode, vis_kind(c), c->affine, a, c->stream);
      return hifr::launch_roe_fp64(c->nd, c->cfg.order, mode, vis_kind(c), c->affine, a, c->stream);
    }
  }
  if (c->cfg.ele_type == HIFR_TET) {
    if constexpr (std::is_same_v<R, float>)
      return hifr::launch_tet_mixed(c->cfg.order, mode, vis_kind(c), c->affine, les_kind(c), a, c->stream);
    else
      return hifr::launch_tet_fp64(c->cfg.order, mode, vis_kind(c), c->affine, les_kind(c), a, c->stream);
  }
  if constexpr (std::is_same_v<R, float>)
    return hifr::launch_mixed(c->nd, c->cfg.order, mode, vis_kind(c), c->affine, les_kind(c), a, c->stream);
  else
    return hifr::launch_fp64(c->nd, c->cfg.order, mode, vis_kind(c), c->affine, les_kind(c), a, c->stream);
}

// Fills the arguments of one stage launch, checks that every array the variant
// uses is set (hifr::missing_input), and launches it.
template <class R>
hifr_status launch_io(hifr_ctx* c, int mode, const StageIO& io) {
  StageArgsT<R> a = base_args<R>(c);
  a.uf_in = io.uf_in >= 0 ? (const R*)c->uf[io.uf_in] : nullptr;
  a.uf_out = io.uf_out >= 0 ? (R*)c->uf[io.uf_out] : nullptr;
  a.fv_in = io.fv_in ? (const R*)c->fv : nullptr;
  a.fv_out = io.fv_out ? (R*)c->fv : nullptr;
  a.gf_out = io.gf_out ? (R*)c->gf : nullptr;
  a.gu_out = io.gu_out ? c->gu : nullptr;
  a.div = io.div ? c->div : nullptr;
  a.adv_type = io.adv_type;
  a.rk_a = io.rk_a;
  a.rk_b = io.rk_b;
  a.dt = io.dt;
  const hifr::LaunchPtrs ptrs{a.u,     a.res,      a.div,      a.uf_in,    a.uf_out,    a.fv_in, a.fv_out,
                              a.detj,  a.jg,       a.nbr,      a.fgeo,     a.tI,        a.tG,    a.tD,
                              a.mlpw.w1, a.mlpw.w2, a.mlpw.w3, a.mlpw.w4, a.mlpw.params, a.nut};
  const hifr::LaunchKind kind{mode, c->cfg.viscous != 0, les_kind(c), c->cfg.ele_type == HIFR_TET};
  if (const char* miss = hifr::missing_input(ptrs, kind)) {
    std::fprintf(stderr, "hifr: stage kernel (mode %d) not launched: %s is not set\n", mode, miss);
    return HIFR_ERR_STATE;
  }
  return dev_err(launch_args<R>(c, mode, a), "stage kernel launch");
}

hifr_status launch(hifr_ctx* c, int mode, const StageIO& io) {
  return c->mixed ? launch_io<float>(c, mode, io) : launch_io<double>(c, mode, io);
}

hifr_status timed_launch(hifr_ctx* c, int mode, int kind, const StageIO& io) {
  if (c->timing) {
    if (c->ev_used + 2 > c->ev.size()) return HIFR_ERR_ARG;
    HIP_TRY(hipEventRecord(c->ev[c->ev_used], c->stream));
    c->ev_kind[c->ev_used / 2] = kind;
  }
  const hifr_status ls = launch(c, mode, io);
  if (ls != HIFR_OK) return ls;
  if (c->timing) {
    HIP_TRY(hipEventRecord(c->ev[c->ev_used + 1], c->stream));
    c->ev_used += 2;
  }
  return HIFR_OK;
}

hifr_status trace_u(hifr_ctx* c) {
  StageIO io;
  io.uf_out = c->cur;
  return timed_launch(c, hifr::MODE_TRACE_U, 0, io);
}

// write_grad: also store the corrected gradient at the flux points (for
// hifr_download; the residual itself only needs the viscous normal fluxes)
hifr_status trace_grad(hifr_ctx* c, bool write_grad) {
  StageIO io;
  io.uf_in = c->cur;
  io.fv_out = true;
  io.gf_out = write_grad;
  return timed_launch(c, hifr::MODE_TRACE_GRAD, 1, io);
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
  if (cfg->ele_type != HIFR_HEX && cfg->ele_type != HIFR_QUAD && cfg->ele_type != HIFR_TET)
    return HIFR_ERR_UNSUPPORTED;
  if (!order_supported(cfg->ele_type, cfg->order)) return HIFR_ERR_UNSUPPORTED;
  if (cfg->n_eles <= 0) return HIFR_ERR_ARG;
  // inviscid flux: Rusanov, or Roe in FP64 without LES; viscous flux: LDG
  if (cfg->riemann_solve_type != 0 &&
      !(cfg->riemann_solve_type == 2 && cfg->precision == HIFR_FP64 && !cfg->les))
    return HIFR_ERR_UNSUPPORTED;
  if (cfg->vis_riemann_solve_type != 0) return HIFR_ERR_UNSUPPORTED;
  if (cfg->precision != HIFR_FP64 && cfg->precision != HIFR_MIXED) return HIFR_ERR_UNSUPPORTED;
  if (cfg->les && (!cfg->viscous || cfg->fix_vis != 1.0 ||
                   !(cfg->sgs_model == HIFR_SGS_SMAGORINSKY || cfg->sgs_model == HIFR_SGS_WALE ||
                     cfg->sgs_model == HIFR_SGS_LEARNED)))
    return HIFR_ERR_UNSUPPORTED;
  if (cfg->les && cfg->sgs_model == HIFR_SGS_LEARNED &&
      (!hifr::mlp::kLearnedEnabled || !(cfg->ele_type == HIFR_HEX && (cfg->order == 2 || cfg->order == 3))))
    return HIFR_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HIFR_ERR_DEVICE;
  hifr_ctx* c = new hifr_ctx();
  c->cfg = *cfg;
  c->nd = cfg->ele_type == HIFR_QUAD ? 2 : 3;
  c->nf = c->nd + 2;
  c->n1 = cfg->order + 1;
  if (cfg->ele_type == HIFR_TET) {
    c->nu = c->n1 * (c->n1 + 1) * (c->n1 + 2) / 6;
    c->nfpf = c->n1 * (c->n1 + 1) / 2;
    c->nfaces = 4;
  } else {
    c->nu = c->nd == 3 ? c->n1 * c->n1 * c->n1 : c->n1 * c->n1;
    c->nfpf = c->nd == 3 ? c->n1 * c->n1 : c->n1;
    c->nfaces = 2 * c->nd;
  }
  c->nfp = c->nfaces * c->nfpf;
  c->N = cfg->n_eles;
  c->stream = nullptr;
  c->cur = 0;
  c->timing = false;
  c->ev_used = 0;
  c->mixed = cfg->precision == HIFR_MIXED;
  c->tsz = c->mixed ? sizeof(float) : sizeof(double);
  c->ref = hifr::RefState<double>{0.0, 0.0, 0.0, 0.0};
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nf = c->nf, nd = c->nd, tsz = c->tsz;
  struct { void** p; size_t bytes; } allocs[] = {
      {(void**)&c->u, sizeof(double) * nu * N * nf},
      {(void**)&c->res, sizeof(double) * nu * N * nf},
      {(void**)&c->div, sizeof(double) * nu * N * nf},
      {&c->uf[0], tsz * nfp * N * nf},
      {&c->uf[1], tsz * nfp * N * nf},
      {&c->fv, cfg->viscous ? tsz * nfp * N * nf : 8},
      {&c->gf, cfg->viscous ? tsz * nfp * N * nf * nd : 8},
      {(void**)&c->nbr, sizeof(int) * nfp * N},
      {(void**)&c->red, sizeof(double) * 4096},
  };
  for (auto& al : allocs) *al.p = nullptr;
  c->detj = c->fgeo = nullptr;  // sized by hifr_set_metrics / hifr_set_interfaces
  c->jg = nullptr;
  c->gu = c->cub_opp = c->cub_w = c->cub_detj = c->cub_part = nullptr;  // hifr_set_cubature
  c->ncub = 0;
  c->nut = nullptr;
  c->tI = c->tG = c->tD = nullptr;
  c->mlp_buf = nullptr;
  std::memset(&c->mlpw, 0, sizeof(c->mlpw));
  std::memset(&c->mlpn, 0, sizeof(c->mlpn));
  c->mlp_on = 0;
  c->affine = false;
  for (auto& al : allocs) {
    if (hipMalloc(al.p, al.bytes) != hipSuccess) {
      hifr_destroy(c);
      return HIFR_ERR_DEVICE;
    }
  }
  if (cfg->les && cfg->sgs_model == HIFR_SGS_LEARNED) {
    // the weights the LES kernels read: zero until a network is attached
    if (hipMalloc(&c->mlp_buf, kMlpBufBytes) != hipSuccess || hipMemset(c->mlp_buf, 0, kMlpBufBytes) != hipSuccess) {
      hifr_destroy(c);
      return HIFR_ERR_DEVICE;
    }
    c->mlpw = mlp_weights_in(c->mlp_buf);
    if (hipMalloc(&c->nut, tsz * nu * N) != hipSuccess) {
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
  void* ps[] = {c->u,    c->res, c->div, c->uf[0], c->uf[1], c->fv,      c->gf,    c->detj,   c->jg,
                c->fgeo, c->nbr, c->red, c->tI,    c->tG,    c->tD,    c->mlp_buf, c->gu,
                c->cub_opp, c->cub_w, c->cub_detj, c->cub_part, c->nut};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  drop_graphs(c);
  if (c->gstream) (void)hipStreamDestroy(c->gstream);
  delete c;
  return HIFR_OK;
}

hifr_status hifr_set_stream(hifr_ctx* c, void* s) {
  if (!c) return HIFR_ERR_ARG;
  drop_graphs(c);
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

}  // extern "C"

namespace {

// Tets: the dense operators in the kernel layouts I [Nfp][Nu], D [d][Nu][Nu],
// L [Nu][Nfp], in the context's arithmetic type.
// Operator A (M x K, a(row, col); K a multiple of 4) in the 16x16x4 MFMA A-operand
// order: tile (mt, ks) holds, for lane l, A[16 mt + l % 16][4 ks + l / 16] (zero for
// rows >= M), 64 values per tile, tiles mt-major.
template <class R, class F>
void mfma_tiles(int M, int K, F a, std::vector<R>& out) {
  const int MT = (M + 15) / 16, KS = K / 4;
  out.assign((size_t)MT * KS * 64, R(0));
  for (int mt = 0; mt < MT; mt++)
    for (int ks = 0; ks < KS; ks++)
      for (int l = 0; l < 64; l++) {
        const int r = 16 * mt + l % 16, col = 4 * ks + l / 16;
        if (r < M) out[((size_t)mt * KS + ks) * 64 + l] = R(a(r, col));
      }
}

// The four combined tet operators of the MFMA flux pass (fr_tet_mfma.hpp), from the
// reference's opp_0 (I), opp_4 (D) and opp_3 (L) and the reference tet's face
// normals (face 0: (1,1,1)/sqrt(3); faces 1-3: -e_r, -e_s, -e_t).
template <int P, class R>
bool tet_tiles(const double* opp_0, const double* opp_3, const double* opp_4, std::vector<R>& tI,
               std::vector<R>& tG, std::vector<R>& tD) {
  using M = hifr::TetMfma<P>;
  constexpr int NU = M::NU, NFP = M::NFP, KU = M::KU, KF = M::KF, NFPF = NFP / 4;
  auto Iv = [&](int f, int m) { return opp_0[f + NFP * m]; };
  auto Dv = [&](int c, int p, int m) { return opp_4[p + NU * (m + NU * c)]; };
  auto Lv = [&](int p, int f) { return opp_3[p + NU * f]; };
  auto tn = [](int face, int c) { return face == 0 ? 0.57735026918962576 : (face - 1 == c ? -1.0 : 0.0); };
  mfma_tiles<R>(NFP, KU, [&](int f, int m) { return m < NU ? Iv(f, m) : 0.0; }, tI);
  mfma_tiles<R>(3 * NU, KU + NFP, [&](int r, int col) {
    const int c = r / NU, p = r % NU;
    if (col < KU) return col < NU ? Dv(c, p, col) : 0.0;
    const int f = col - KU;
    return Lv(p, f) * tn(f / NFPF, c);
  }, tG);
  // divergence with the normal flux folded in: D' = D - L N, N = [tnorm_c o I]
  // (the flux pass applies [D' | L] to [Ft; jump] instead of [D | L] to
  // [Ft; jump - N Ft])
  std::vector<double> Dp((size_t)NU * 3 * NU);
  for (int p = 0; p < NU; p++)
    for (int col = 0; col < 3 * NU; col++) {
      double ln = 0.0;
      for (int f = 0; f < NFP; f++) ln += Lv(p, f) * (tn(f / NFPF, col / NU) * Iv(f, col % NU));
      Dp[(size_t)p * 3 * NU + col] = Dv(col / NU, p, col % NU) - ln;
    }
  mfma_tiles<R>(NU, KF + NFP, [&](int p, int col) {
    if (col < 3 * NU) return Dp[(size_t)p * 3 * NU + col];
    return col < KF ? 0.0 : Lv(p, col - KF);
  }, tD);
  // the kernels skip the tiles that the face normals zero (hifr::tet_tile_nz)
  auto skipped_are_zero = [](const std::vector<R>& t, int KS, auto nz) {
    for (size_t i = 0; i < t.size(); i++) {
      const int tile = (int)(i / 64);
      if (!nz(tile / KS, tile % KS) && t[i] != R(0)) return false;
    }
    return true;
  };
  return skipped_are_zero(tG, M::KS_G, [](int mt, int ks) { return hifr::tet_tile_nz<P, hifr::TET_OP_G>(mt, ks); });
}

template <class R>
hifr_status upload_dense(hifr_ctx* c, const double* opp_0, const double* opp_3, const double* opp_4) {
  const size_t nu = c->nu, nfp = c->nfp, nd = c->nd;
  for (size_t i = 0; i < nfp * nu; i++)
    if (!std::isfinite(opp_0[i]) || !std::isfinite(opp_3[i])) return HIFR_ERR_ARG;
  for (size_t i = 0; i < nd * nu * nu; i++)
    if (!std::isfinite(opp_4[i])) return HIFR_ERR_ARG;
  std::vector<R> tI, tG, tD;
  bool ok = false;
  switch (c->cfg.order) {
    case 1: ok = tet_tiles<1>(opp_0, opp_3, opp_4, tI, tG, tD); break;
    case 2: ok = tet_tiles<2>(opp_0, opp_3, opp_4, tI, tG, tD); break;
    case 3: ok = tet_tiles<3>(opp_0, opp_3, opp_4, tI, tG, tD); break;
    default: return HIFR_ERR_UNSUPPORTED;  // hifr_create admits tets of order 1-3 only
  }
  if (!ok) return HIFR_ERR_ARG;
  struct { void** p; const std::vector<R>* v; } bufs[] = {{&c->tI, &tI}, {&c->tG, &tG},
                                                          {&c->tD, &tD}};
  for (auto& b : bufs) {
    if (*b.p) (void)hipFree(*b.p);
    *b.p = nullptr;
    HIP_TRY(hipMalloc(b.p, sizeof(R) * b.v->size()));
    HIP_TRY(hipMemcpy(*b.p, b.v->data(), sizeof(R) * b.v->size(), hipMemcpyHostToDevice));
  }
  return HIFR_OK;
}

}  // namespace

extern "C" {

hifr_status hifr_set_operators(hifr_ctx* c, const double* opp_0, const double* opp_3, const double* opp_4) {
  if (!c || !opp_0 || !opp_3 || !opp_4) return HIFR_ERR_ARG;
  drop_graphs(c);
  {
    const size_t nb = (size_t)c->nfp * c->nu, nd4 = (size_t)c->nu * c->nu * c->nd;
    for (size_t i = 0; i < nb; i++)
      if (!std::isfinite(opp_0[i]) || !std::isfinite(opp_3[i])) return HIFR_ERR_ARG;
    for (size_t i = 0; i < nd4; i++)
      if (!std::isfinite(opp_4[i])) return HIFR_ERR_ARG;
  }
  if (c->cfg.ele_type == HIFR_TET) {
    hifr_status st = c->mixed ? upload_dense<float>(c, opp_0, opp_3, opp_4) : upload_dense<double>(c, opp_0, opp_3, opp_4);
    if (st != HIFR_OK) return st;
    c->have_ops = true;
    return HIFR_OK;
  }
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
  hifr::Ops1DT<float>& f = c->opsf;
  for (int a = 0; a < hifr::MAXN1; a++) {
    f.lL[a] = float(o.lL[a]);
    f.lR[a] = float(o.lR[a]);
    f.gL[a] = float(o.gL[a]);
    f.gR[a] = float(o.gR[a]);
    for (int m = 0; m < hifr::MAXN1; m++) f.D[a][m] = float(o.D[a][m]);
  }
  c->have_ops = true;
  return HIFR_OK;
}

}  // extern "C"

namespace {

// True when every element's volume metrics are the same at all its solution points
// and every face's normal and tdA the same at all its flux points, to `rel`
// relative to the largest magnitude of the quantity in that element / face.
bool all_affine(const hifr_ctx* c, const double* detjac, const double* JGinv, const double* tdA,
                const double* norm, double rel) {
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nd = c->nd, nfpf = c->nfpf;
  for (size_t e = 0; e < N; e++) {
    const double* J = JGinv + nd * nd * nu * e;
    double mx = 0.0, dev = 0.0;
    for (size_t p = 0; p < nu; p++)
      for (size_t k = 0; k < nd * nd; k++) {
        mx = std::fmax(mx, std::fabs(J[k + nd * nd * p]));
        dev = std::fmax(dev, std::fabs(J[k + nd * nd * p] - J[k]));
      }
    if (!(dev <= rel * mx)) return false;
    const double* dj = detjac + nu * e;
    for (size_t p = 0; p < nu; p++)
      if (!(std::fabs(dj[p] - dj[0]) <= rel * std::fabs(dj[0]))) return false;
    for (size_t f = 0; f < (size_t)c->nfaces; f++) {
      const size_t f0 = f * nfpf + nfp * e;
      for (size_t q = 0; q < nfpf; q++) {
        const size_t fe = f0 + q;
        if (!(std::fabs(tdA[fe] - tdA[f0]) <= rel * std::fabs(tdA[f0]))) return false;
        for (size_t m = 0; m < nd; m++)
          if (!(std::fabs(norm[fe + nfp * N * m] - norm[f0 + nfp * N * m]) <= rel)) return false;
      }
    }
  }
  return true;
}

template <class V>
hifr_status realloc_dev(V** p, size_t bytes) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  HIP_TRY(hipMalloc((void**)p, bytes));
  return HIFR_OK;
}

}  // namespace

extern "C" {

hifr_status hifr_set_metrics(hifr_ctx* c, const double* detjac, const double* JGinv, const double* tdA,
                             const double* norm) {
  if (!c || !detjac || !JGinv || !tdA || !norm) return HIFR_ERR_ARG;
  drop_graphs(c);
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nd = c->nd;
  c->have_metrics = false;
  c->have_faces = false;
  c->affine = c->cfg.metric_storage == HIFR_METRICS_AUTO && all_affine(c, detjac, JGinv, tdA, norm, 1e-13);
  // per element: the values at the first solution point; per point: as given
  const size_t nv = c->affine ? N : nu * N, stride = c->affine ? nu : 1;
  std::vector<double> dj(nv), jg(nd * nd * nv);
  // JGinv(l,m,j,i) (first index fastest) -> component-major jg[(l*nd+m)][j + nu*i]
  for (size_t v = 0; v < nv; v++) {
    const size_t pe = v * stride;
    dj[v] = detjac[pe];
    for (size_t l = 0; l < nd; l++)
      for (size_t m = 0; m < nd; m++) jg[(l * nd + m) * nv + v] = JGinv[l + nd * m + nd * nd * pe];
  }
  hifr_status st;
  if ((st = realloc_dev(&c->detj, sizeof(double) * nv)) != HIFR_OK) return st;
  if ((st = realloc_dev(&c->jg, c->tsz * nd * nd * nv)) != HIFR_OK) return st;
  HIP_TRY(hipMemcpy(c->detj, dj.data(), sizeof(double) * nv, hipMemcpyHostToDevice));
  if (c->mixed) {
    std::vector<float> jf(jg.begin(), jg.end());
    HIP_TRY(hipMemcpy(c->jg, jf.data(), sizeof(float) * jf.size(), hipMemcpyHostToDevice));
  } else {
    HIP_TRY(hipMemcpy(c->jg, jg.data(), sizeof(double) * jg.size(), hipMemcpyHostToDevice));
  }
  c->h_tdA.assign(tdA, tdA + nfp * N);
  c->h_norm.assign(norm, norm + nfp * N * nd);
  c->have_metrics = true;
  return HIFR_OK;
}

hifr_status hifr_metrics_per_element(const hifr_ctx* c, int* per_element) {
  if (!c || !per_element) return HIFR_ERR_ARG;
  if (!c->have_metrics) return HIFR_ERR_STATE;
  *per_element = c->affine ? 1 : 0;
  return HIFR_OK;
}

hifr_status hifr_set_boundaries(hifr_ctx* c, int n_pts, const int32_t* fpt, const int32_t* bctype) {
  if (!c || n_pts < 0 || (n_pts > 0 && (!fpt || !bctype))) return HIFR_ERR_ARG;
  drop_graphs(c);
  const size_t tot = (size_t)c->nfp * c->N;
  for (int q = 0; q < n_pts; q++) {
    if (fpt[q] < 0 || (size_t)fpt[q] >= tot) return HIFR_ERR_ARG;
    if (bctype[q] != 11) return HIFR_ERR_UNSUPPORTED;
  }
  c->h_bdy_fpt.assign(fpt, fpt + n_pts);
  c->h_bdy_type.assign(bctype, bctype + n_pts);
  c->have_faces = false;
  return HIFR_OK;
}

hifr_status hifr_set_interfaces(hifr_ctx* c, int n_pts, const int32_t* left, const int32_t* right) {
  if (!c || n_pts < 0 || (n_pts > 0 && (!left || !right))) return HIFR_ERR_ARG;
  if (!c->have_metrics) return HIFR_ERR_STATE;
  drop_graphs(c);
  const size_t N = c->N, nfp = c->nfp, nd = c->nd, tot = nfp * N, nfpf = c->nfpf;
  // face geometry per flux point, or per face when the metrics are per element
  const size_t gtot = c->affine ? (size_t)c->nfaces * N : tot;
  auto gidx = [&](size_t fe) { return c->affine ? fe / nfpf : fe; };
  std::vector<int> nbr(tot, -1);
  std::vector<double> geo((nd + 1) * gtot, 0.0);
  std::vector<char> seen(tot, 0);
  // physical boundary points: own outward normal and tdA
  for (size_t q = 0; q < c->h_bdy_fpt.size(); q++) {
    const size_t B = (size_t)c->h_bdy_fpt[q];
    if (seen[B]) return HIFR_ERR_ARG;
    seen[B] = 1;
    nbr[B] = hifr::BDY_FLAG | c->h_bdy_type[q];
    const size_t gb = gidx(B);
    for (size_t m = 0; m < nd; m++) geo[m * gtot + gb] = c->h_norm[B + tot * m];
    geo[nd * gtot + gb] = c->h_tdA[B];
  }
  for (int q = 0; q < n_pts; q++) {
    const long L = left[q], R = right[q];
    if (L < 0 || R < 0 || (size_t)L >= tot || (size_t)R >= tot || seen[L] || seen[R] || L == R) return HIFR_ERR_ARG;
    seen[L] = seen[R] = 1;
    double nl[3] = {0, 0, 0};
    for (size_t m = 0; m < nd; m++) nl[m] = c->h_norm[L + tot * m];
    // one normal (the left one) serves both sides of the point, as in the
    // reference; per-face storage keeps the value of the face's last point
    const size_t gl = gidx(L), gr = gidx(R);
    for (size_t m = 0; m < nd; m++) {
      geo[m * gtot + gl] = nl[m];
      geo[m * gtot + gr] = -nl[m];
    }
    geo[nd * gtot + gl] = c->h_tdA[L];
    geo[nd * gtot + gr] = c->h_tdA[R];
  }
  for (size_t i = 0; i < tot; i++)
    if (!seen[i]) {
      std::fprintf(stderr, "hifr: flux point %zu of element %zu has no interface or boundary\n", i % nfp,
                   i / nfp);
      return HIFR_ERR_UNSUPPORTED;
    }
  // LDG switch direction of the left normal (src/inters.cpp:650-660), from the
  // normal the kernels read (per point: the reference's own; per face: the face's)
  for (int q = 0; q < n_pts; q++) {
    const long L = left[q], R = right[q];
    const size_t gl = gidx(L);
    const double n0 = geo[gl], n1 = geo[gtot + gl], n2 = nd == 3 ? geo[2 * gtot + gl] : 0.0;
    const double sw = nd == 3 ? (n0 + n1 + std::sqrt(2.) * n2) : (n0 + n1);
    const bool neg_l = sw < 0.;
    nbr[L] = (int)R | (neg_l ? (int)0x80000000u : 0);
    nbr[R] = (int)L | (!neg_l ? (int)0x80000000u : 0);
  }
  HIP_TRY(hipMemcpy(c->nbr, nbr.data(), sizeof(int) * tot, hipMemcpyHostToDevice));
  hifr_status st = realloc_dev(&c->fgeo, sizeof(double) * geo.size());
  if (st != HIFR_OK) return st;
  HIP_TRY(hipMemcpy(c->fgeo, geo.data(), sizeof(double) * geo.size(), hipMemcpyHostToDevice));
  c->have_faces = true;
  return HIFR_OK;
}

hifr_status hifr_upload_state(hifr_ctx* c, const double* u) {
  if (!c || !u) return HIFR_ERR_ARG;
  if (!c->have_ops) return HIFR_ERR_STATE;
  drop_graphs(c);  // the mixed path's reference state is a kernel argument
  const size_t n = (size_t)c->nu * c->N * c->nf;
  if (c->mixed) {
    // reference state at rest for the FP32 deviations: the mean density and
    // total energy of the uploaded state
    const size_t np = (size_t)c->nu * c->N;
    double rho = 0.0, E = 0.0;
    for (size_t i = 0; i < np; i++) {
      rho += u[i];
      E += u[i + np * (c->nf - 1)];
    }
    rho /= (double)np;
    E /= (double)np;
    const double p = (c->cfg.gamma - 1.0) * E;
    c->ref = hifr::RefState<double>{rho, E, p, E + p};
  }
  HIP_TRY(hipMemcpyAsync(c->u, u, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(c->res, 0, sizeof(double) * n, c->stream));
  c->cur = 0;
  c->have_state = true;
  hifr_status s = trace_u(c);
  if (s != HIFR_OK) return s;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return HIFR_OK;
}

// The metrics as the kernels read them, expanded to the per-point layouts of
// hifr_set_metrics (HIFR_DETJAC_UPTS .. HIFR_NORM_FPTS).
static hifr_status download_metrics(hifr_ctx* c, int what, double* host) {
  if (!c->have_metrics) return HIFR_ERR_STATE;
  if ((what == HIFR_TDA_FPTS || what == HIFR_NORM_FPTS) && !c->have_faces) return HIFR_ERR_STATE;
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nd = c->nd, nfpf = c->nfpf;
  const size_t nv = c->affine ? N : nu * N;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (what == HIFR_DETJAC_UPTS) {
    std::vector<double> dj(nv);
    HIP_TRY(hipMemcpy(dj.data(), c->detj, sizeof(double) * nv, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nu * N; i++) host[i] = dj[c->affine ? i / nu : i];
    return HIFR_OK;
  }
  if (what == HIFR_JGINV_UPTS) {
    std::vector<double> jg(nd * nd * nv);
    if (c->mixed) {
      std::vector<float> jf(jg.size());
      HIP_TRY(hipMemcpy(jf.data(), c->jg, sizeof(float) * jf.size(), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < jf.size(); i++) jg[i] = jf[i];
    } else {
      HIP_TRY(hipMemcpy(jg.data(), c->jg, sizeof(double) * jg.size(), hipMemcpyDeviceToHost));
    }
    // component-major jg[(l*nd+m)][v] -> JGinv(l,m,p,e)
    for (size_t pe = 0; pe < nu * N; pe++) {
      const size_t v = c->affine ? pe / nu : pe;
      for (size_t l = 0; l < nd; l++)
        for (size_t m = 0; m < nd; m++) host[l + nd * m + nd * nd * pe] = jg[(l * nd + m) * nv + v];
    }
    return HIFR_OK;
  }
  const size_t gtot = c->affine ? (size_t)c->nfaces * N : nfp * N;
  std::vector<double> geo((nd + 1) * gtot);
  HIP_TRY(hipMemcpy(geo.data(), c->fgeo, sizeof(double) * geo.size(), hipMemcpyDeviceToHost));
  for (size_t fe = 0; fe < nfp * N; fe++) {
    const size_t g = c->affine ? fe / nfpf : fe;
    if (what == HIFR_TDA_FPTS) host[fe] = geo[nd * gtot + g];
    else
      for (size_t m = 0; m < nd; m++) host[fe + nfp * N * m] = geo[m * gtot + g];
  }
  return HIFR_OK;
}

hifr_status hifr_download(hifr_ctx* c, int what, double* host) {
  if (!c || !host) return HIFR_ERR_ARG;
  const size_t N = c->N, nu = c->nu, nfp = c->nfp, nf = c->nf, nd = c->nd;
  const void* src = nullptr;
  size_t n = 0, esz = sizeof(double);
  switch (what) {
    case HIFR_DISU_UPTS: src = c->u; n = nu * N * nf; break;
    case HIFR_RK_REGISTER: src = c->res; n = nu * N * nf; break;
    case HIFR_DIV_TCONF_UPTS: src = c->div; n = nu * N * nf; break;
    case HIFR_DISU_FPTS: src = c->uf[c->cur]; n = nfp * N * nf; esz = c->tsz; break;
    case HIFR_GRAD_DISU_FPTS:
      if (!c->cfg.viscous) return HIFR_ERR_ARG;
      src = c->gf; n = nfp * N * nf * nd; esz = c->tsz; break;
    case HIFR_DETJAC_UPTS:
    case HIFR_JGINV_UPTS:
    case HIFR_TDA_FPTS:
    case HIFR_NORM_FPTS: return download_metrics(c, what, host);
    default: return HIFR_ERR_ARG;
  }
  if (esz == sizeof(double) && what != HIFR_DISU_FPTS) {
    HIP_TRY(hipMemcpyAsync(host, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HIFR_OK;
  }
  // FP32 buffers (mixed path), and the field-interleaved state traces [f * nf + k],
  // which go out in the reference layout (Nfp, N, nf); mixed state traces are
  // deviations from the reference state
  std::vector<double> tmp(n);
  if (esz == sizeof(double)) {
    HIP_TRY(hipMemcpyAsync(tmp.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  } else {
    std::vector<float> tf(n);
    HIP_TRY(hipMemcpyAsync(tf.data(), src, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) tmp[i] = tf[i];
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (what == HIFR_DISU_FPTS) {
    const size_t np = nfp * N;
    const double add_rho = c->mixed ? c->ref.rho : 0.0, add_E = c->mixed ? c->ref.E : 0.0;
    for (size_t f = 0; f < np; f++)
      for (size_t k = 0; k < nf; k++)
        host[f + np * k] = tmp[f * nf + k] + (k == 0 ? add_rho : (k == nf - 1 ? add_E : 0.0));
  } else {
    for (size_t i = 0; i < n; i++) host[i] = tmp[i];
  }
  return HIFR_OK;
}

hifr_status hifr_residual(hifr_ctx* c) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (c->cfg.viscous) {
    s = trace_grad(c, true);
    if (s != HIFR_OK) return s;
  }
  StageIO io;
  io.uf_in = c->cur;
  io.fv_in = c->cfg.viscous != 0;
  io.div = true;
  return timed_launch(c, hifr::MODE_RESID, 2, io);
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
  StageIO io;
  io.uf_in = c->cur;
  io.uf_out = 1 - c->cur;
  io.fv_in = c->cfg.viscous != 0;
  io.div = write_div;
  io.adv_type = adv_type;
  io.dt = dt;
  io.rk_a = adv_type == 3 ? rk4a[stage] : 0.0;
  io.rk_b = adv_type == 3 ? rk4b[stage] : 0.0;
  hifr_status s = timed_launch(c, hifr::MODE_ADVANCE, 2, io);
  if (s != HIFR_OK) return s;
  c->cur = 1 - c->cur;
  return HIFR_OK;
}

hifr_status hifr_advance(hifr_ctx* c, int stage, int adv_type, double dt) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  return advance_impl(c, stage, adv_type, dt, true);
}

// One time step (adv_type's stages, no divergence output) as a HIP graph replay.
// The graph of the current trace-buffer parity is captured on first use from the
// same launches as the eager path, on an internal stream, and replayed on the
// context's stream.
static hifr_status graph_step(hifr_ctx* c, int adv_type, double dt) {
  const int nst = adv_type == 3 ? 5 : 1;
  if (c->g_adv_type != adv_type || c->g_dt != dt) drop_graphs(c);
  const int par = c->cur;
  if (!c->gexec[par]) {
    if (!c->gstream) HIP_TRY(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    const hipStream_t user = c->stream;
    c->stream = c->gstream;
    hipError_t e = hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal);
    hifr_status s = e == hipSuccess ? HIFR_OK : dev_err(e, "graph capture");
    for (int st = 0; st < nst && s == HIFR_OK && e == hipSuccess; st++) s = advance_impl(c, st, adv_type, dt, false);
    hipGraph_t g = nullptr;
    if (e == hipSuccess) e = hipStreamEndCapture(c->gstream, &g);
    c->stream = user;
    c->cur = par;  // capture launches nothing: the parity is the replay's to change
    if (s == HIFR_OK && e == hipSuccess) e = hipGraphInstantiate(&c->gexec[par], g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (s != HIFR_OK) return s;
    HIP_TRY(e);
    c->g_adv_type = adv_type;
    c->g_dt = dt;
  }
  HIP_TRY(hipGraphLaunch(c->gexec[par], c->stream));
  if (nst % 2) c->cur = 1 - c->cur;
  return HIFR_OK;
}

// Meshes of at most this many solution points step by graph replay: their
// kernels last a few microseconds, and one graph launch per step replaces ten
// kernel launches (20^2 quads P=3: 54 instead of 58 us per RK45 step).  Larger
// ones launch eagerly, which measured 2 % faster from 8^3 hexes P=3 up.
constexpr size_t kGraphMaxPoints = 16384;

hifr_status hifr_run(hifr_ctx* c, int n_steps, int adv_type, double dt) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (n_steps < 0) return HIFR_ERR_ARG;
  if (adv_type != 0 && adv_type != 3) return HIFR_ERR_UNSUPPORTED;
  const int nst = adv_type == 3 ? 5 : 1;
  int it = 0;
  // small meshes: every step but the last as a graph replay; the last runs
  // eagerly and writes the divergence
  if ((size_t)c->N * c->nu <= kGraphMaxPoints)
    for (; it < n_steps - 1; it++) {
      s = graph_step(c, adv_type, dt);
      if (s != HIFR_OK) return s;
    }
  for (; it < n_steps; it++)
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
  if (n_steps < 0) return HIFR_ERR_ARG;
  if (adv_type != 0 && adv_type != 3) return HIFR_ERR_UNSUPPORTED;
  const int nst = adv_type == 3 ? 5 : 1;
  const size_t need = (size_t)n_steps * 2 * 2 + 2;
  if (need > (size_t)EV_POOL) return HIFR_ERR_ARG;
  while (c->ev.size() < need) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c->ev.push_back(e);
  }
  c->ev_kind.assign(c->ev.size() / 2 + 1, -1);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  hipError_t e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e == hipSuccess) e = hipEventRecord(t0, c->stream);
  if (e != hipSuccess) {
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    HIP_TRY(e);
  }
  // events bracket the two launches of each step's first stage: around every
  // launch they would lengthen the loop by about 9 % (an event pair between
  // dependent kernels), and every stage runs the same kernels on the same arrays
  c->ev_used = 0;
  for (int it = 0; it < n_steps && s == HIFR_OK; it++)
    for (int st = 0; st < nst && s == HIFR_OK; st++) {
      c->timing = st == 0;
      s = advance_impl(c, st, adv_type, dt, it == n_steps - 1 && st == nst - 1);
    }
  c->timing = false;
  e = hipEventRecord(t1, c->stream);
  if (e == hipSuccess) e = hipEventSynchronize(t1);
  float ms = 0.f;
  if (e == hipSuccess && s == HIFR_OK) e = hipEventElapsedTime(&ms, t0, t1);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (s != HIFR_OK) return s;
  HIP_TRY(e);
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
  return HIFR_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// residual norms (CalcNormResidual): two-pass deterministic reduction

namespace {

template <int NF>
// detj holds one value per point (pts_per_det = 1) or per element (pts_per_det = Nu)
__global__ void __launch_bounds__(256) norm_partial_kernel(const double* __restrict__ div, const double* __restrict__ detj,
                                                           long npts, int pts_per_det, int norm_type, double const_src,
                                                           double* out) {
  __shared__ double sm[NF][256];
  double acc[NF];
  for (int k = 0; k < NF; k++) acc[k] = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < npts; i += (long)gridDim.x * 256L) {
    const double inv = 1.0 / detj[i / pts_per_det];
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
  const int ppd = c->affine ? c->nu : 1;  // detj entries: per element or per point
  const int nb = 512;
  if (c->nf == 5)
    hipLaunchKernelGGL(norm_partial_kernel<5>, dim3(nb), dim3(256), 0, c->stream, c->div, c->detj, npts, ppd,
                       norm_type, c->cfg.const_src, c->red);
  else
    hipLaunchKernelGGL(norm_partial_kernel<4>, dim3(nb), dim3(256), 0, c->stream, c->div, c->detj, npts, ppd,
                       norm_type, c->cfg.const_src, c->red);
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
// residual norms without allocation: the partial sums of norm_partial_kernel go to
// a file-scope device array instead of the context's reduction buffer

namespace {

constexpr int kNormBlocks = 512;
constexpr int kNormMaxFields = 5;

// per-block partials (kNormBlocks x fields); every GPU the library runs on holds
// its own copy until the process ends.  Host code uses the device address from
// hipGetSymbolAddress (norm_scratch_ptr), never this name.
static __device__ __align__(256) double scratch[kNormBlocks * kNormMaxFields];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
hifr_status norm_scratch_ptr(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return HIFR_ERR_DEVICE;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      std::fprintf(stderr, "hifr: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
      return HIFR_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = (double*)g_scratch_ptr[dev];
  return HIFR_OK;
}

// out[k] from the per-block partials, summed in block order (the order the
// allocating entry point sums them on the host)
__global__ void __launch_bounds__(64) norm_final_kernel(const double* __restrict__ part, int nb, int nf,
                                                        int norm_type, double npts, double* __restrict__ out) {
  const int k = threadIdx.x;
  if (k >= nf) return;
  double v = 0.0;
  for (int b = 0; b < nb; b++) {
    const double x = part[b * nf + k];
    v = norm_type == 0 ? fmax(v, x) : v + x;
  }
  if (norm_type == 1) v = v / npts;
  else if (norm_type == 2) v = sqrt(v) / npts;
  out[k] = v;
}

}  // namespace

extern "C" hifr_status hifr_norm_residual_noalloc(const double* div, const double* detjac, long n_pts,
                                                  int pts_per_detjac, int n_fields, int norm_type, double const_src,
                                                  void* hip_stream, double* out) {
  if (!div || !detjac || !out || n_pts <= 0 || pts_per_detjac <= 0 || norm_type < 0 || norm_type > 2)
    return HIFR_ERR_ARG;
  if (n_fields != 4 && n_fields != 5) {
    std::fprintf(stderr, "hifr_norm_residual_noalloc: %d fields; the scratch array holds 4 or 5 per block\n",
                 n_fields);
    return HIFR_ERR_ARG;
  }
  double* part = nullptr;
  hifr_status st = norm_scratch_ptr(&part);
  if (st != HIFR_OK) return st;
  const hipStream_t stream = (hipStream_t)hip_stream;
  // every block writes its partials, so nothing needs clearing between calls
  if (n_fields == 5)
    hipLaunchKernelGGL(norm_partial_kernel<5>, dim3(kNormBlocks), dim3(256), 0, stream, div, detjac, n_pts,
                       pts_per_detjac, norm_type, const_src, part);
  else
    hipLaunchKernelGGL(norm_partial_kernel<4>, dim3(kNormBlocks), dim3(256), 0, stream, div, detjac, n_pts,
                       pts_per_detjac, norm_type, const_src, part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(64), 0, stream, part, kNormBlocks, n_fields, norm_type,
                     (double)n_pts, out);
  HIP_TRY(hipGetLastError());
  return HIFR_OK;
}

// ---------------------------------------------------------------------------
// integral diagnostics (CalcIntegralQuantities, src/eles.cpp:6503-6643)

namespace {

constexpr int kMaxIntegrals = 5;
constexpr int kIntegralBlock = 256;

struct IntegralKinds {
  int n;
  int kind[kMaxIntegrals];
};

// One workgroup per element: the element's state and corrected gradient staged in
// LDS, one thread per cubature point (looping when there are more points than
// threads), the integrands weighted by w_j detJ_j and summed over the workgroup.
template <int ND>
__global__ void __launch_bounds__(kIntegralBlock)
integral_partial_kernel(int N, int nu, int ncub, const double* __restrict__ u, const double* __restrict__ gu,
                        const double* __restrict__ opp, const double* __restrict__ w,
                        const double* __restrict__ detj, IntegralKinds q, double gamma,
                        double* __restrict__ part) {
  constexpr int NF = ND + 2;
  extern __shared__ double sm[];
  double* s_u = sm;                 // [k][p]
  double* s_g = sm + NF * nu;       // [k][m][p]
  __shared__ double red[kMaxIntegrals][kIntegralBlock];
  const int e = blockIdx.x, tid = threadIdx.x;
  for (int i = tid; i < NF * nu; i += kIntegralBlock) {
    const int k = i / nu, p = i - k * nu;
    s_u[i] = u[p + (long)nu * (e + (long)N * k)];
  }
  for (int i = tid; i < NF * ND * nu; i += kIntegralBlock) {
    const int km = i / nu, p = i - km * nu;  // km = k * ND + m
    const int k = km / ND, m = km - k * ND;
    s_g[i] = gu[p + (long)nu * (e + (long)N * (k + NF * m))];
  }
  __syncthreads();
  double acc[kMaxIntegrals];
  for (int i = 0; i < kMaxIntegrals; i++) acc[i] = 0.0;
  for (int j = tid; j < ncub; j += kIntegralBlock) {
    double uc[NF], g[NF][ND];
    for (int k = 0; k < NF; k++) {
      uc[k] = 0.0;
      for (int m = 0; m < ND; m++) g[k][m] = 0.0;
    }
    for (int p = 0; p < nu; p++) {
      const double o = opp[j + (long)ncub * p];
      for (int k = 0; k < NF; k++) {
        uc[k] += o * s_u[k * nu + p];
        for (int m = 0; m < ND; m++) g[k][m] += o * s_g[(k * ND + m) * nu + p];
      }
    }
    const double irho = 1.0 / uc[0];
    // velocity gradient dv_i/dx_m = (d(rho v_i)/dx_m - v_i drho/dx_m) / rho
    double dv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int i = 0; i < ND; i++)
      for (int m = 0; m < ND; m++) dv[i][m] = irho * (g[1 + i][m] - uc[1 + i] * irho * g[0][m]);
    double tke = 0.0;
    for (int k = 1; k < NF - 1; k++) tke += 0.5 * uc[k] * uc[k];
    const double wt = w[j] * detj[j + (long)ncub * e];
    for (int iq = 0; iq < q.n; iq++) {
      double diag = 0.0;
      switch (q.kind[iq]) {
        case HIFR_KINETIC_ENERGY:
          diag = irho * tke;
          break;
        case HIFR_VORTICITY: {
          const double wz = dv[1][0] - dv[0][1];
          diag = wz * wz;
          if (ND == 3) {
            const double wx = dv[2][1] - dv[1][2], wy = dv[0][2] - dv[2][0];
            diag += wx * wx + wy * wy;
          }
          diag *= 0.5 / irho;
          break;
        }
        case HIFR_PRESSURE_DILATATION: {
          const double pr = (gamma - 1.0) * (uc[NF - 1] - irho * tke);
          double div = 0.0;
          for (int i = 0; i < ND; i++) div += dv[i][i];
          diag = pr * div;
          break;
        }
        default: {  // strain-rate (deviatoric) colon product
          double S[3][3], tr = 0.0;
          for (int i = 0; i < ND; i++)
            for (int m = 0; m < ND; m++) S[i][m] = i == m ? dv[i][i] : (dv[i][m] + dv[m][i]) / 2.0;
          for (int i = 0; i < ND; i++) tr += S[i][i] / 3.0;
          if (q.kind[iq] == HIFR_DEV_STRAIN_COLON)
            for (int i = 0; i < ND; i++) S[i][i] -= tr;
          for (int i = 0; i < ND; i++)
            for (int m = 0; m < ND; m++) diag += S[i][m] * S[i][m];
        }
      }
      acc[iq] += diag * wt;
    }
  }
  for (int iq = 0; iq < kMaxIntegrals; iq++) red[iq][tid] = acc[iq];
  __syncthreads();
  for (int st = kIntegralBlock / 2; st > 0; st >>= 1) {
    if (tid < st)
      for (int iq = 0; iq < kMaxIntegrals; iq++) red[iq][tid] += red[iq][tid + st];
    __syncthreads();
  }
  if (tid < q.n) part[(long)tid * N + e] = red[tid][0];
}

// sum of the per-element partials of integral iq = blockIdx.x (fixed order)
__global__ void __launch_bounds__(kIntegralBlock)
integral_sum_kernel(int N, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[kIntegralBlock];
  double acc = 0.0;
  for (int e = threadIdx.x; e < N; e += kIntegralBlock) acc += part[(long)blockIdx.x * N + e];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = kIntegralBlock / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

}  // namespace

extern "C" hifr_status hifr_set_cubature(hifr_ctx* c, int n_cub, const double* opp, const double* weight,
                                         const double* detjac) {
  if (!c || n_cub <= 0 || !opp || !weight || !detjac) return HIFR_ERR_ARG;
  if (c->cfg.ele_type == HIFR_TET) return HIFR_ERR_UNSUPPORTED;
  drop_graphs(c);
  const size_t nu = c->nu, N = c->N, nf = c->nf, nd = c->nd;
  if ((size_t)(nf + nf * nd) * nu * sizeof(double) > 64 * 1024) return HIFR_ERR_UNSUPPORTED;
  hifr_status st;
  if ((st = realloc_dev(&c->cub_opp, sizeof(double) * n_cub * nu)) != HIFR_OK) return st;
  if ((st = realloc_dev(&c->cub_w, sizeof(double) * n_cub)) != HIFR_OK) return st;
  if ((st = realloc_dev(&c->cub_detj, sizeof(double) * n_cub * N)) != HIFR_OK) return st;
  if ((st = realloc_dev(&c->cub_part, sizeof(double) * kMaxIntegrals * N)) != HIFR_OK) return st;
  if (!c->gu) {
    if ((st = realloc_dev(&c->gu, sizeof(double) * nu * N * nf * nd)) != HIFR_OK) return st;
    HIP_TRY(hipMemset(c->gu, 0, sizeof(double) * nu * N * nf * nd));
  }
  HIP_TRY(hipMemcpy(c->cub_opp, opp, sizeof(double) * n_cub * nu, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->cub_w, weight, sizeof(double) * n_cub, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->cub_detj, detjac, sizeof(double) * n_cub * N, hipMemcpyHostToDevice));
  c->ncub = n_cub;
  return HIFR_OK;
}

extern "C" hifr_status hifr_integral_quantities(hifr_ctx* c, int n_q, const int* kinds, double* out) {
  hifr_status s = check_ready(c);
  if (s != HIFR_OK) return s;
  if (n_q < 0 || n_q > kMaxIntegrals || (n_q > 0 && (!kinds || !out))) return HIFR_ERR_ARG;
  if (c->ncub == 0) return HIFR_ERR_STATE;
  IntegralKinds q;
  q.n = n_q;
  bool grad = false;
  for (int i = 0; i < kMaxIntegrals; i++) q.kind[i] = 0;
  for (int i = 0; i < n_q; i++) {
    if (kinds[i] < HIFR_KINETIC_ENERGY || kinds[i] > HIFR_DEV_STRAIN_COLON) return HIFR_ERR_ARG;
    q.kind[i] = kinds[i];
    grad = grad || kinds[i] != HIFR_KINETIC_ENERGY;
  }
  if (n_q == 0) return HIFR_OK;
  // the gradient comes from the viscous gradient pass of the current state
  if (grad && !c->cfg.viscous) return HIFR_ERR_UNSUPPORTED;
  if (grad) {
    if ((s = trace_u(c)) != HIFR_OK) return s;
    StageIO io;
    io.uf_in = c->cur;
    io.fv_out = true;
    io.gu_out = true;
    if ((s = launch(c, hifr::MODE_TRACE_GRAD, io)) != HIFR_OK) return s;
  }
  const size_t lds = sizeof(double) * (c->nf + c->nf * c->nd) * c->nu;
  if (c->nd == 3)
    hipLaunchKernelGGL(integral_partial_kernel<3>, dim3(c->N), dim3(kIntegralBlock), lds, c->stream, c->N, c->nu,
                       c->ncub, c->u, c->gu, c->cub_opp, c->cub_w, c->cub_detj, q, c->cfg.gamma, c->cub_part);
  else
    hipLaunchKernelGGL(integral_partial_kernel<2>, dim3(c->N), dim3(kIntegralBlock), lds, c->stream, c->N, c->nu,
                       c->ncub, c->u, c->gu, c->cub_opp, c->cub_w, c->cub_detj, q, c->cfg.gamma, c->cub_part);
  HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(integral_sum_kernel, dim3(n_q), dim3(kIntegralBlock), 0, c->stream, c->N, c->cub_part, c->red);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, c->red, sizeof(double) * n_q, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return HIFR_OK;
}

// ---------------------------------------------------------------------------
// streaming copy used to calibrate byte counters

namespace {

// one element per thread over a grid that covers the buffer: on MI355X this form
// copies at about 6.3 TB/s, against 4.6-5.3 TB/s for grid-stride loops over
// 1K-32K blocks
__global__ void __launch_bounds__(256) stream_copy_kernel(double* __restrict__ dst, const double* __restrict__ src, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" hifr_status hifr_calibrate_copy(long n, int reps, double* ms) {
  if (n <= 0 || n > (1L << 34) || reps <= 0 || !ms) return HIFR_ERR_ARG;  // grid of at most 2^26 blocks
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
    const long grid = (n + 255) / 256;
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

extern "C" hifr_status hifr_set_sgs_network(hifr_ctx* c, const hifr_mlp* m) {
  if (!hifr::mlp::kLearnedEnabled) return HIFR_ERR_UNSUPPORTED;
  if (!c) return HIFR_ERR_ARG;
  if (!(c->cfg.les && c->cfg.sgs_model == HIFR_SGS_LEARNED) || !c->mlp_buf) return HIFR_ERR_STATE;
  drop_graphs(c);  // the normalisation and mlp_on are kernel arguments
  using namespace hifr::mlp;
  char* b = (char*)c->mlp_buf;
  if (!m) {
    c->mlp_on = 0;
    HIP_TRY(hipMemsetAsync(c->mlp_buf, 0, kMlpBufBytes, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return HIFR_OK;
  }
  // the network's pending work (training on its own stream) ends before the copy
  HIP_TRY(hipStreamSynchronize(m->stream));
  const struct { size_t off; const void* src; size_t bytes; } parts[] = {
      {kMlpW1, m->w1, 2 * H * NPAD}, {kMlpW2, m->w2, 2 * H * H}, {kMlpW3, m->w3, 2 * H * H},
      {kMlpW4, m->w4, 4 * H},        {kMlpP, m->params, 4 * NPARAM}};
  for (const auto& pt : parts) {
    if (!pt.src) return HIFR_ERR_STATE;
    HIP_TRY(hipMemcpyAsync(b + pt.off, pt.src, pt.bytes, hipMemcpyDeviceToDevice, c->stream));
  }
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->mlpn = m->norm;
  c->mlp_on = 1;
  return HIFR_OK;
}

hifr_status hifr::ctx_view(const hifr_ctx* c, CtxView& v) {
  if (!c) return HIFR_ERR_ARG;
  v.ele_type = c->cfg.ele_type;
  v.order = c->cfg.order;
  v.N = c->N;
  v.have_state = c->have_state;
  v.have_metrics = c->have_metrics;
  v.affine = c->affine;
  v.mixed = c->mixed;
  v.stream = c->stream;
  v.u = c->u;
  v.detj = c->detj;
  v.jg = c->mixed ? nullptr : (const double*)c->jg;
  v.mu = c->cfg.mu_inf;
  return HIFR_OK;
}
