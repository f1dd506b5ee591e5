// Training samples of the learned eddy viscosity from a high-order hexahedral
// solution (hifr_sgs_samples, include/hifr_sgs.h): per element, a sharp modal
// filter onto degree p_lo, the Favre-filtered velocity, the SGS stress of the
// filtered products and the least-squares eddy viscosity of its deviatoric part.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/hifr_sgs.h"
#include "fr_physics.hpp"
#include "hifr_ctx_view.hpp"
#include "sgs_mlp_state.hpp"

namespace {

constexpr int MAXN = 6, MAXP = MAXN * MAXN * MAXN, NFIELD = 11, BLOCK = 256;

// One workgroup per element.  F (n_lo x n_hi) filters nodal values at the
// element's Gauss points onto the truncated Legendre expansion at the degree-p_lo
// Gauss points; D (n_lo x n_lo) differentiates the degree-p_lo interpolant there.
__global__ void __launch_bounds__(BLOCK)
sgs_samples_kernel(const double* __restrict__ u, int N, int n_hi, int n_lo, const double* __restrict__ F,
                   const double* __restrict__ D, const double* __restrict__ detj, const double* __restrict__ jg,
                   double filter_ratio, double mu, float* __restrict__ x, float* __restrict__ c_ex,
                   float* __restrict__ c_w, double* __restrict__ u_lo) {
  __shared__ double s_src[NFIELD][MAXP];  // rho, m_1..3, E, m_i m_j / rho (6)
  __shared__ double s_flt[NFIELD][MAXP];  // filtered values at the low-order points
  __shared__ double s_F[MAXN * MAXN], s_D[MAXN * MAXN];
  const int e = blockIdx.x, tid = threadIdx.x;
  const int nh3 = n_hi * n_hi * n_hi, nl3 = n_lo * n_lo * n_lo;
  const long NN = N;
  for (int i = tid; i < n_lo * n_hi; i += BLOCK) s_F[i] = F[i];
  for (int i = tid; i < n_lo * n_lo; i += BLOCK) s_D[i] = D[i];
  for (int p = tid; p < nh3; p += BLOCK) {
    double q[5];
#pragma unroll
    for (int k = 0; k < 5; k++) q[k] = u[p + nh3 * (e + NN * k)];
#pragma unroll
    for (int k = 0; k < 5; k++) s_src[k][p] = q[k];
    const double ir = 1.0 / q[0];
    s_src[5][p] = q[1] * q[1] * ir;
    s_src[6][p] = q[2] * q[2] * ir;
    s_src[7][p] = q[3] * q[3] * ir;
    s_src[8][p] = q[1] * q[2] * ir;
    s_src[9][p] = q[1] * q[3] * ir;
    s_src[10][p] = q[2] * q[3] * ir;
  }
  __syncthreads();
  // filtered value of field f at low-order point (a, b, c) = (t, s, r) indices
  for (int task = tid; task < NFIELD * nl3; task += BLOCK) {
    const int f = task / nl3, q = task - f * nl3;
    const int a = q / (n_lo * n_lo), b = (q / n_lo) % n_lo, c = q % n_lo;
    double v = 0.0;
    for (int i = 0; i < n_hi; i++) {
      double vi = 0.0;
      for (int j = 0; j < n_hi; j++) {
        double vj = 0.0;
        for (int k = 0; k < n_hi; k++) vj += s_F[c * n_hi + k] * s_src[f][(i * n_hi + j) * n_hi + k];
        vi += s_F[b * n_hi + j] * vj;
      }
      v += s_F[a * n_hi + i] * vi;
    }
    s_flt[f][q] = v;
  }
  __syncthreads();
  if (tid >= nl3) return;
  const int q = tid, a = q / (n_lo * n_lo), b = (q / n_lo) % n_lo, c = q % n_lo;
  const double rb = s_flt[0][q];
  double vt[3];
#pragma unroll
  for (int i = 0; i < 3; i++) vt[i] = s_flt[1 + i][q] / rb;
  // reference derivatives of v~ = m-bar / rho-bar along r (c), s (b), t (a)
  double dref[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    double dr = 0.0, ds = 0.0, dt = 0.0;
    for (int m = 0; m < n_lo; m++) {
      const int qr = (a * n_lo + b) * n_lo + m, qs = (a * n_lo + m) * n_lo + c, qt = (m * n_lo + b) * n_lo + c;
      dr += s_D[c * n_lo + m] * s_flt[1 + i][qr] / s_flt[0][qr];
      ds += s_D[b * n_lo + m] * s_flt[1 + i][qs] / s_flt[0][qs];
      dt += s_D[a * n_lo + m] * s_flt[1 + i][qt] / s_flt[0][qt];
    }
    dref[i][0] = dr;
    dref[i][1] = ds;
    dref[i][2] = dt;
  }
  // physical gradient with dr_d/dx_m = JGinv(d,m) / detJ (affine element)
  const double dj = detj[e];
  double vg[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int m = 0; m < 3; m++) {
      double s = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) s += dref[i][d] * jg[(d * 3 + m) * NN + e] / dj;
      vg[i][m] = s;
    }
  // SGS stress tau_ij = (m_i m_j / rho)-bar - m-bar_i m-bar_j / rho-bar
  const int pidx[3][3] = {{5, 8, 9}, {8, 6, 10}, {9, 10, 7}};
  double tau[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) tau[i][j] = s_flt[pidx[i][j]][q] - s_flt[1 + i][q] * s_flt[1 + j][q] / rb;
  double trv = vg[0][0] + vg[1][1] + vg[2][2], trt = tau[0][0] + tau[1][1] + tau[2][2];
  hifr::SgsPoint<3, double> sp;
  sp.rho = rb;
  double ss = 0.0, ts = 0.0, G2 = 0.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      sp.vg[i][j] = vg[i][j];
      sp.Sd[i][j] = 0.5 * (vg[i][j] + vg[j][i]) - (i == j ? trv / 3.0 : 0.0);
      const double td = tau[i][j] - (i == j ? trt / 3.0 : 0.0);
      ss += sp.Sd[i][j] * sp.Sd[i][j];
      ts += td * sp.Sd[i][j];
      G2 += vg[i][j] * vg[i][j];
    }
  const double G = sqrt(G2);
  const double delta = filter_ratio * cbrt(dj * 8.0) / n_lo;
  const double nu_ex = ss > 0.0 ? -ts / (2.0 * rb * ss) : 0.0;
  const double sc = delta * delta * G;
  const long si = q + (long)nl3 * e;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) x[si * 10 + 3 * i + j] = G > 0.0 ? float(vg[i][j] / G) : 0.0f;
  x[si * 10 + 9] = G > 0.0 ? float(log(sc * rb / mu)) : 0.0f;
  c_ex[si] = G > 0.0 ? float(nu_ex / sc) : 0.0f;
  c_w[si] = G > 0.0 ? float(hifr::wale_nu<3, double>(sp, delta) / sc) : 0.0f;
  if (u_lo) {
#pragma unroll
    for (int k = 0; k < 5; k++) u_lo[q + nl3 * (e + NN * k)] = s_flt[k][q];
  }
}

}  // namespace

extern "C" hifr_status hifr_sgs_samples(hifr_ctx* c, int order_lo, const double* filter_1d, const double* deriv_1d,
                                        double filter_ratio, float* x, float* c_exact, float* c_wale,
                                        double* u_lo) {
  hifr::CtxView v;
  if (!hifr::mlp::kLearnedEnabled) return HIFR_ERR_UNSUPPORTED;
  if (!c || !filter_1d || !deriv_1d || !x || !c_exact || !c_wale || !(filter_ratio > 0.0)) return HIFR_ERR_ARG;
  if (hifr::ctx_view(c, v) != HIFR_OK) return HIFR_ERR_ARG;
  if (v.ele_type != HIFR_HEX) return HIFR_ERR_UNSUPPORTED;
  if (!v.have_state || !v.have_metrics) return HIFR_ERR_STATE;
  if (!v.affine || v.mixed) return HIFR_ERR_UNSUPPORTED;  // FP64, per-element metrics (affine hexes)
  const int n_hi = v.order + 1, n_lo = order_lo + 1;
  if (order_lo < 0 || n_lo > n_hi || n_hi > MAXN) return HIFR_ERR_ARG;
  const long N = v.N, nl3 = (long)n_lo * n_lo * n_lo;
  double *dF = nullptr, *dD = nullptr, *du = nullptr;
  float *dx = nullptr, *dc = nullptr, *dw = nullptr;
  hifr_status st = HIFR_OK;
  auto cleanup = [&]() {
    void* ps[] = {dF, dD, du, dx, dc, dw};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  };
  if (hipMalloc(&dF, sizeof(double) * n_lo * n_hi) != hipSuccess ||
      hipMalloc(&dD, sizeof(double) * n_lo * n_lo) != hipSuccess ||
      hipMalloc(&dx, sizeof(float) * 10 * nl3 * N) != hipSuccess ||
      hipMalloc(&dc, sizeof(float) * nl3 * N) != hipSuccess || hipMalloc(&dw, sizeof(float) * nl3 * N) != hipSuccess ||
      (u_lo && hipMalloc(&du, sizeof(double) * 5 * nl3 * N) != hipSuccess)) {
    cleanup();
    return HIFR_ERR_DEVICE;
  }
  if (hipMemcpy(dF, filter_1d, sizeof(double) * n_lo * n_hi, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dD, deriv_1d, sizeof(double) * n_lo * n_lo, hipMemcpyHostToDevice) != hipSuccess ||
      hipStreamSynchronize(v.stream) != hipSuccess)
    st = HIFR_ERR_DEVICE;
  if (st == HIFR_OK) {
    hipLaunchKernelGGL(sgs_samples_kernel, dim3(N), dim3(BLOCK), 0, v.stream, v.u, (int)N, n_hi, n_lo, dF, dD,
                       v.detj, v.jg, filter_ratio, v.mu, dx, dc, dw, du);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(v.stream) != hipSuccess) st = HIFR_ERR_DEVICE;
  }
  if (st == HIFR_OK &&
      (hipMemcpy(x, dx, sizeof(float) * 10 * nl3 * N, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(c_exact, dc, sizeof(float) * nl3 * N, hipMemcpyDeviceToHost) != hipSuccess ||
       hipMemcpy(c_wale, dw, sizeof(float) * nl3 * N, hipMemcpyDeviceToHost) != hipSuccess ||
       (u_lo && hipMemcpy(u_lo, du, sizeof(double) * 5 * nl3 * N, hipMemcpyDeviceToHost) != hipSuccess)))
    st = HIFR_ERR_DEVICE;
  cleanup();
  return st;
}
