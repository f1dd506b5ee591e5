// Pointwise physics of the FR residual on the device: Euler and Navier-Stokes
// fluxes, the Rusanov interface flux and the LDG common solution / viscous flux.
// Formulas follow src/flux.cpp:36-431 and src/inters.cpp:311-365, 550-678.
//
// Interface functions work from the point of view of one element ("own" side,
// outward normal n) and only ever need normal components f.n of the fluxes, which
// keeps the register footprint of the face kernels small.  Both elements sharing
// a face evaluate the same expressions with (own, neighbour, n) swapped and
// negated, so they obtain the same common flux with opposite sign.
#pragma once
#include <hip/hip_runtime.h>

namespace hifr {

struct Phys {
  double gamma, prandtl, mu_inf, rt_inf, c_sth, fix_vis, tau, pen_fact, const_src;
};

// Euler flux projected on an arbitrary vector w: out[k] = sum_m f(k,m) w_m
// (src/flux.cpp:36-128).
template <int ND>
__device__ __forceinline__ void inv_flux_dot(const double* u, const double* w, double gamma, double* out) {
  const double ir = 1.0 / u[0];
  double vw = 0.0, ke = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const double v = u[1 + i] * ir;
    vw += v * w[i];
    ke += v * v;
  }
  const double p = (gamma - 1.0) * (u[ND + 1] - 0.5 * u[0] * ke);
  out[0] = u[0] * vw;
#pragma unroll
  for (int i = 0; i < ND; i++) out[1 + i] = u[1 + i] * vw + p * w[i];
  out[ND + 1] = vw * (u[ND + 1] + p);
}

// Dynamic viscosity.  SUTH = false: constant viscosity mu_inf (fix_vis = 1, every
// workload here); SUTH = true: Sutherland's law blended by fix_vis
// (src/flux.cpp:308-310), with rt^1.5 formed as rt*sqrt(rt), within an ulp of pow.
// The choice is a template parameter so that the constant-viscosity kernels do not
// carry the register footprint of the Sutherland branch.
template <bool SUTH>
__device__ __forceinline__ double viscosity(const Phys& ph, double inte) {
  if constexpr (!SUTH) {
    return ph.mu_inf;
  } else {
    const double rt = (ph.gamma - 1.0) * inte / ph.rt_inf;
    const double mu = ph.mu_inf * (rt * sqrt(rt)) * (1.0 + ph.c_sth) / (rt + ph.c_sth);
    return mu + ph.fix_vis * (ph.mu_inf - mu);
  }
}

// Laminar viscous stress tau and energy flux column of the Navier-Stokes viscous
// flux (src/flux.cpp:263-431): F_v(1+i, m) = -tau(i,m), F_v(nf-1, m) = qe(m).
template <int ND>
struct ViscT {
  double tau[ND][ND];
  double qe[ND];
};

// g(k, m) = d u_k / d x_m (physical gradient of the conservative state)
template <int ND, bool SUTH>
__device__ __forceinline__ void visc_tensor(const double* u, const double (*g)[ND], const Phys& ph, ViscT<ND>& t) {
  const double rho = u[0], irho = 1.0 / rho;
  double v[ND];
  double ke = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    v[i] = u[1 + i] * irho;
    ke += v[i] * v[i];
  }
  const double inte = u[ND + 1] * irho - 0.5 * ke;
  const double mu = viscosity<SUTH>(ph, inte);
  double dv[ND][ND];
  double de[ND];
#pragma unroll
  for (int m = 0; m < ND; m++) {
    double rdv = 0.0;
#pragma unroll
    for (int i = 0; i < ND; i++) {
      dv[i][m] = (g[1 + i][m] - g[0][m] * v[i]) * irho;
      rdv += v[i] * dv[i][m];
    }
    const double dke = 0.5 * ke * g[0][m] + rho * rdv;
    de[m] = (g[ND + 1][m] - dke - g[0][m] * inte) * irho;
  }
  double diag = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) diag += dv[i][i];
  diag = diag / 3.0;
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++)
      t.tau[i][j] = (i == j) ? 2.0 * mu * (dv[i][i] - diag) : mu * (dv[i][j] + dv[j][i]);
  const double kq = mu / ph.prandtl * ph.gamma;
#pragma unroll
  for (int m = 0; m < ND; m++) {
    double w = 0.0;
#pragma unroll
    for (int i = 0; i < ND; i++) w += v[i] * t.tau[i][m];
    t.qe[m] = -(w + kq * de[m]);
  }
}


// Same as visc_tensor with the gradient supplied one column (d/dx_m of all
// fields) at a time by gcol(m, out[NF]); keeps only one column live.
template <int ND, bool SUTH, class GCol>
__device__ __forceinline__ void visc_tensor_cols(const double* u, GCol&& gcol, const Phys& ph, ViscT<ND>& t) {
  constexpr int NF = ND + 2;
  const double rho = u[0], irho = 1.0 / rho;
  double v[ND];
  double ke = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    v[i] = u[1 + i] * irho;
    ke += v[i] * v[i];
  }
  const double inte = u[ND + 1] * irho - 0.5 * ke;
  const double mu = viscosity<SUTH>(ph, inte);
  double dv[ND][ND];
  double de[ND];
#pragma unroll
  for (int m = 0; m < ND; m++) {
    double g[NF];
    gcol(m, g);
    double rdv = 0.0;
#pragma unroll
    for (int i = 0; i < ND; i++) {
      dv[i][m] = (g[1 + i] - g[0] * v[i]) * irho;
      rdv += v[i] * dv[i][m];
    }
    const double dke = 0.5 * ke * g[0] + rho * rdv;
    de[m] = (g[ND + 1] - dke - g[0] * inte) * irho;
    __builtin_amdgcn_sched_barrier(0);
  }
  double diag = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) diag += dv[i][i];
  diag = diag / 3.0;
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++)
      t.tau[i][j] = (i == j) ? 2.0 * mu * (dv[i][i] - diag) : mu * (dv[i][j] + dv[j][i]);
  const double kq = mu / ph.prandtl * ph.gamma;
#pragma unroll
  for (int m = 0; m < ND; m++) {
    double w = 0.0;
#pragma unroll
    for (int i = 0; i < ND; i++) w += v[i] * t.tau[i][m];
    t.qe[m] = -(w + kq * de[m]);
  }
}

// Viscous flux projected on w: out[k] = sum_m F_v(k,m) w_m
template <int ND>
__device__ __forceinline__ void visc_flux_dot(const ViscT<ND>& t, const double* w, double* out) {
  out[0] = 0.0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    double s = 0.0;
#pragma unroll
    for (int m = 0; m < ND; m++) s += t.tau[i][m] * w[m];
    out[1 + i] = -s;
  }
  double s = 0.0;
#pragma unroll
  for (int m = 0; m < ND; m++) s += t.qe[m] * w[m];
  out[ND + 1] = s;
}

// Rusanov common normal flux for outward normal n (src/inters.cpp:311-365, static mesh).
template <int ND>
__device__ __forceinline__ void rusanov(const double* uo, const double* un, const double* n, double gamma,
                                        double* fn) {
  constexpr int NF = ND + 2;
  double fo[NF], fb[NF];
  inv_flux_dot<ND>(uo, n, gamma, fo);
  inv_flux_dot<ND>(un, n, gamma, fb);
  double vno = 0.0, vnn = 0.0, ko = 0.0, kn = 0.0;
  const double iro = 1.0 / uo[0], irn = 1.0 / un[0];
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const double a = uo[1 + i] * iro, b = un[1 + i] * irn;
    vno += a * n[i];
    vnn += b * n[i];
    ko += a * a;
    kn += b * b;
  }
  const double po = (gamma - 1.0) * (uo[ND + 1] - 0.5 * uo[0] * ko);
  const double pn = (gamma - 1.0) * (un[ND + 1] - 0.5 * un[0] * kn);
  const double vn_av = fabs(0.5 * (vno + vnn));
  const double c_av = sqrt(gamma * (po + pn) / (uo[0] + un[0]));
  const double eig = fabs(vn_av + c_av);
#pragma unroll
  for (int k = 0; k < NF; k++) fn[k] = 0.5 * ((fo[k] + fb[k]) - eig * (un[k] - uo[k]));
}

// LDG viscous common normal flux, interior form (flux_spec 0, src/inters.cpp:550-641):
//   f_c = 1/2 (f_o + f_n) + pen n [n.(f_o - f_n)] + tau n (u_o - u_n),  fn = f_c . n
// written with the normal components a = f_o.n, b = f_n.n and nn = n.n:
//   fn = 1/2 (a + b) + (pen (a - b) + tau (u_o - u_n)) nn
template <int ND>
__device__ __forceinline__ double ldg_normal(double a, double b, double uo, double un, double nn, double pen,
                                             double tau) {
  return 0.5 * (a + b) + (pen * (a - b) + tau * (uo - un)) * nn;
}

}  // namespace hifr
