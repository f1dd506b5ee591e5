// Pointwise physics of the FR residual on the device: Euler and Navier-Stokes
// fluxes, the Rusanov interface flux and the LDG common solution / viscous flux.
// Formulas follow src/flux.cpp:36-431 and src/inters.cpp:311-365, 550-678.
//
// Every function is templated on the arithmetic type T (double for the FP64
// path, float for the mixed-precision path).  The mixed path works with the
// deviation du = u - u_ref from a constant reference state at rest (density
// rho_r, energy E_r): the *_dev variants return F(u) - F(u_ref), written so that
// no O(p_ref) term is formed and then cancelled.  A constant flux has zero
// divergence and zero interface jump, so the residual is unchanged.
//
// Interface functions work from the point of view of one element ("own" side,
// outward normal n) and only ever need normal components f.n of the fluxes, which
// keeps the register footprint of the face kernels small.  Both elements sharing
// a face evaluate the same expressions with (own, neighbour, n) swapped and
// negated, so they obtain the same common flux with opposite sign.
#pragma once
#include <hip/hip_runtime.h>

namespace hifr {

// Reciprocal of a positive normal number without the IEEE division sequence.
// FP64: the hardware estimate refined by two Newton-Raphson steps
// r <- r + r (1 - x r), within about an ulp of 1/x.  FP32 (mixed path): the
// hardware estimate, within an ulp.
__device__ __forceinline__ double rcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Square root: FP64 correctly rounded; FP32 the hardware estimate, within an ulp.
__device__ __forceinline__ double sqrt_fast(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }

struct Phys {
  double gamma, prandtl, mu_inf, rt_inf, c_sth, fix_vis, tau, pen_fact, const_src;
  double R_ref, T_wall;  // isothermal walls
};

// Ghost state of an isothermal no-slip wall on a static mesh (bctype 11,
// src/bdy_inters.cpp:391-420, 668-697, 884-889): pressure extrapolated from the
// trace ul, temperature T_wall, rho = p / (R_ref T_wall), velocity zero.
template <int ND, class T>
__device__ __forceinline__ void wall_ghost(const T* ul, T gamma, T R_ref, T T_wall, T* ur) {
  T v_sq = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const T v = ul[i + 1] * rcp(ul[0]);
    v_sq += v * v;
  }
  const T p = (gamma - T(1)) * (ul[ND + 1] - T(0.5) * ul[0] * v_sq);
  const T rho = p * rcp(R_ref * T_wall);
  ur[0] = rho;
#pragma unroll
  for (int i = 0; i < ND; i++) ur[i + 1] = T(0);
  ur[ND + 1] = p * rcp(gamma - T(1));
}

// Reference state at rest of the mixed-precision path (rho_r, E_r; p_r and the
// total enthalpy H_r = E_r + p_r follow).
template <class T>
struct RefState {
  T rho, E, p, H;
};

// Euler flux projected on an arbitrary vector w: out[k] = sum_m f(k,m) w_m
// (src/flux.cpp:36-128).
template <int ND, class T>
__device__ __forceinline__ void inv_flux_dot(const T* u, const T* w, T gamma, T* out) {
  const T ir = rcp(u[0]);
  T vw = 0, ke = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const T v = u[1 + i] * ir;
    vw += v * w[i];
    ke += v * v;
  }
  const T p = (gamma - T(1)) * (u[ND + 1] - T(0.5) * u[0] * ke);
  out[0] = u[0] * vw;
#pragma unroll
  for (int i = 0; i < ND; i++) out[1 + i] = u[1 + i] * vw + p * w[i];
  out[ND + 1] = vw * (u[ND + 1] + p);
}

// Deviation form: du = u - u_ref (u_ref at rest); out = (F(u) - F(u_ref)) . w.
// p - p_r = (gamma-1)(dE - rho|v|^2/2); the energy flux is v.w (H_r + dE + p - p_r).
template <int ND, class T>
__device__ __forceinline__ void inv_flux_dot_dev(const T* du, const T* w, T gamma, const RefState<T>& r, T* out) {
  const T ir = rcp(r.rho + du[0]);
  T vw = 0, mv = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const T v = du[1 + i] * ir;
    vw += v * w[i];
    mv += du[1 + i] * v;
  }
  const T dp = (gamma - T(1)) * (du[ND + 1] - T(0.5) * mv);
  out[0] = du[1] * w[0];
#pragma unroll
  for (int i = 1; i < ND; i++) out[0] += du[1 + i] * w[i];
#pragma unroll
  for (int i = 0; i < ND; i++) out[1 + i] = du[1 + i] * vw + dp * w[i];
  out[ND + 1] = vw * (r.H + (du[ND + 1] + dp));
}

// Full state from a deviation (for the viscous terms of the mixed path).
template <int ND, class T>
__device__ __forceinline__ void add_ref(const T* du, const RefState<T>& r, T* u) {
  u[0] = r.rho + du[0];
#pragma unroll
  for (int i = 0; i < ND; i++) u[1 + i] = du[1 + i];
  u[ND + 1] = r.E + du[ND + 1];
}

// Dynamic viscosity.  SUTH = false: constant viscosity mu_inf (fix_vis = 1, every
// workload here); SUTH = true: Sutherland's law blended by fix_vis
// (src/flux.cpp:308-310), with rt^1.5 formed as rt*sqrt(rt), within an ulp of pow.
// The choice is a template parameter so that the constant-viscosity kernels do not
// carry the register footprint of the Sutherland branch.
template <bool SUTH, class T>
__device__ __forceinline__ T viscosity(const Phys& ph, T inte) {
  if constexpr (!SUTH) {
    return T(ph.mu_inf);
  } else {
    const T rt = (T(ph.gamma) - T(1)) * inte / T(ph.rt_inf);
    const T mu = T(ph.mu_inf) * (rt * sqrt(rt)) * (T(1) + T(ph.c_sth)) / (rt + T(ph.c_sth));
    return mu + T(ph.fix_vis) * (T(ph.mu_inf) - mu);
  }
}

// Laminar viscous stress tau and energy flux column of the Navier-Stokes viscous
// flux (src/flux.cpp:263-431): F_v(1+i, m) = -tau(i,m), F_v(nf-1, m) = qe(m).
template <int ND, class T = double>
struct ViscT {
  T tau[ND][ND];
  T qe[ND];
};

// Viscous stress and heat-flux column from the state u and the physical gradient
// g(k, m) = d u_k / d x_m, supplied one column (d/dx_m of all fields) at a time by
// gcol(m, out[NF]) so that only one column is live.
template <int ND, bool SUTH, class T, class GCol>
__device__ __forceinline__ void visc_tensor_cols(const T* u, GCol&& gcol, const Phys& ph, ViscT<ND, T>& t) {
  constexpr int NF = ND + 2;
  const T rho = u[0], irho = rcp(rho);
  T v[ND];
  T ke = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    v[i] = u[1 + i] * irho;
    ke += v[i] * v[i];
  }
  const T inte = u[ND + 1] * irho - T(0.5) * ke;
  const T mu = viscosity<SUTH, T>(ph, inte);
  T dv[ND][ND];
  T de[ND];
#pragma unroll
  for (int m = 0; m < ND; m++) {
    T g[NF];
    gcol(m, g);
    T rdv = 0;
#pragma unroll
    for (int i = 0; i < ND; i++) {
      dv[i][m] = (g[1 + i] - g[0] * v[i]) * irho;
      rdv += v[i] * dv[i][m];
    }
    const T dke = T(0.5) * ke * g[0] + rho * rdv;
    de[m] = (g[ND + 1] - dke - g[0] * inte) * irho;
    __builtin_amdgcn_sched_barrier(0);
  }
  T diag = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) diag += dv[i][i];
  diag = diag * T(1.0 / 3.0);
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++)
      t.tau[i][j] = (i == j) ? T(2) * mu * (dv[i][i] - diag) : mu * (dv[i][j] + dv[j][i]);
  const T kq = mu * T(ph.gamma / ph.prandtl);
#pragma unroll
  for (int m = 0; m < ND; m++) {
    T w = 0;
#pragma unroll
    for (int i = 0; i < ND; i++) w += v[i] * t.tau[i][m];
    t.qe[m] = -(w + kq * de[m]);
  }
}

// The viscous tensor of visc_tensor_cols in three steps, for callers that receive
// the gradient columns in separate passes: primitives of the state, one gradient
// column at a time, then stress and heat flux.  (visc_tensor_cols keeps its fused
// form: its callers are register-bound kernels.)
template <int ND, class T>
struct ViscPrim {
  T rho, irho, v[ND], ke, inte;
};
template <int ND, class T>
__device__ __forceinline__ void visc_prim(const T* u, ViscPrim<ND, T>& q) {
  q.rho = u[0];
  q.irho = rcp(q.rho);
  q.ke = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    q.v[i] = u[1 + i] * q.irho;
    q.ke += q.v[i] * q.v[i];
  }
  q.inte = u[ND + 1] * q.irho - T(0.5) * q.ke;
}
// dv[i][m] = d v_i / d x_m and de = d e / d x_m from g[k] = d u_k / d x_m
template <int ND, class T>
__device__ __forceinline__ void visc_column(const ViscPrim<ND, T>& q, const T* g, T (&dv)[ND][ND], int m, T& de) {
  T rdv = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    dv[i][m] = (g[1 + i] - g[0] * q.v[i]) * q.irho;
    rdv += q.v[i] * dv[i][m];
  }
  const T dke = T(0.5) * q.ke * g[0] + q.rho * rdv;
  de = (g[ND + 1] - dke - g[0] * q.inte) * q.irho;
}
template <int ND, bool SUTH, class T>
__device__ __forceinline__ void visc_finish(const ViscPrim<ND, T>& q, const T (&dv)[ND][ND], const T* de,
                                            const Phys& ph, ViscT<ND, T>& t) {
  const T mu = viscosity<SUTH, T>(ph, q.inte);
  T diag = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) diag += dv[i][i];
  diag = diag * T(1.0 / 3.0);
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++)
      t.tau[i][j] = (i == j) ? T(2) * mu * (dv[i][i] - diag) : mu * (dv[i][j] + dv[j][i]);
  const T kq = mu * T(ph.gamma / ph.prandtl);
#pragma unroll
  for (int m = 0; m < ND; m++) {
    T w = 0;
#pragma unroll
    for (int i = 0; i < ND; i++) w += q.v[i] * t.tau[i][m];
    t.qe[m] = -(w + kq * de[m]);
  }
}

// Eddy-viscosity SGS flux at a solution point, as the reference CPU build
// computes it (src/eles.cpp:2456-2778, no wall model): MODEL 1 Smagorinsky
// (C_s = 0.1), 2 WALE (C_w = 0.5, eps = 1e-12), turbulent Prandtl number 0.5,
// filter width delta = filter_ratio * (detJ * vref)^(1/d) / (P+1) with the
// reference-element volume vref (4 quads, 8 hexes).  Two index slips of that code
// are kept: its "velocity" is u_i = U_i / rho for i < d (density / rho in slot 0,
// src/eles.cpp:2483) and du(i,j) takes the density derivative along x_j
// (src/eles.cpp:2656).  u: conservative state; g[k][m] = d U_k / d x_m.
// Output: momentum flux M[i][j] = F_sgs(1+i, j), energy flux E[j] = F_sgs(d+1, j);
// the density flux is zero.
template <int ND, int MODEL, class T>
__device__ __forceinline__ void sgs_flux(const T* u, const T (*g)[ND], T detj_vref, T filter_ratio, int order,
                                         T gamma, T (*M)[ND], T* E) {
  // Divisions are products with a reciprocal (rcp: within an ulp) and the powers
  // are products of square roots: a few ulps from the reference's quotients and
  // pow(), for a quarter of the instructions.
  constexpr int NF = ND + 2;
  const T rho = u[0], ir = rcp(rho);
  const T third = T(1.0 / 3.0);
  T uu[ND];
  T ke = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    uu[i] = u[i] * ir;
    ke += T(0.5) * (uu[i] * uu[i]);
  }
  const T inte = u[NF - 1] * ir - ke;
  const T delta = filter_ratio * (ND == 3 ? cbrt(detj_vref) : sqrt(detj_vref)) * rcp(T(order) + T(1));
  T du[ND][ND], de[ND], S[ND][ND];
#pragma unroll
  for (int i = 0; i < ND; i++) {
    T dke = ke * g[0][i];
#pragma unroll
    for (int j = 0; j < ND; j++) {
      du[i][j] = (g[j + 1][i] - uu[j] * g[0][j]) * ir;
      dke += rho * uu[j] * du[i][j];
    }
    de[i] = (g[NF - 1][i] - dke - g[0][i] * inte) * ir;
  }
  T diag = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
#pragma unroll
    for (int j = 0; j < ND; j++) S[i][j] = (du[i][j] + du[j][i]) * T(0.5);
    diag += S[i][i] * third;
  }
#pragma unroll
  for (int i = 0; i < ND; i++) S[i][i] -= diag;
  T mu_t;
  if constexpr (MODEL == 1) {
    T smod = 0;
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) smod += T(2) * S[i][j] * S[i][j];
    smod = sqrt(smod);
    const T cs = T(0.1);
    mu_t = rho * cs * cs * delta * delta * smod;
  } else {
    T num = 0, den = 0, d2 = 0;
    T Sq[ND][ND];
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) {
        T q = 0;
#pragma unroll
        for (int k = 0; k < ND; ++k) q += (du[i][k] * du[k][j] + du[j][k] * du[k][i]) * T(0.5);
        Sq[i][j] = q;
        d2 += du[i][j] * du[j][i] * third;
      }
#pragma unroll
    for (int i = 0; i < ND; i++) Sq[i][i] -= d2;
#pragma unroll
    for (int i = 0; i < ND; i++)
#pragma unroll
      for (int j = 0; j < ND; j++) {
        num += Sq[i][j] * Sq[i][j];
        den += S[i][j] * S[i][j];
      }
    // den^(5/2) + num^(5/4) and num^(3/2)
    const T sn = sqrt(num);
    den = den * den * sqrt(den) + num * sqrt(sn);
    num = num * sn;
    const T cw = T(0.5);
    mu_t = rho * cw * cw * delta * delta * num * rcp(den + T(1.e-12));
  }
  const T ek = -gamma * mu_t / T(0.5);
#pragma unroll
  for (int j = 0; j < ND; j++) {
    E[j] = ek * de[j];
#pragma unroll
    for (int i = 0; i < ND; i++) M[i][j] = T(-2) * mu_t * S[i][j];
  }
}

// Eddy-viscosity closure in the form of the reference's GPU build
// (SGS_flux_kernel, src/cuda_kernels.cu:1466-1660, with the velocity and energy
// gradients of vis_NS_flux, :880-960): velocity gradient vg[i][j] = d v_i / d x_j,
// deviatoric strain Sd, internal-energy gradient de.  Used by the learned closure
// and its WALE fallback.
template <int ND, class T>
struct SgsPoint {
  T rho;
  T vg[ND][ND];
  T Sd[ND][ND];
  T de[ND];
};

template <int ND, class T>
__device__ __forceinline__ void sgs_point(const T* u, const T (*g)[ND], SgsPoint<ND, T>& o) {
  constexpr int NF = ND + 2;
  const T rho = u[0], ir = rcp(rho);
  T v[ND], ke = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    v[i] = u[1 + i] * ir;
    ke += v[i] * v[i];
  }
  o.rho = rho;
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) o.vg[i][j] = (g[1 + i][j] - v[i] * g[0][j]) * ir;
  // e = E/rho - |v|^2/2:  de_j = (dE_j - (E/rho) drho_j) / rho - v_i vg[i][j]
  const T Er = u[NF - 1] * ir;
#pragma unroll
  for (int j = 0; j < ND; j++) {
    T s = (g[NF - 1][j] - Er * g[0][j]) * ir;
#pragma unroll
    for (int i = 0; i < ND; i++) s -= v[i] * o.vg[i][j];
    o.de[j] = s;
  }
  T tr = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) tr += o.vg[i][i];
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) o.Sd[i][j] = T(0.5) * (o.vg[i][j] + o.vg[j][i]) - (i == j ? tr * T(1.0 / 3.0) : T(0));
  (void)ke;
}

// WALE kinematic eddy viscosity (C_w = 0.5, eps = 1e-10) of the GPU form:
// nu_t = (C_w delta)^2 (Sq:Sq)^(3/2) / ((Sd:Sd)^(5/2) + (Sq:Sq)^(5/4) + eps) with
// Sq = sym(vg vg) - tr(vg vg)/3 I.
template <int ND, class T>
__device__ __forceinline__ T wale_nu(const SgsPoint<ND, T>& o, T delta) {
  T num = 0, den = 0, d2 = 0;
  T Sq[ND][ND];
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) {
      T q = 0;
#pragma unroll
      for (int k = 0; k < ND; k++) q += T(0.5) * (o.vg[i][k] * o.vg[k][j] + o.vg[j][k] * o.vg[k][i]);
      Sq[i][j] = q;
    }
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int k = 0; k < ND; k++) d2 += o.vg[i][k] * o.vg[k][i];
#pragma unroll
  for (int i = 0; i < ND; i++) Sq[i][i] -= d2 * T(1.0 / 3.0);
#pragma unroll
  for (int i = 0; i < ND; i++)
#pragma unroll
    for (int j = 0; j < ND; j++) {
      num += Sq[i][j] * Sq[i][j];
      den += o.Sd[i][j] * o.Sd[i][j];
    }
  // the powers as products of square roots, the quotient with a reciprocal
  const T cw = T(0.5), sn = sqrt(num);
  return cw * cw * delta * delta * (num * sn) * rcp(den * den * sqrt(den) + num * sqrt(sn) + T(1.e-10));
}

// SGS flux of an eddy viscosity mu_t (GPU form): M[i][j] = -2 mu_t Sd[i][j],
// E[j] = -gamma mu_t / Pr_t de[j] with Pr_t = 0.5.
template <int ND, class T>
__device__ __forceinline__ void sgs_flux_of(const SgsPoint<ND, T>& o, T mu_t, T gamma, T (*M)[ND], T* E) {
  const T ek = -gamma * mu_t / T(0.5);
#pragma unroll
  for (int j = 0; j < ND; j++) {
    E[j] = ek * o.de[j];
#pragma unroll
    for (int i = 0; i < ND; i++) M[i][j] = T(-2) * mu_t * o.Sd[i][j];
  }
}

// Viscous flux projected on w: out[k] = sum_m F_v(k,m) w_m
template <int ND, class T>
__device__ __forceinline__ void visc_flux_dot(const ViscT<ND, T>& t, const T* w, T* out) {
  out[0] = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    T s = 0;
#pragma unroll
    for (int m = 0; m < ND; m++) s += t.tau[i][m] * w[m];
    out[1 + i] = -s;
  }
  T s = 0;
#pragma unroll
  for (int m = 0; m < ND; m++) s += t.qe[m] * w[m];
  out[ND + 1] = s;
}

// Rusanov common normal flux for outward normal n (src/inters.cpp:311-365, static mesh).
template <int ND, class T>
__device__ __forceinline__ void rusanov(const T* uo, const T* un, const T* n, T gamma, T* fn) {
  constexpr int NF = ND + 2;
  T fo[NF], fb[NF];
  inv_flux_dot<ND, T>(uo, n, gamma, fo);
  inv_flux_dot<ND, T>(un, n, gamma, fb);
  T vno = 0, vnn = 0, ko = 0, kn = 0;
  const T iro = rcp(uo[0]), irn = rcp(un[0]);
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const T a = uo[1 + i] * iro, b = un[1 + i] * irn;
    vno += a * n[i];
    vnn += b * n[i];
    ko += a * a;
    kn += b * b;
  }
  const T po = (gamma - T(1)) * (uo[ND + 1] - T(0.5) * uo[0] * ko);
  const T pn = (gamma - T(1)) * (un[ND + 1] - T(0.5) * un[0] * kn);
  const T vn_av = fabs(T(0.5) * (vno + vnn));
  const T c_av = sqrt_fast(gamma * (po + pn) * rcp(uo[0] + un[0]));
  const T eig = fabs(vn_av + c_av);
#pragma unroll
  for (int k = 0; k < NF; k++) fn[k] = T(0.5) * ((fo[k] + fb[k]) - eig * (un[k] - uo[k]));
}

// Roe common normal flux with Harten's entropy fix for the left state ul, right
// state ur and the left element's outward normal n (src/inters.cpp:390-520, static
// mesh).  The reference evaluates the central part and the pressures in 2-D only;
// in 3-D the same expressions take the third velocity component.  Unlike the
// Rusanov flux, its Roe averages are not symmetric in rounding under a swap of the
// sides, so both elements of an interface evaluate it in the left element's
// orientation (the right one negates the result).
template <int ND, class T>
__device__ __forceinline__ void roe(const T* ul, const T* ur, const T* n, T gamma, T* fn) {
  constexpr int NF = ND + 2;
  const T g1 = gamma - T(1);
  const T irl = T(1) / ul[0], irr = T(1) / ur[0];
  T vl[ND], vr[ND], kl = 0, kr = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    vl[i] = ul[1 + i] * irl;
    vr[i] = ur[1 + i] * irr;
    kl += vl[i] * vl[i];
    kr += vr[i] * vr[i];
  }
  const T pl = g1 * (ul[ND + 1] - T(0.5) * ul[0] * kl);
  const T pr = g1 * (ur[ND + 1] - T(0.5) * ur[0] * kr);
  const T hl = (ul[ND + 1] + pl) * irl, hr = (ur[ND + 1] + pr) * irr;
  // Roe averages
  const T sq = sqrt(ur[0] * irl);
  const T w = T(1) / (sq + T(1));
  T um[ND], usq = 0, unm = 0, mnl = 0, mnr = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    um[i] = w * (vl[i] + sq * vr[i]);
    usq += T(0.5) * um[i] * um[i];
    unm += um[i] * n[i];
    mnl += ul[1 + i] * n[i];
    mnr += ur[1 + i] * n[i];
  }
  const T hm = w * (hl + sq * hr);
  const T am2 = g1 * (hm - usq);
  const T am = sqrt(am2);
  // central part (twice the average of the two normal fluxes)
  fn[0] = mnl + mnr;
#pragma unroll
  for (int i = 0; i < ND; i++) fn[1 + i] = mnl * vl[i] + mnr * vr[i] + (pl + pr) * n[i];
  fn[ND + 1] = mnl * hl + mnr * hr;
  // wave speeds with the entropy fix
  T l0 = fabs(unm), lp = fabs(unm + am), lm = fabs(unm - am);
  const T eps = T(0.5) * (fabs(mnl * irl - mnr * irr) + fabs(sqrt(gamma * pl * irl) - sqrt(gamma * pr * irr)));
  const T e2 = T(2) * eps;
  if (l0 < e2) l0 = T(0.25) * l0 * l0 / eps + eps;
  if (lp < e2) lp = T(0.25) * lp * lp / eps + eps;
  if (lm < e2) lm = T(0.25) * lm * lm / eps + eps;
  const T a2 = T(0.5) * (lp + lm) - l0;
  const T a3 = T(0.5) * (lp - lm) / am;
  const T a1 = a2 * g1 / am2;
  const T a4 = a3 * g1;
  T du[NF];
#pragma unroll
  for (int k = 0; k < NF; k++) du[k] = ur[k] - ul[k];
  T a5 = usq * du[0], a6 = unm * du[0];
#pragma unroll
  for (int i = 0; i < ND; i++) {
    a5 -= um[i] * du[1 + i];
    a6 -= n[i] * du[1 + i];
  }
  a5 += du[ND + 1];
  const T aL = a1 * a5 - a3 * a6;
  const T bL = a4 * a5 - a2 * a6;
  fn[0] = T(0.5) * (fn[0] - (l0 * du[0] + aL));
#pragma unroll
  for (int i = 0; i < ND; i++) fn[1 + i] = T(0.5) * (fn[1 + i] - (l0 * du[1 + i] + aL * um[i] + bL * n[i]));
  fn[ND + 1] = T(0.5) * (fn[ND + 1] - (l0 * du[ND + 1] + aL * hm + bL * unm));
}

// Roe flux at an interface point seen from one side: own trace index fe, state
// uo and outward normal n; the neighbour's trace index nb and state un.  Both sides
// evaluate roe() with the state of the smaller trace index first and that side's
// normal (the other side's normal is its exact negation), so the two get the same
// value with opposite signs and the interface flux is exactly conservative.
template <int ND, class T>
__device__ __forceinline__ void roe_pair(long fe, long nb, const T* uo, const T* un, const T* n, T gamma, T* fn) {
  if (fe < nb) {
    roe<ND, T>(uo, un, n, gamma, fn);
  } else {
    T nl[ND];
#pragma unroll
    for (int l = 0; l < ND; l++) nl[l] = -n[l];
    roe<ND, T>(un, uo, nl, gamma, fn);
#pragma unroll
    for (int k = 0; k < ND + 2; k++) fn[k] = -fn[k];
  }
}

// Rusanov flux of the mixed path on deviations: (F_c - F(u_ref)) . n.
template <int ND, class T>
__device__ __forceinline__ void rusanov_dev(const T* duo, const T* dun, const T* n, T gamma, const RefState<T>& r,
                                            T* fn) {
  constexpr int NF = ND + 2;
  T fo[NF], fb[NF];
  inv_flux_dot_dev<ND, T>(duo, n, gamma, r, fo);
  inv_flux_dot_dev<ND, T>(dun, n, gamma, r, fb);
  const T ro = r.rho + duo[0], rn = r.rho + dun[0];
  const T iro = rcp(ro), irn = rcp(rn);
  T vno = 0, vnn = 0, mvo = 0, mvn = 0;
#pragma unroll
  for (int i = 0; i < ND; i++) {
    const T a = duo[1 + i] * iro, b = dun[1 + i] * irn;
    vno += a * n[i];
    vnn += b * n[i];
    mvo += duo[1 + i] * a;
    mvn += dun[1 + i] * b;
  }
  const T po = r.p + (gamma - T(1)) * (duo[ND + 1] - T(0.5) * mvo);
  const T pn = r.p + (gamma - T(1)) * (dun[ND + 1] - T(0.5) * mvn);
  const T vn_av = fabs(T(0.5) * (vno + vnn));
  const T c_av = sqrt_fast(gamma * (po + pn) * rcp(ro + rn));
  const T eig = fabs(vn_av + c_av);
#pragma unroll
  for (int k = 0; k < NF; k++) fn[k] = T(0.5) * ((fo[k] + fb[k]) - eig * (dun[k] - duo[k]));
}

// LDG viscous common normal flux, interior form (flux_spec 0, src/inters.cpp:550-641):
//   f_c = 1/2 (f_o + f_n) + pen n [n.(f_o - f_n)] + tau n (u_o - u_n),  fn = f_c . n
// written with the normal components a = f_o.n, b = f_n.n and nn = n.n:
//   fn = 1/2 (a + b) + (pen (a - b) + tau (u_o - u_n)) nn
template <class T>
__device__ __forceinline__ T ldg_normal(T a, T b, T uo, T un, T nn, T pen, T tau) {
  return T(0.5) * (a + b) + (pen * (a - b) + tau * (uo - un)) * nn;
}

}  // namespace hifr
