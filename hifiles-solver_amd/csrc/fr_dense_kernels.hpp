// Element-block kernels of the FR residual for tetrahedra on gfx950: the same
// stage split as the tensor-product kernels (fr_tensor_kernels.hpp), with the dense
// element operators of the reference (src/eles.cpp:3620-4052, src/eles_tets.cpp:
// 972-1131) applied as small matrix-vector products from LDS:
//   I = opp_0 (Nfp x Nu)   solution -> flux points
//   D = opp_4 (d x Nu x Nu) reference-space derivative (opp_2 = opp_4)
//   L = opp_3 (Nu x Nfp)   correction lift (DG for tets); opp_5 = L o tnorm
// and the reference tet's face normals tnorm (face 0: (1,1,1)/sqrt(3), faces 1-3:
// -e_r, -e_s, -e_t; src/eles_tets.cpp:589-620), so opp_1 = I o tnorm.
// One thread per solution point for the volume passes and one task per flux point
// for the face passes; EPB elements per workgroup share the operators in LDS.
#pragma once
#include "fr_tensor_kernels.hpp"

namespace hifr {

// one wavefront per workgroup: EPB = 64 / Nu elements (16, 6, 3 at P = 1, 2, 3)
template <int P>
struct TetP {
  static constexpr int NU = (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NFPF = (P + 1) * (P + 2) / 2;
  static constexpr int NFP = 4 * NFPF;
  static constexpr int EPB = 64 / NU > 0 ? 64 / NU : 1;
  static constexpr int BLOCK = ((EPB * NU + 63) / 64) * 64;
};

// reference normal of a tet face: component c
template <class R>
__device__ __forceinline__ R tet_tnorm(int face, int c) {
  if (face == 0) return R(0.57735026918962576);  // 1/sqrt(3)
  return face - 1 == c ? R(-1) : R(0);
}

template <int P, int VIS, int MODE, bool AFF, bool MIX, int LES>
__global__ void __launch_bounds__((TetP<P>::BLOCK))
fr_dense_kernel(const StageArgsT<std::conditional_t<MIX, float, double>> a) {
  using R = std::conditional_t<MIX, float, double>;
  constexpr int ND = 3, NF = 5, NFACE = 4;
  constexpr bool VISC = VIS != 0;
  constexpr bool SUTH = VIS == 2;
  static_assert(LES == 0 || VISC, "LES needs the viscous kernels");
  using T = TetP<P>;
  constexpr int NU = T::NU, NFP = T::NFP, NFPF = T::NFPF, EPB = T::EPB, BLOCK = T::BLOCK;
  // LDS: operators, then per element u, face values, gradient / flux, SGS flux
  constexpr int SZ_I = NFP * NU, SZ_D = ND * NU * NU, SZ_L = NU * NFP;
  constexpr int SZ_U = EPB * NF * NU, SZ_Q = (MODE == MODE_TRACE_U) ? 0 : EPB * NFP * NF;
  constexpr int SZ_G = (MODE == MODE_TRACE_U) ? 0 : EPB * NF * ND * NU;
  constexpr int SZ_S = (MODE == MODE_TRACE_GRAD && LES) ? EPB * NF * ND * NU : 0;
  __shared__ R smem[SZ_I + SZ_D + SZ_L + SZ_U + SZ_Q + SZ_G + SZ_S];
  R* s_I = smem;          // [NFP][NU]
  R* s_D = s_I + SZ_I;    // [ND][NU][NU]: D[c][j][m]
  R* s_L = s_D + SZ_D;    // [NU][NFP]
  R* s_u = s_L + SZ_L;    // [EPB][NF][NU]
  R* s_q = s_u + SZ_U;    // [EPB][NFP][NF]  LDG delta, then interface jumps
  R* s_g = s_q + SZ_Q;    // [EPB][NF*ND][NU] gradient (TRACE_GRAD) / transformed flux
  R* s_s = s_g + SZ_G;    // [EPB][NF*ND][NU] transformed SGS flux (TRACE_GRAD, LES)

  const R gam = R(a.ph.gamma);
  RefState<R> ref;
  ref.rho = R(a.ref.rho);
  ref.E = R(a.ref.E);
  ref.p = R(a.ref.p);
  ref.H = R(a.ref.H);
  auto ref_of = [&](int k) -> double { return k == 0 ? a.ref.rho : (k == NF - 1 ? a.ref.E : 0.0); };
  auto full = [&](const R* w, R* out) {
    if constexpr (MIX) {
      add_ref<ND, R>(w, ref, out);
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) out[k] = w[k];
    }
  };
  auto ghost = [&](const R* wo, R* wr) {
    R uo_full[NF];
    full(wo, uo_full);
    wall_ghost<ND, R>(uo_full, gam, R(a.ph.R_ref), R(a.ph.T_wall), wr);
    if constexpr (MIX) {
      wr[0] -= ref.rho;
      wr[NF - 1] -= ref.E;
    }
  };

  const int N = a.N;
  const long NN = (long)N;
  const int tid = threadIdx.x;
  const int el = tid / NU;
  const int p = tid - el * NU;
  const int e = blockIdx.x * EPB + el;
  const bool active = (tid < EPB * NU) && (e < N);
  auto vol_idx = [&](long pe, int ee) -> long { return AFF ? (long)ee : pe; };
  auto geo_idx = [&](int fpt, int ee) -> long { return AFF ? (long)(fpt / NFPF) + NFACE * (long)ee
                                                             : (long)fpt + NFP * (long)ee; };
  constexpr long GEO_PER_ELE = AFF ? NFACE : NFP;

  // ---- operators and state into LDS ----
  for (int i = tid; i < SZ_I; i += BLOCK) s_I[i] = a.opI[i];
  for (int i = tid; i < SZ_D; i += BLOCK) s_D[i] = a.opD[i];
  for (int i = tid; i < SZ_L; i += BLOCK) s_L[i] = a.opL[i];
  if (active) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const double v = a.u[p + NU * (e + NN * k)];
      s_u[(el * NF + k) * NU + p] = MIX ? R(v - ref_of(k)) : R(v);
    }
  }
  __syncthreads();

  // u at flux point f of local element tl
  auto interp_u = [&](int tl, int f, R* out) {
    const R* row = s_I + f * NU;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      R v = 0;
      for (int j = 0; j < NU; j++) v += row[j] * s_u[(tl * NF + k) * NU + j];
      out[k] = v;
    }
  };

  if constexpr (MODE == MODE_TRACE_U) {
#pragma unroll 1
    for (int task = tid; task < EPB * NFP; task += BLOCK) {
      const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
      if (te >= N) continue;
      R uo[NF];
      interp_u(tl, fp, uo);
#pragma unroll
      for (int k = 0; k < NF; k++) a.uf_out[(fp + NFP * (long)te) * NF + k] = uo[k];
    }
    return;
  } else {
    // ---- LDG common solution -> delta = u_c - u_own at each flux point ----
    if constexpr (VISC) {
#pragma unroll 1
      for (int task = tid; task < EPB * NFP; task += BLOCK) {
        const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
        if (te >= N) continue;
        int nb = a.nbr[fp + NFP * te];
        const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
        nb &= 0x7fffffff;
        R uo[NF];
        interp_u(tl, fp, uo);
        if (nb & BDY_FLAG) {
          R ur[NF];
          ghost(uo, ur);
#pragma unroll
          for (int k = 0; k < NF; k++) s_q[(tl * NFP + fp) * NF + k] = R(0.5) * (ur[k] + uo[k]) - uo[k];
        } else {
#pragma unroll
          for (int k = 0; k < NF; k++) {
            const R un = a.uf_in[(long)nb * NF + k];
            const R uc = R(0.5) * (uo[k] + un) - pen * (uo[k] - un);
            s_q[(tl * NFP + fp) * NF + k] = uc - uo[k];
          }
        }
      }
      __syncthreads();
    }

    // ---- per solution point: corrected physical gradient, flux, transformed flux ----
    if (active) {
      const long pe = p + NU * (long)e;
      const long vi = vol_idx(pe, e);
      constexpr long VOL = AFF ? 1 : NU;
      R Jg[ND][ND];
#pragma unroll
      for (int l = 0; l < ND; l++)
#pragma unroll
        for (int m = 0; m < ND; m++) Jg[l][m] = a.jg[(l * ND + m) * NN * VOL + vi];
      R u_own[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) u_own[k] = s_u[(el * NF + k) * NU + p];
      ViscT<ND, R> vt;
      if constexpr (VISC) {
        const R inv_detj = R(rcp(a.detj[vi]));
        // reference gradient D u plus the LDG correction (L o tnorm) delta
        // (src/eles.cpp:1667-1870)
        R gr[NF][ND];
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
          for (int c = 0; c < ND; c++) {
            R r = 0;
            const R* drow = s_D + (c * NU + p) * NU;
            for (int m = 0; m < NU; m++) r += drow[m] * s_u[(el * NF + k) * NU + m];
            R q = 0;
            const R* lrow = s_L + p * NFP;
            for (int f = 0; f < NFP; f++) q += lrow[f] * tet_tnorm<R>(f / NFPF, c) * s_q[(el * NFP + f) * NF + k];
            gr[k][c] = r + q;
          }
        R Ji[ND][ND];
#pragma unroll
        for (int d = 0; d < ND; d++)
#pragma unroll
          for (int m = 0; m < ND; m++) Ji[d][m] = Jg[d][m] * inv_detj;
        auto gcol = [&](int m, R* gc) {
#pragma unroll
          for (int k = 0; k < NF; k++) {
            R v = 0;
#pragma unroll
            for (int d = 0; d < ND; d++) v += gr[k][d] * Ji[d][m];
            gc[k] = v;
          }
        };
        R sgM[ND][ND], sgE[ND];
        if constexpr (LES != 0) {
          R gf[NF][ND];
#pragma unroll
          for (int m = 0; m < ND; m++) {
            R gc[NF];
            gcol(m, gc);
#pragma unroll
            for (int k = 0; k < NF; k++) gf[k][m] = gc[k];
          }
          R uu[NF];
          full(u_own, uu);
          // calc_ele_vol for tets: detJ * 8/6
          sgs_flux<ND, LES, R>(uu, gf, R(a.detj[vi] * 8. / 6.), R(a.filter_ratio), P, gam, sgM, sgE);
        }
        if constexpr (MODE == MODE_TRACE_GRAD) {
#pragma unroll
          for (int m = 0; m < ND; m++) {
            R gc[NF];
            gcol(m, gc);
#pragma unroll
            for (int k = 0; k < NF; k++) s_g[(el * NF * ND + k * ND + m) * NU + p] = gc[k];
          }
          if constexpr (LES != 0) {
#pragma unroll
            for (int l = 0; l < ND; l++) {
              s_s[(el * NF * ND + 0 * ND + l) * NU + p] = R(0);
#pragma unroll
              for (int i = 0; i < ND; i++) {
                R v = 0;
#pragma unroll
                for (int m = 0; m < ND; m++) v += Jg[l][m] * sgM[i][m];
                s_s[(el * NF * ND + (1 + i) * ND + l) * NU + p] = v;
              }
              R v = 0;
#pragma unroll
              for (int m = 0; m < ND; m++) v += Jg[l][m] * sgE[m];
              s_s[(el * NF * ND + (NF - 1) * ND + l) * NU + p] = v;
            }
          }
        } else {
          R uu[NF];
          full(u_own, uu);
          visc_tensor_cols<ND, SUTH, R>(uu, gcol, a.ph, vt);
          if constexpr (LES != 0) {
#pragma unroll
            for (int i = 0; i < ND; i++)
#pragma unroll
              for (int j = 0; j < ND; j++) vt.tau[i][j] -= sgM[i][j];
#pragma unroll
            for (int j = 0; j < ND; j++) vt.qe[j] += sgE[j];
          }
        }
      }
      if constexpr (MODE != MODE_TRACE_GRAD) {
        // transformed flux tdisf(k,l) = sum_m JGinv(l,m) F(k,m)
#pragma unroll
        for (int l = 0; l < ND; l++) {
          R fi[NF];
          if constexpr (MIX) inv_flux_dot_dev<ND, R>(u_own, Jg[l], gam, ref, fi);
          else inv_flux_dot<ND, R>(u_own, Jg[l], gam, fi);
          if constexpr (VISC) {
            R fv[NF];
            visc_flux_dot<ND, R>(vt, Jg[l], fv);
#pragma unroll
            for (int k = 0; k < NF; k++) fi[k] += fv[k];
          }
#pragma unroll
          for (int k = 0; k < NF; k++) s_g[(el * NF * ND + k * ND + l) * NU + p] = fi[k];
        }
      }
    }
    __syncthreads();

    if constexpr (MODE == MODE_TRACE_GRAD) {
      // ---- viscous normal flux F_v(u_f, grad u_f).n at the flux points ----
#pragma unroll 1
      for (int task = tid; task < EPB * NFP; task += BLOCK) {
        const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
        if (te >= N) continue;
        const long fe = fp + NFP * (long)te;
        const R* row = s_I + fp * NU;
        R uo[NF];
        interp_u(tl, fp, uo);
        const long gi = geo_idx(fp, te);
        R nrm[ND];
#pragma unroll
        for (int l = 0; l < ND; l++) nrm[l] = R(a.fgeo[l * NN * GEO_PER_ELE + gi]);
        R* gout = a.gf_out;
        ViscT<ND, R> t;
        R uu[NF];
        full(uo, uu);
        visc_tensor_cols<ND, SUTH, R>(uu, [&](int c, R* gc) {
#pragma unroll
          for (int k = 0; k < NF; k++) {
            R w = 0;
            for (int j = 0; j < NU; j++) w += row[j] * s_g[(tl * NF * ND + k * ND + c) * NU + j];
            gc[k] = w;
            if (gout) gout[fe + NFP * NN * (k + NF * c)] = w;
          }
        }, a.ph, t);
        R fa[NF];
        visc_flux_dot<ND, R>(t, nrm, fa);
        if constexpr (LES != 0) {
#pragma unroll
          for (int k = 1; k < NF; k++) {
            R acc = 0;
#pragma unroll
            for (int l = 0; l < ND; l++) {
              R w = 0;
              for (int j = 0; j < NU; j++) w += row[j] * s_s[(tl * NF * ND + k * ND + l) * NU + j];
              acc += nrm[l] * w;
            }
            fa[k] += acc;
          }
        }
#pragma unroll
        for (int k = 0; k < NF; k++) a.fv_out[fe * NF + k] = fa[k];
      }
      return;
    } else {
      // ---- interface fluxes and jumps at the flux points ----
#pragma unroll 1
      for (int task = tid; task < EPB * NFP; task += BLOCK) {
        const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
        if (te >= N) continue;
        const int face = fp / NFPF;
        const long fe = fp + NFP * (long)te;
        int nb = a.nbr[fe];
        const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
        nb &= 0x7fffffff;
        const long gi = geo_idx(fp, te);
        R nrm[ND];
#pragma unroll
        for (int l = 0; l < ND; l++) nrm[l] = R(a.fgeo[l * NN * GEO_PER_ELE + gi]);
        const R tdA = R(a.fgeo[ND * NN * GEO_PER_ELE + gi]);
        R uo[NF], un[NF], fn[NF];
        interp_u(tl, fp, uo);
        const bool wall = (nb & BDY_FLAG) != 0;
        if (wall) {
          ghost(uo, un);
          R fo[NF], fg[NF];
          if constexpr (MIX) {
            inv_flux_dot_dev<ND, R>(uo, nrm, gam, ref, fo);
            inv_flux_dot_dev<ND, R>(un, nrm, gam, ref, fg);
          } else {
            inv_flux_dot<ND, R>(uo, nrm, gam, fo);
            inv_flux_dot<ND, R>(un, nrm, gam, fg);
          }
#pragma unroll
          for (int k = 0; k < NF; k++) fn[k] = R(0.5) * (fo[k] + fg[k]);
        } else {
#pragma unroll
          for (int k = 0; k < NF; k++) un[k] = a.uf_in[(long)nb * NF + k];
          if constexpr (MIX) rusanov_dev<ND, R>(uo, un, nrm, gam, ref, fn);
          else rusanov<ND, R>(uo, un, nrm, gam, fn);
        }
        R jump[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) jump[k] = fn[k] * tdA;
        if constexpr (VISC) {
          R nn = 0;
#pragma unroll
          for (int l = 0; l < ND; l++) nn += nrm[l] * nrm[l];
          const R tau = R(a.ph.tau);
          if (wall) {
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const R fa = a.fv_in[fe * NF + k];
              jump[k] += (fa + tau * nn * (uo[k] - un[k])) * tdA;
            }
          } else {
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const R fa = a.fv_in[fe * NF + k];
              const R fb = -a.fv_in[(long)nb * NF + k];
              jump[k] += ldg_normal<R>(fa, fb, uo[k], un[k], nn, pen, tau) * tdA;
            }
          }
        }
        // minus the normal transformed discontinuous flux (opp_1 = I o tnorm)
        const R* row = s_I + fp * NU;
#pragma unroll
        for (int k = 0; k < NF; k++) {
          R v = 0;
#pragma unroll
          for (int l = 0; l < ND; l++) {
            const R tn = tet_tnorm<R>(face, l);
            if (tn == R(0)) continue;
            R w = 0;
            for (int j = 0; j < NU; j++) w += row[j] * s_g[(tl * NF * ND + k * ND + l) * NU + j];
            v += tn * w;
          }
          s_q[(tl * NFP + fp) * NF + k] = jump[k] - v;
        }
      }
      __syncthreads();

      // ---- divergence plus correction lift, then output ----
      if (active) {
        R div[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) {
          R v = 0;
#pragma unroll
          for (int l = 0; l < ND; l++) {
            const R* drow = s_D + (l * NU + p) * NU;
            for (int m = 0; m < NU; m++) v += drow[m] * s_g[(el * NF * ND + k * ND + l) * NU + m];
          }
          const R* lrow = s_L + p * NFP;
          R c = 0;
          for (int f = 0; f < NFP; f++) c += lrow[f] * s_q[(el * NFP + f) * NF + k];
          div[k] = v + c;
        }
        const long pe = p + NU * (long)e;
        if (MODE == MODE_RESID || a.div) {
#pragma unroll
          for (int k = 0; k < NF; k++) a.div[pe + NU * NN * k] = double(div[k]);
        }
        if constexpr (MODE == MODE_ADVANCE) {
          const double idetj = rcp(a.detj[vol_idx(pe, e)]);
          if (a.adv_type == 3) {
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const long idx = pe + NU * NN * k;
              const double rhs = -double(div[k]) * idetj + a.ph.const_src;
              const double r = a.rk_a * a.res[idx] + a.dt * rhs;
              a.res[idx] = r;
              if constexpr (MIX) {
                const double unew = a.u[idx] + a.rk_b * r;
                a.u[idx] = unew;
                s_u[(el * NF + k) * NU + p] = R(unew - ref_of(k));
              } else {
                const double unew = s_u[(el * NF + k) * NU + p] + a.rk_b * r;
                a.u[idx] = unew;
                s_u[(el * NF + k) * NU + p] = unew;
              }
            }
          } else {
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const long idx = pe + NU * NN * k;
              if constexpr (MIX) {
                const double unew = a.u[idx] - a.dt * (double(div[k]) * idetj - a.ph.const_src);
                a.u[idx] = unew;
                s_u[(el * NF + k) * NU + p] = R(unew - ref_of(k));
              } else {
                const double unew = s_u[(el * NF + k) * NU + p] - a.dt * (div[k] * idetj - a.ph.const_src);
                a.u[idx] = unew;
                s_u[(el * NF + k) * NU + p] = unew;
              }
            }
          }
        }
      }
      if constexpr (MODE == MODE_ADVANCE) {
        __syncthreads();
#pragma unroll 1
        for (int task = tid; task < EPB * NFP; task += BLOCK) {
          const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
          if (te >= N) continue;
          R uo[NF];
          interp_u(tl, fp, uo);
#pragma unroll
          for (int k = 0; k < NF; k++) a.uf_out[(fp + NFP * (long)te) * NF + k] = uo[k];
        }
      }
    }
  }
}

}  // namespace hifr
