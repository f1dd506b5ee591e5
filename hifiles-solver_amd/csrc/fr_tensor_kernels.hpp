// Element-block kernels of the FR residual for tensor-product elements (quads,
// hexes) on gfx950.
//
// One thread per solution point, EPB elements per workgroup.  The dense element
// operators of the reference (opp_0..opp_6, src/eles.cpp:3620-4052) are applied in
// sum-factorised form from their 1-D factors: every flux point lies on one line of
// solution points, so interpolation to a face, the reference-space derivative and
// the VCJH correction lift are length-(P+1) sums along that line, read from LDS.
//
// A residual needs data from the neighbours twice (the state traces, and for
// Navier-Stokes the viscous flux traces), so one RK stage is split at those points:
//   TRACE_U    u -> u at the element's flux points                  (first stage only)
//   TRACE_GRAD u, neighbour u_f -> corrected physical grad u -> viscous normal flux
//              F_v(u_f, grad u_f).n at every flux point (5 values instead of the
//              15-component gradient trace: the LDG flux only needs normal
//              components, and the neighbour's value is the exact negation)
//   RESID      u, neighbour u_f, viscous normal fluxes -> div of the continuous flux
//   ADVANCE    RESID + RK update (src/eles.cpp:1091-1188) + traces of the new u,
//              so the next stage starts from TRACE_GRAD.
// Each kernel recomputes what it needs element-locally instead of storing
// intermediates (u is read twice per stage, every trace is written once and read
// once by the neighbour).
//
// Layouts (first index fastest): u, res, div (Nu,N,nf); u_f and the viscous normal
// flux traces (Nfp,N,nf); optional gradient traces (Nfp,N,nf,d) = the reference's
// grad_disu_fpts.  Metrics, per point: detJ (Nu,N), JGinv component-major
// (d*d, Nu*N), face geometry (d+1, Nfp*N); per element when every element is
// affine (AFF): detJ (N), JGinv (d*d, N), face geometry (d+1, 2d*N).
//
// MIX selects the mixed-precision variant: arithmetic, LDS, JGinv and the traces
// in FP32 on the deviation u - u_ref (see fr_physics.hpp); the state u, the RK
// register, detJ, the face normals / tdA and the RK update stay in FP64.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fr_physics.hpp"
#include "sgs_mlp.hpp"

namespace hifr {

constexpr int MAXN1 = 8;

// 1-D factors of the element operators (Lagrange basis on the solution points).
template <class R>
struct Ops1DT {
  R lL[MAXN1], lR[MAXN1];  // l_m(-1), l_m(+1): interpolation to the faces
  R D[MAXN1][MAXN1];       // D[a][m] = l'_m(xi_a): differentiation
  R gL[MAXN1], gR[MAXN1];  // g_L'(xi_a), g_R'(xi_a): VCJH correction derivatives
};
using Ops1D = Ops1DT<double>;

// The same factors for n points per line, as a kernel keeps them in LDS: its
// threads index them by their own position along a line, and per-lane indexed
// reads of the kernel arguments would be vector-memory loads.
template <class R, int n>
struct OpsN {
  R lL[n], lR[n];
  R D[n][n];
  R gL[n], gR[n];
};

enum { MODE_TRACE_U = 0, MODE_TRACE_GRAD = 1, MODE_RESID = 2, MODE_ADVANCE = 3 };

// nbr entry of a flux point on a physical boundary (isothermal wall)
constexpr int BDY_FLAG = 1 << 30;

// R: arithmetic / trace type (double, or float on the mixed path)
template <class R>
struct StageArgsT {
  int N;
  double* u;             // state (read; written by ADVANCE)
  double* res;           // RK register (ADVANCE)
  double* div;           // div_tconf_upts output (RESID; ADVANCE if non-null)
  // Traces (state u_f, viscous normal flux F_v.n) are stored field-interleaved:
  // value of field k at global flux point f = fpt + Nfp * element is at [f * nf + k],
  // so a face's points and fields form one contiguous run for the neighbour's gather.
  const R* uf_in;        // traces of the current state (deviations on the mixed path)
  R* uf_out;             // traces written (TRACE_U, ADVANCE)
  const R* fv_in;        // viscous normal flux traces (RESID/ADVANCE, viscous)
  R* fv_out;             // viscous normal flux traces written (TRACE_GRAD)
  R* gf_out;             // optional gradient traces (TRACE_GRAD, may be null)
  double* gu_out;        // optional corrected physical gradient at the solution points,
                         // (Nu,N,nf,d) = the reference's grad_disu_upts (TRACE_GRAD,
                         // quads/hexes, may be null)
  const double* detj;    // per point, or per element (AFF)
  const R* jg;           // per point, or per element (AFF)
  const int* nbr;        // neighbour trace index per element flux point; bit 31: LDG switch -;
                         // boundary points: BDY_FLAG | bctype
  const double* fgeo;    // own-side outward normal (d) and tdA, component-major; per flux
                         // point, or per face (AFF)
  double rk_a, rk_b, dt;
  int adv_type;          // 3 RK45 stage, 0 forward Euler
  Ops1DT<R> ops;
  Phys ph;
  RefState<double> ref;  // mixed path: reference state of the deviations
  double filter_ratio;   // LES filter width factor
  // dense element operators of tets (I = opp_0, D = opp_4, L = opp_3), combined and
  // laid out as 16x16x4 MFMA A-operand tiles: I, [D_c | L o tnorm_c], and
  // [D' | L] with D' = [D_0 D_1 D_2] - L [tnorm_c o I] (the normal flux folded in)
  const R* tI;
  const R* tG;
  const R* tD;
  // learned eddy viscosity (LES 3): network weights, input standardisation and
  // training range, output scale; mlp_on 0 gives the WALE fallback everywhere
  mlp::Weights mlpw;
  float mlp_mean[mlp::NIN], mlp_scale[mlp::NIN], mlp_lo[mlp::NIN], mlp_hi[mlp::NIN];
  float mlp_out_scale;
  int mlp_on;
  // learned closure: eddy viscosity at the solution points (Nu, N), written by the
  // gradient-trace kernel and read by the flux kernels of the same stage, which
  // see the same state and traces and so the same value
  R* nut;
};
using StageArgs = StageArgsT<double>;

template <int P, int ND, int EPB>
struct TP {
  static constexpr int n = P + 1;
  static constexpr int NU = ND == 3 ? n * n * n : n * n;
  static constexpr int NFPF = ND == 3 ? n * n : n;
  static constexpr int NFP = 2 * ND * NFPF;
  static constexpr int NF = ND + 2;
  static constexpr int BLOCK = ((EPB * NU + 63) / 64) * 64;
  static constexpr int stride(int dir) { return dir == 0 ? 1 : (dir == 1 ? n : n * n); }
};

// Flux point -> (direction of its line, side 0 = -1 end / 1 = +1 end, first
// solution point of the line).  Face layouts of src/eles_hexas.cpp:266-324 and
// src/eles_quads.cpp:259-300.
template <int P, int ND>
__device__ __forceinline__ void face_line(int fp, int& dir, int& side, int& base) {
  constexpr int n = P + 1;
  if constexpr (ND == 3) {
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
}

// Solution point p, direction dir -> position along its line, first point of the
// line, and the flux points at the -1 and +1 ends.
template <int P, int ND>
__device__ __forceinline__ void line_of(int p, int dir, int& pos, int& base, int& fpL, int& fpR) {
  constexpr int n = P + 1;
  const int a = p % n, b = (p / n) % n;
  if constexpr (ND == 3) {
    const int c = p / (n * n);
    if (dir == 0) { pos = a; base = p - a; fpL = 4 * n * n + (P - b) + n * c; fpR = 2 * n * n + b + n * c; }
    else if (dir == 1) { pos = b; base = p - n * b; fpL = n * n + a + n * c; fpR = 3 * n * n + (P - a) + n * c; }
    else { pos = c; base = p - n * n * c; fpL = (P - a) + n * b; fpR = 5 * n * n + a + n * b; }
  } else {
    if (dir == 0) { pos = a; base = p - a; fpL = 3 * n + P - b; fpR = n + b; }
    else { pos = b; base = p - n * b; fpL = a; fpR = 2 * n + P - a; }
  }
}

// LDS of one block (VB bytes per value).  A pass may write its result over what no
// later pass reads: TRACE_GRAD always writes the gradient over u and the LDG delta.
// RESID / ADVANCE write the transformed flux over u (ALIAS_F: the face pass then
// takes the own state trace from the trace array, and the new state goes back into
// u after the divergence pass) only where the plain layout would fit fewer than 6
// blocks in a CU's 160 KB, the most that the registers of those kernels allow;
// elsewhere the extra barriers cost more than they gain.
template <int P, int ND, int EPB, int MODE, int LES, int VB, bool SPL = false>
struct SmemLayout {
  using T = TP<P, ND, EPB>;
  static constexpr int mx(int x, int y) { return x > y ? x : y; }
  static constexpr bool FLUX = MODE == MODE_RESID || MODE == MODE_ADVANCE;
  static constexpr int U = EPB * T::NF * T::NU;
  static constexpr int Q = (MODE == MODE_TRACE_U) ? 0 : EPB * T::NFP * T::NF;
  static constexpr int G = (MODE == MODE_TRACE_GRAD) ? EPB * T::NF * ND * T::NU : 0;
  static constexpr int F = FLUX ? EPB * T::NF * ND * T::NU : 0;
  // LES: transformed SGS flux at the solution points (extrapolated to the faces)
  static constexpr int SG = (MODE == MODE_TRACE_GRAD && LES) ? EPB * T::NF * ND * T::NU : 0;
  static constexpr bool SPLIT_OK = SPL && FLUX && ND == 3 && P == 3 && EPB == 2;
  static constexpr bool ALIAS_F = FLUX && ((U + Q + F) * VB > 163840 / 6 || SPLIT_OK);
  // SPLIT (SPL: the advance kernels of hexes at P = 3 with constant viscosity
  // and no LES, FP64 and mixed, which run under an LDS-derived register budget):
  // the transformed flux goes over u and is staged two directions at a time,
  // directions 0 and 1 with the faces across them (4 of the 6 faces: one
  // flux-point task per thread), then direction 2 with the other two faces over
  // the same rows.  A block's LDS is 18 KB in FP64 (8 blocks per CU instead of 7
  // at 128 VGPRs) and 9 KB mixed (14 instead of 11 at 72 VGPRs)
  static constexpr bool SPLIT = SPLIT_OK;
  static constexpr int NS = SPLIT ? 2 : ND;  // direction slots of the flux rows
  static constexpr int FS = EPB * T::NF * NS * T::NU;
  static constexpr int Q_AT = FLUX ? 0 : U;
  static constexpr int U_AT = FLUX ? Q : 0;
  static constexpr int G_AT = 0;
  static constexpr int F_AT = ALIAS_F ? Q : Q + U;
  static constexpr int S_AT = mx(U + Q, G);
  static constexpr int TOTAL = FLUX ? (ALIAS_F ? Q + mx(U, FS) : Q + U + F)
                                    : (MODE == MODE_TRACE_GRAD ? S_AT + SG : U);
  // waves per SIMD that the LDS allows (rounded up, at most 8; SPLIT: at most 4
  // in FP64 and 7 mixed, the budgets under which its face passes do not spill)
  static constexpr int LDS_BLOCKS = 163840 / (TOTAL * VB);
  static constexpr int WAVES_LDS = mx(1, (LDS_BLOCKS * (T::BLOCK / 64) + 3) / 4 > 8 ? 8
                                             : (LDS_BLOCKS * (T::BLOCK / 64) + 3) / 4);
  static constexpr int WAVES_MAX = VB == 8 ? 4 : 7;
  static constexpr int WAVES = SPLIT && WAVES_LDS > WAVES_MAX ? WAVES_MAX : WAVES_LDS;
};

// Interpolate NF fields from LDS rows (stride NU apart) to the end of a line.
template <int n, int NF, int NU, class R>
__device__ __forceinline__ void interp_line(const R* rows, const R* lw, int base, int st, R* out) {
#pragma unroll
  for (int k = 0; k < NF; k++) {
    R v = 0;
#pragma unroll
    for (int m = 0; m < n; m++) v += lw[m] * rows[k * NU + base + m * st];
    out[k] = v;
  }
}

// VIS: 0 Euler, 1 Navier-Stokes with constant viscosity, 2 with Sutherland's law.
// AFF: metrics stored per element / per face (every element affine).
// MIX: mixed precision (FP32 arithmetic on deviations, FP64 state and update).
// LES: 0 none, 1 Smagorinsky, 2 WALE eddy viscosity (viscous kernels only), as the
// reference CPU build computes them; 3 learned eddy viscosity with the WALE model
// of the reference GPU build as fallback (hexes).
// Register budget (waves per SIMD) of a kernel variant.  The advance kernels of
// affine hexes up to P = 4 (constant viscosity, no LES; mixed: P = 3, 4), and of
// hexes with per-point metrics at P = 3, are compiled for the number of waves
// that their LDS allows, so that registers do not bind first (FP64 P = 3 affine:
// 128 VGPRs and 8 blocks per CU instead of 151 and 6; per-point: 128 VGPRs with
// an 8-byte spill); every other variant would spill under that budget and keeps
// the compiler's choice.
template <int P, int ND, int EPB, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE>
constexpr bool split_flux() {
  return MODE == MODE_ADVANCE && LES == 0 && VIS != 2 && RIE == 0;
}
template <int P, int ND, int EPB, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE>
constexpr int kernel_waves() {
  using S = SmemLayout<P, ND, EPB, MODE, LES, MIX ? 4 : 8, split_flux<P, ND, EPB, VIS, MODE, AFF, MIX, LES, RIE>()>;
  return (MODE == MODE_ADVANCE && ND == 3 && P <= 4 && (!MIX || P >= 3) && (AFF || (P == 3 && EPB == 2)) && LES == 0 && VIS != 2 && RIE == 0)
             ? S::WAVES
             : 1;
}

// RIE: inviscid interface flux, 0 Rusanov, 2 Roe (riemann_solve_type)
template <int P, int ND, int EPB, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE = 0>
__global__ void __launch_bounds__((TP<P, ND, EPB>::BLOCK),
                                  (kernel_waves<P, ND, EPB, VIS, MODE, AFF, MIX, LES, RIE>()))
fr_tensor_kernel(const StageArgsT<std::conditional_t<MIX, float, double>> a) {
  using R = std::conditional_t<MIX, float, double>;
  constexpr bool VISC = VIS != 0;
  constexpr bool SUTH = VIS == 2;
  static_assert(LES == 0 || VISC, "LES needs the viscous kernels");
  static_assert(RIE == 0 || (RIE == 2 && !MIX), "Roe flux: FP64 kernels only");
  using T = TP<P, ND, EPB>;
  using S = SmemLayout<P, ND, EPB, MODE, LES, (int)sizeof(R), split_flux<P, ND, EPB, VIS, MODE, AFF, MIX, LES, RIE>()>;
  constexpr int n = T::n, NU = T::NU, NFP = T::NFP, NF = T::NF, BLOCK = T::BLOCK;
  constexpr int NFPF = T::NFPF, NFACE = 2 * ND;
  __shared__ R smem[S::TOTAL];
  R* s_u = smem + S::U_AT;   // [EPB][NF][NU]  u, or u - u_ref (MIX)
  R* s_q = smem + S::Q_AT;   // [EPB][NFP][NF]   LDG delta, then interface jumps
  R* s_g = smem + S::G_AT;   // [EPB][NF*ND][NU] physical gradient (TRACE_GRAD, over s_u and s_q)
  R* s_f = smem + S::F_AT;   // [EPB][NF*ND][NU] transformed flux (RESID/ADVANCE, over s_u)
  R* s_s = smem + S::S_AT;   // [EPB][NF*ND][NU] transformed SGS flux (TRACE_GRAD, LES)
  // row of direction d of the transformed flux: NS = ND slots, or with SPLIT two
  // (directions 0 and 1, then direction 2 in slot 0)
  constexpr int NS = S::NS;
  constexpr auto fslot = [](int d) { return S::SPLIT && d == 2 ? 0 : d; };
  // learned closure (the gradient-trace kernel evaluates the network): weights,
  // standardised inputs and outputs per point
  constexpr bool NET = LES == 3 && MODE == MODE_TRACE_GRAD;
  __shared__ std::conditional_t<NET, mlp::Lds, char> s_mlp;
  __shared__ float s_x[NET ? BLOCK * mlp::NIN : 1];
  __shared__ float s_ml[NET ? BLOCK : 1];
  const R gam = R(a.ph.gamma);
  RefState<R> ref;
  ref.rho = R(a.ref.rho);
  ref.E = R(a.ref.E);
  ref.p = R(a.ref.p);
  ref.H = R(a.ref.H);
  // reference value of field k (zero momentum); only used by MIX
  auto ref_of = [&](int k) -> double { return k == 0 ? a.ref.rho : (k == NF - 1 ? a.ref.E : 0.0); };
  // full state from what the LDS / traces hold
  auto full = [&](const R* w, R* out) {
    if constexpr (MIX) {
      add_ref<ND, R>(w, ref, out);
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) out[k] = w[k];
    }
  };
  // wall ghost state in the representation of the traces (deviations on MIX)
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
  // the 1-D factors: in LDS for the flux kernels of affine hexes (FP64; mixed at
  // P = 3, 4) (every use follows at least one barrier; 1-11 % faster); the other
  // kernels read them from the arguments, where the LDS copy would cost the
  // gradient-trace kernels a wave per SIMD in registers and others a block per CU
  // in LDS
  constexpr bool OPS_LDS =
      (!MIX || P == 3 || P == 4) && ND == 3 && AFF && LES == 0 && VIS != 2 && RIE == 0 &&
      (MODE == MODE_RESID || MODE == MODE_ADVANCE);
  __shared__ std::conditional_t<OPS_LDS, OpsN<R, n>, char> s_ops;
  if constexpr (OPS_LDS) {
    constexpr int CNT = 4 * n + n * n;
    static_assert(CNT <= BLOCK, "one value per thread");
    if (tid < CNT) {
      R v;
      if (tid < n) v = a.ops.lL[tid];
      else if (tid < 2 * n) v = a.ops.lR[tid - n];
      else if (tid < 2 * n + n * n) v = a.ops.D[(tid - 2 * n) / n][(tid - 2 * n) % n];
      else if (tid < 3 * n + n * n) v = a.ops.gL[tid - 2 * n - n * n];
      else v = a.ops.gR[tid - 3 * n - n * n];
      reinterpret_cast<R*>(&s_ops)[tid] = v;
    }
  }
  using OpsT = std::conditional_t<OPS_LDS, OpsN<R, n>, Ops1DT<R>>;
  const OpsT* ops_p;
  if constexpr (OPS_LDS) ops_p = &s_ops;
  else ops_p = &a.ops;
  const OpsT& ops = *ops_p;
  // metric index of (point p of element e) and of (flux point fp of element te)
  auto vol_idx = [&](long pe, int ee) -> long { return AFF ? (long)ee : pe; };
  auto geo_idx = [&](int fpt, int ee) -> long { return AFF ? (long)(fpt / NFPF) + NFACE * (long)ee
                                                             : (long)fpt + NFP * (long)ee; };
  constexpr long GEO_PER_ELE = AFF ? NFACE : NFP;

  // ---- this thread's flux-point tasks (task = tid + t * BLOCK): in FP64 the
  //      neighbour index is fetched before u, so that its latency overlaps the u
  //      load and the face passes start with the neighbour-trace load (the FP32
  //      kernels would lose a wave per SIMD to the extra registers) ----
  constexpr int MT = (EPB * NFP + BLOCK - 1) / BLOCK;
  constexpr bool PREF = !MIX && MODE != MODE_TRACE_U;
  [[maybe_unused]] int pnb[MT];
  // ... and, in the flux kernels, each task's flux point, local element and line
  // (face_line) are decoded once for both face passes:
  // fp | tl << 9 | dir << 15 | side << 17 | base << 18
  constexpr bool FTC = PREF && MODE != MODE_TRACE_GRAD;
  [[maybe_unused]] int pft[MT];
  static_assert(NFP < 512 && EPB * NFP + BLOCK < 64 * NFP && NU < (1 << 13), "task fields fit their bits");
  // task t of this thread: tid + t * BLOCK; with SPLIT, t = 0 is one of the flux
  // points of faces 1-4 (directions 0 and 1: flux points NFPF .. 5 NFPF - 1) and
  // t = 1 one of faces 0 and 5 (direction 2), for the first 2 NFPF EPB threads
  [[maybe_unused]] auto task_of = [&](const int t) -> int {
    if constexpr (S::SPLIT) {
      static_assert(MT == 2 && 4 * NFPF == NU && BLOCK == EPB * NU, "SPLIT task map");
      if (t == 0) return (tid / NU) * NFP + NFPF + tid % NU;
      if (tid >= 2 * NFPF * EPB) return EPB * NFP;
      const int j = tid % (2 * NFPF);
      return (tid / (2 * NFPF)) * NFP + (j < NFPF ? j : 4 * NFPF + j);
    } else {
      return tid + t * BLOCK;
    }
  };
  if constexpr (PREF) {
#pragma unroll
    for (int t = 0; t < MT; t++) {
      const int task = task_of(t), tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
      pnb[t] = (task < EPB * NFP && te < N) ? a.nbr[fp + NFP * te] : BDY_FLAG;
      if constexpr (FTC) {
        int dir, side, base;
        face_line<P, ND>(fp, dir, side, base);
        pft[t] = fp | (tl << 9) | (dir << 15) | (side << 17) | (base << 18);
      }
    }
  }
  struct FaceTask {
    int tl, fp, dir, side, base;
  };
  [[maybe_unused]] auto face_task = [&](const int t, const int task) {
    FaceTask f;
    if constexpr (FTC) {
      const int k = pft[t];
      f.fp = k & 511;
      f.tl = (k >> 9) & 63;
      f.dir = (k >> 15) & 3;
      f.side = (k >> 17) & 1;
      f.base = k >> 18;
    } else {
      f.tl = task / NFP;
      f.fp = task - f.tl * NFP;
      face_line<P, ND>(f.fp, f.dir, f.side, f.base);
    }
    return f;
  };
  // body(t, task) over this thread's flux-point tasks: unrolled when pnb holds the
  // prefetched indices (registers need a constant t), a plain loop otherwise
  [[maybe_unused]] auto each_task = [&](auto&& body) {
    if constexpr (PREF) {
#pragma unroll
      for (int t = 0; t < MT; t++) body(t, task_of(t));
    } else {
#pragma unroll 1
      for (int task = tid; task < EPB * NFP; task += BLOCK) body(0, task);
    }
  };

  // ---- per-element metrics of affine meshes (JGinv, detJ), staged with the state
  //      so that the gradient pass reads them from LDS (3-13 % faster TRACE_GRAD;
  //      in the flux kernels it measured slower: a spill in FP64 at P = 3, 5 % in
  //      the mixed P = 3 kernel) ----
  constexpr bool SMET = AFF && MODE == MODE_TRACE_GRAD;
  __shared__ R s_jg[SMET ? EPB * ND * ND : 1];
  __shared__ double s_detj[SMET ? EPB : 1];
  if constexpr (SMET) {
    for (int i = tid; i < EPB * (ND * ND + 1); i += BLOCK) {
      if (i < EPB * ND * ND) {
        const int ee = blockIdx.x * EPB + i / (ND * ND);
        if (ee < N) s_jg[i] = a.jg[(i % (ND * ND)) * NN + ee];
      } else {
        const int ee = blockIdx.x * EPB + (i - EPB * ND * ND);
        if (ee < N) s_detj[i - EPB * ND * ND] = a.detj[ee];
      }
    }
  }
  // metric of point p of local element el (vi: its index in the metric arrays)
  [[maybe_unused]] auto jg_at = [&](int c, long vi) -> R {
    if constexpr (SMET) return s_jg[el * ND * ND + c];
    else return a.jg[c * NN * (AFF ? 1 : NU) + vi];
  };
  [[maybe_unused]] auto detj_at = [&](long vi) -> double {
    if constexpr (SMET) return s_detj[el];
    else return a.detj[vi];
  };

  // ---- load u ----
  if (active) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const double v = a.u[p + NU * (e + NN * k)];
      s_u[(el * NF + k) * NU + p] = MIX ? R(v - ref_of(k)) : R(v);
    }
  }
  // the network's weights are read only when one is attached (a.mlp_on is the
  // same for every thread of the launch)
  if constexpr (NET)
    if (a.mlp_on) mlp::stage(a.mlpw, s_mlp, tid, BLOCK);
  // EARLY: the LDG delta pass takes the own state trace from the trace array, not
  // from u in LDS, so it needs no barrier after the state load: the state load
  // and the neighbour gathers are in flight together, and one barrier covers both
  // (1-4 % faster).  Not in the FP64 advance kernels up to P = 3 with the three
  // flux directions staged at once, which would spill under their register
  // budget (with two-direction staging: an 8-byte spill, about 1 % faster).
  constexpr bool EARLY = VISC && MODE != MODE_TRACE_U && LES != 3 && (S::SPLIT || !(MODE == MODE_ADVANCE && !MIX && P <= 3));
  if constexpr (!EARLY) __syncthreads();

  if constexpr (MODE == MODE_TRACE_U) {
    // ---- traces of u only (start of the first stage) ----
#pragma unroll 1
    for (int task = tid; task < EPB * NFP; task += BLOCK) {
      const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
      if (te >= N) continue;
      int dir, side, base;
      face_line<P, ND>(fp, dir, side, base);
      R uo[NF];
      interp_line<n, NF, NU>(s_u + tl * NF * NU, side ? ops.lR : ops.lL, base, T::stride(dir), uo);
#pragma unroll
      for (int k = 0; k < NF; k++) a.uf_out[(fp + NFP * (long)te) * NF + k] = uo[k];
    }
    return;
  } else {
    // ---- LDG common solution -> delta = u_c - u_own at each flux point
    //      (int_inters::calculate_common_invFlux viscous part, src/int_inters.cpp:325-346) ----
    if constexpr (VISC) {
      each_task([&](const int t, const int task) {
        if (task >= EPB * NFP) return;
        const FaceTask ft = face_task(t, task);
        const int tl = ft.tl, fp = ft.fp, te = blockIdx.x * EPB + tl, dir = ft.dir, side = ft.side, base = ft.base;
        if (te >= N) return;
        int nb;
        if constexpr (PREF) nb = pnb[t];
        else nb = a.nbr[fp + NFP * te];
        const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
        nb &= 0x7fffffff;
        R uo[NF];
        if constexpr (EARLY) {
#pragma unroll
          for (int k = 0; k < NF; k++) uo[k] = a.uf_in[(fp + NFP * (long)te) * NF + k];
        } else {
          interp_line<n, NF, NU>(s_u + tl * NF * NU, side ? ops.lR : ops.lL, base, T::stride(dir), uo);
        }
        if (nb & BDY_FLAG) {
          // wall: Dirichlet common solution u_c = (u_l + u_r)/2 (src/inters.cpp:664-668)
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
      });
      __syncthreads();
    }

    // reference-space corrected gradient of point p of local element el
    // (reference gradient + LDG correction, src/eles.cpp:1667-1870)
    [[maybe_unused]] auto point_grad = [&](R (*gr)[ND]) {
      int pos[ND], lb[ND], fL[ND], fR[ND];
#pragma unroll
      for (int d = 0; d < ND; d++) line_of<P, ND>(p, d, pos[d], lb[d], fL[d], fR[d]);
#pragma unroll
      for (int k = 0; k < NF; k++)
#pragma unroll
        for (int d = 0; d < ND; d++) {
          const int st = T::stride(d);
          R r = 0;
#pragma unroll
          for (int q = 0; q < n; q++) r += ops.D[pos[d]][q] * s_u[(el * NF + k) * NU + lb[d] + q * st];
          r += ops.gL[pos[d]] * s_q[(el * NFP + fL[d]) * NF + k] + ops.gR[pos[d]] * s_q[(el * NFP + fR[d]) * NF + k];
          gr[k][d] = r;
        }
    };

    // ---- learned closure: eddy viscosity of every point before the flux pass.
    //      Inputs x = (g_ij / |g|, log(delta^2 |g| / nu)) of the velocity gradient
    //      g; the network runs on the block's points as 16-point MFMA tiles;
    //      outside the training range (or without a network) the WALE value.  The
    //      gradient-trace kernel evaluates it and stores it per point; the flux
    //      kernels of the stage read it back ----
    [[maybe_unused]] R nu_t_pt = R(0);
    [[maybe_unused]] SgsPoint<ND, R> sp;
    if constexpr (LES == 3 && VISC && !NET) {
      if (active) {
        const long vi = vol_idx(p + NU * (long)e, e);
        R gr[NF][ND];
        point_grad(gr);
        const R inv_detj = R(rcp(detj_at(vi)));
        R gf[NF][ND];
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
          for (int m = 0; m < ND; m++) {
            R v = 0;
#pragma unroll
            for (int d = 0; d < ND; d++) v += gr[k][d] * (jg_at(d * ND + m, vi) * inv_detj);
            gf[k][m] = v;
          }
        R uo[NF], uu[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) uo[k] = s_u[(el * NF + k) * NU + p];
        full(uo, uu);
        sgs_point<ND, R>(uu, gf, sp);
        nu_t_pt = a.nut[p + NU * (long)e];
      }
    }
    if constexpr (NET && VISC) {
      bool use_net = false;
      R scale_pt = R(0), nu_w = R(0);
      if (active) {
        const long vi = vol_idx(p + NU * (long)e, e);
        R gr[NF][ND];
        point_grad(gr);
        const R inv_detj = R(rcp(detj_at(vi)));
        R gf[NF][ND];
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
          for (int m = 0; m < ND; m++) {
            R v = 0;
#pragma unroll
            for (int d = 0; d < ND; d++) v += gr[k][d] * (jg_at(d * ND + m, vi) * inv_detj);
            gf[k][m] = v;
          }
        R uo[NF], uu[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) uo[k] = s_u[(el * NF + k) * NU + p];
        full(uo, uu);
        sgs_point<ND, R>(uu, gf, sp);
        const R vref = ND == 3 ? R(8) : R(4);
        const R vol = R(detj_at(vi)) * vref;
        const R delta = R(a.filter_ratio) * (ND == 3 ? cbrt(vol) : sqrt(vol)) * R(1.0 / (P + 1));
        R G2 = 0;
#pragma unroll
        for (int i = 0; i < ND; i++)
#pragma unroll
          for (int j = 0; j < ND; j++) G2 += sp.vg[i][j] * sp.vg[i][j];
        const R G = sqrt(G2);
        R x[mlp::NIN];
        use_net = a.mlp_on != 0 && G > R(0);
        if (use_net) {
#pragma unroll
          for (int i = 0; i < ND; i++)
#pragma unroll
            for (int j = 0; j < ND; j++) x[ND * i + j] = sp.vg[i][j] / G;
          x[mlp::NIN - 1] = log(delta * delta * G * sp.rho / R(a.ph.mu_inf));
#pragma unroll
          for (int f = 0; f < mlp::NIN; f++) use_net = use_net && x[f] >= R(a.mlp_lo[f]) && x[f] <= R(a.mlp_hi[f]);
        }
#pragma unroll
        for (int f = 0; f < mlp::NIN; f++)
          s_x[tid * mlp::NIN + f] = use_net ? float((x[f] - R(a.mlp_mean[f])) / R(a.mlp_scale[f])) : 0.0f;
        scale_pt = delta * delta * G;
        nu_w = wale_nu<ND, R>(sp, delta);
      } else {
#pragma unroll
        for (int f = 0; f < mlp::NIN; f++) s_x[tid * mlp::NIN + f] = 0.0f;
      }
      __syncthreads();
      if (a.mlp_on) {
        const int lane = tid & 63, wv = tid >> 6;
#pragma unroll 1
        for (int t4 = 0; t4 < 4; t4++) {
          const int pt = wv * 64 + t4 * 16 + (lane & 15);
          mlp::Tile tl;
          mlp::forward(s_mlp, tl, [&](int f) { return s_x[pt * mlp::NIN + f]; }, lane);
          if ((lane >> 4) == 0) s_ml[pt] = mlp::softplusf(tl.z4);
        }
      }
      __syncthreads();
      if (active) {
        nu_t_pt = use_net ? R(s_ml[tid]) * R(a.mlp_out_scale) * scale_pt : nu_w;
        a.nut[p + NU * (long)e] = nu_t_pt;
      }
    }

    // ---- per solution point: corrected physical gradient, flux, transformed flux ----
    // TRACE_GRAD keeps the gradient (and the transformed SGS flux) in registers until
    // every point has read u and the delta, then writes it over them
    [[maybe_unused]] R gx[NF][ND];
    [[maybe_unused]] R sx[LES != 0 && MODE == MODE_TRACE_GRAD ? NF - 1 : 1][ND];
    // RESID / ADVANCE likewise keep the transformed flux, and FP64 ADVANCE the state
    [[maybe_unused]] R fx[ND][NF];
    [[maybe_unused]] R ukeep[NF];
    if (active) {
      const long pe = p + NU * (long)e;
      const long vi = vol_idx(pe, e);
      R Jg[ND][ND];
#pragma unroll
      for (int l = 0; l < ND; l++)
#pragma unroll
        for (int m = 0; m < ND; m++) Jg[l][m] = jg_at(l * ND + m, vi);
      R u_own[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) u_own[k] = s_u[(el * NF + k) * NU + p];
      if constexpr (MODE == MODE_ADVANCE && !MIX && S::ALIAS_F) {
#pragma unroll
        for (int k = 0; k < NF; k++) ukeep[k] = u_own[k];
      }
      ViscT<ND, R> vt;
      if constexpr (VISC) {
        const R inv_detj = R(rcp(detj_at(vi)));
        int pos[ND], lb[ND], fL[ND], fR[ND];
#pragma unroll
        for (int d = 0; d < ND; d++) line_of<P, ND>(p, d, pos[d], lb[d], fL[d], fR[d]);
        // physical gradient column by column: d/dx_m = sum_d (JGinv(d,m)/J) d/dr_d
        // (reference gradient + LDG correction, src/eles.cpp:1667-1870)
        R gr[NF][ND];
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
          for (int d = 0; d < ND; d++) {
            const int st = T::stride(d);
            R r = 0;
#pragma unroll
            for (int q = 0; q < n; q++) r += ops.D[pos[d]][q] * s_u[(el * NF + k) * NU + lb[d] + q * st];
            r += ops.gL[pos[d]] * s_q[(el * NFP + fL[d]) * NF + k] + ops.gR[pos[d]] * s_q[(el * NFP + fR[d]) * NF + k];
            gr[k][d] = r;
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
        // SGS eddy-viscosity flux (src/eles.cpp:2369-2417): physical momentum flux
        // M and energy flux E from the full corrected gradient
        R sgM[ND][ND], sgE[ND];
        if constexpr (LES == 3) {
          sgs_flux_of<ND, R>(sp, sp.rho * nu_t_pt, gam, sgM, sgE);
        } else if constexpr (LES != 0) {
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
          const R vref = ND == 3 ? R(8) : R(4);
          sgs_flux<ND, LES, R>(uu, gf, R(detj_at(vi)) * vref, R(a.filter_ratio), P, gam, sgM, sgE);
        }
        if constexpr (MODE == MODE_TRACE_GRAD) {
#pragma unroll
          for (int m = 0; m < ND; m++) {
            R gc[NF];
            gcol(m, gc);
#pragma unroll
            for (int k = 0; k < NF; k++) gx[k][m] = gc[k];
          }
          if (a.gu_out) {
            // (a uniform branch: the integral diagnostics ask for the gradient)
#pragma unroll
            for (int k = 0; k < NF; k++)
#pragma unroll
              for (int m = 0; m < ND; m++) a.gu_out[pe + (long)NU * N * (k + NF * m)] = double(gx[k][m]);
          }
          if constexpr (LES != 0) {
            // transformed SGS flux sgsf(k,l) = sum_m JGinv(l,m) F_sgs(k,m), k >= 1
#pragma unroll
            for (int l = 0; l < ND; l++) {
#pragma unroll
              for (int i = 0; i < ND; i++) {
                R v = 0;
#pragma unroll
                for (int m = 0; m < ND; m++) v += Jg[l][m] * sgM[i][m];
                sx[i][l] = v;
              }
              R v = 0;
#pragma unroll
              for (int m = 0; m < ND; m++) v += Jg[l][m] * sgE[m];
              sx[NF - 2][l] = v;
            }
          }
        } else {
          R uu[NF];
          full(u_own, uu);
          visc_tensor_cols<ND, SUTH, R>(uu, gcol, a.ph, vt);
          if constexpr (LES != 0) {
            // F_v(1+i, j) = -tau(i, j), F_v(d+1, j) = qe(j): add the SGS flux
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
        // transformed flux tdisf(k,l) = sum_m JGinv(l,m) F(k,m)  (src/eles.cpp:1397-1405, 2437-2444)
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
          for (int k = 0; k < NF; k++) {
            if constexpr (S::ALIAS_F) fx[l][k] = fi[k];
            else s_f[(el * NF * NS + k * NS + l) * NU + p] = fi[k];
          }
        }
      }
    }
    if constexpr (S::ALIAS_F) {
      __syncthreads();
      if (active) {
#pragma unroll
        for (int l = 0; l < NS; l++)
#pragma unroll
          for (int k = 0; k < NF; k++) s_f[(el * NF * NS + k * NS + l) * NU + p] = fx[l][k];
      }
    }
    if constexpr (MODE == MODE_TRACE_GRAD) {
      __syncthreads();
      if (active) {
#pragma unroll
        for (int k = 0; k < NF; k++)
#pragma unroll
          for (int m = 0; m < ND; m++) s_g[(el * NF * ND + k * ND + m) * NU + p] = gx[k][m];
        if constexpr (LES != 0) {
#pragma unroll
          for (int l = 0; l < ND; l++) {
            s_s[(el * NF * ND + 0 * ND + l) * NU + p] = R(0);
#pragma unroll
            for (int k = 1; k < NF; k++) s_s[(el * NF * ND + k * ND + l) * NU + p] = sx[k - 1][l];
          }
        }
      }
    }
    __syncthreads();

    if constexpr (MODE == MODE_TRACE_GRAD) {
      // ---- viscous normal flux F_v(u_f, grad u_f).n at the flux points
      //      (the interface-side viscous flux of src/int_inters.cpp:392-404) ----
#pragma unroll 1
      for (int task = tid; task < EPB * NFP; task += BLOCK) {
        const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
        if (te >= N) continue;
        int dir, side, base;
        face_line<P, ND>(fp, dir, side, base);
        const int st = T::stride(dir);
        const R* lw = side ? ops.lR : ops.lL;
        const long fe = fp + NFP * (long)te;
        // own state trace, as the state pass wrote it (s_u now holds the gradient)
        R uo[NF];
#pragma unroll
        for (int k = 0; k < NF; k++) uo[k] = a.uf_in[fe * NF + k];
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
#pragma unroll
            for (int q = 0; q < n; q++) w += lw[q] * s_g[(tl * NF * ND + k * ND + c) * NU + base + q * st];
            gc[k] = w;
            if (gout) gout[fe + NFP * NN * (k + NF * c)] = w;
          }
        }, a.ph, t);
        R fa[NF];
        visc_flux_dot<ND, R>(t, nrm, fa);
        if constexpr (LES != 0) {
          // plus the extrapolated transformed SGS flux dotted with the physical
          // normal, as the reference adds it (src/int_inters.cpp:415-427)
#pragma unroll
          for (int k = 1; k < NF; k++) {
            R acc = 0;
#pragma unroll
            for (int l = 0; l < ND; l++) {
              R w = 0;
#pragma unroll
              for (int q = 0; q < n; q++) w += lw[q] * s_s[(tl * NF * ND + k * ND + l) * NU + base + q * st];
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
      // ---- interface fluxes and jumps at the flux points
      //      (common_invFlux / common_viscFlux, extrapolate_totalFlux, daxpy) ----
      auto face_body = [&](const int t, const int task) {
        if (task >= EPB * NFP) return;
        const FaceTask ft = face_task(t, task);
        const int tl = ft.tl, fp = ft.fp, te = blockIdx.x * EPB + tl, dir = ft.dir, side = ft.side, base = ft.base;
        if (te >= N) return;
        const int st = T::stride(dir);
        const R* lw = side ? ops.lR : ops.lL;
        const long fe = fp + NFP * (long)te;
        int nb;
        if constexpr (PREF) nb = pnb[t];
        else nb = a.nbr[fe];
        const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
        nb &= 0x7fffffff;
        const long gi = geo_idx(fp, te);
        R nrm[ND];
#pragma unroll
        for (int l = 0; l < ND; l++) nrm[l] = R(a.fgeo[l * NN * GEO_PER_ELE + gi]);
        const R tdA = R(a.fgeo[ND * NN * GEO_PER_ELE + gi]);
        // own state trace: with ALIAS_F as the state pass wrote it (s_u holds the flux)
        R uo[NF], un[NF], fn[NF];
        if constexpr (S::ALIAS_F) {
#pragma unroll
          for (int k = 0; k < NF; k++) uo[k] = a.uf_in[fe * NF + k];
        } else {
          interp_line<n, NF, NU>(s_u + tl * NF * NU, lw, base, st, uo);
        }
        const bool wall = (nb & BDY_FLAG) != 0;
        if (wall) {
          ghost(uo, un);
          if constexpr (RIE == 2) {
            // Roe flux against the ghost state (src/bdy_inters.cpp:330)
            roe<ND, R>(uo, un, nrm, gam, fn);
          } else {
            // isothermal wall: central flux between the trace and the ghost state
            // (convective_flux_boundary, src/inters.cpp:368-387)
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
          }
        } else {
#pragma unroll
          for (int k = 0; k < NF; k++) un[k] = a.uf_in[(long)nb * NF + k];
          if constexpr (RIE == 2) roe_pair<ND, R>(fe, nb, uo, un, nrm, gam, fn);
          else if constexpr (MIX) rusanov_dev<ND, R>(uo, un, nrm, gam, ref, fn);
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
            // Dirichlet LDG flux f_c = f_l + tau n (u_l - u_r) (src/inters.cpp:605-619)
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const R fa = a.fv_in[fe * NF + k];
              jump[k] += (fa + tau * nn * (uo[k] - un[k])) * tdA;
            }
          } else {
#pragma unroll
            for (int k = 0; k < NF; k++) {
              // own F_v.n and the neighbour's, whose normal is the negated own normal
              const R fa = a.fv_in[fe * NF + k];
              const R fb = -a.fv_in[(long)nb * NF + k];
              jump[k] += ldg_normal<R>(fa, fb, uo[k], un[k], nn, pen, tau) * tdA;
            }
          }
        }
        // minus the normal transformed discontinuous flux (opp_1 = opp_0 o tnorm)
        const R sgn = side ? R(1) : R(-1);
#pragma unroll
        for (int k = 0; k < NF; k++) {
          R v = 0;
#pragma unroll
          for (int m = 0; m < n; m++) v += lw[m] * s_f[(tl * NF * NS + k * NS + fslot(dir)) * NU + base + m * st];
          s_q[(tl * NFP + fp) * NF + k] = jump[k] - sgn * v;
        }
      };
      // divergence of the transformed flux in directions [d0, d1) plus their
      // correction lift, into div
      [[maybe_unused]] R div[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) div[k] = 0;
      auto div_dirs = [&](const int d0, const int d1) {
#pragma unroll
        for (int d = 0; d < ND; d++) {
          if (d < d0 || d >= d1) continue;
          int pos, lb, fpL, fpR;
          line_of<P, ND>(p, d, pos, lb, fpL, fpR);
          const int st = T::stride(d);
#pragma unroll
          for (int k = 0; k < NF; k++) {
            R v = 0;
#pragma unroll
            for (int m = 0; m < n; m++) v += ops.D[pos][m] * s_f[(el * NF * NS + k * NS + fslot(d)) * NU + lb + m * st];
            div[k] += v + (ops.gR[pos] * s_q[(el * NFP + fpR) * NF + k] - ops.gL[pos] * s_q[(el * NFP + fpL) * NF + k]);
          }
        }
      };
      if constexpr (S::SPLIT) {
        // directions 0 and 1: faces 1-4, their divergence; then direction 2
        // over the same flux rows: faces 0 and 5 (tasks of the first 2 NFPF EPB
        // threads); its divergence completes div below
        face_body(0, task_of(0));
        __syncthreads();
        if (active) div_dirs(0, 2);
        __syncthreads();
        if (active) {
#pragma unroll
          for (int k = 0; k < NF; k++) s_f[(el * NF * NS + k * NS + fslot(2)) * NU + p] = fx[2][k];
        }
        __syncthreads();
        face_body(1, task_of(1));
      } else {
        each_task(face_body);
      }
      __syncthreads();

      // ---- divergence of the transformed flux (calculate_divergence) plus the
      //      correction lift (calculate_corrected_divergence), then output ----
      [[maybe_unused]] R unew_r[NF];
      // state before the update (FP64: as loaded; MIX: the FP64 array) and the new
      // state's LDS copy (in place, or after the barrier below with ALIAS_F)
      [[maybe_unused]] auto old_u = [&](int k, long idx) -> double {
        if constexpr (MIX) return a.u[idx];
        else if constexpr (S::ALIAS_F) return ukeep[k];
        else return s_u[(el * NF + k) * NU + p];
      };
      [[maybe_unused]] auto put_u = [&](int k, double unew) {
        const R v = R(unew - (MIX ? ref_of(k) : 0.0));
        if constexpr (S::ALIAS_F) unew_r[k] = v;
        else s_u[(el * NF + k) * NU + p] = v;
      };
      if (active) {
        // the RK register and detJ are requested before the divergence, so that
        // their latency overlaps it (the compiler cannot move these loads above
        // the div store by itself).  The RK register is next read two kernels
        // later: its loads and stores are non-temporal, which leaves the caches
        // to the traces and the state that the next kernel reads
        [[maybe_unused]] double resv[NF];
        [[maybe_unused]] double detjv = 0.0;
        if constexpr (MODE == MODE_ADVANCE) {
          const long pe0 = p + NU * (long)e;
          detjv = detj_at(vol_idx(pe0, e));
          if (a.adv_type == 3) {
#pragma unroll
            for (int k = 0; k < NF; k++) resv[k] = __builtin_nontemporal_load(&a.res[pe0 + NU * NN * k]);
          }
        }
        div_dirs(S::SPLIT ? 2 : 0, ND);
        const long pe = p + NU * (long)e;
        if (MODE == MODE_RESID || a.div) {
#pragma unroll
          for (int k = 0; k < NF; k++) a.div[pe + NU * NN * k] = double(div[k]);
        }
        if constexpr (MODE == MODE_ADVANCE) {
          const double idetj = rcp(detjv);
          if (a.adv_type == 3) {
            // low-storage RK45 stage (src/eles.cpp:1155-1178)
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const long idx = pe + NU * NN * k;
              const double rhs = -double(div[k]) * idetj + a.ph.const_src;
              const double r = a.rk_a * resv[k] + a.dt * rhs;
              __builtin_nontemporal_store(r, &a.res[idx]);
              const double unew = old_u(k, idx) + a.rk_b * r;
              a.u[idx] = unew;
              put_u(k, unew);
            }
          } else {
            // forward Euler (src/eles.cpp:1067-1079)
#pragma unroll
            for (int k = 0; k < NF; k++) {
              const long idx = pe + NU * NN * k;
              const double unew = old_u(k, idx) - a.dt * (double(div[k]) * idetj - a.ph.const_src);
              a.u[idx] = unew;
              put_u(k, unew);
            }
          }
        }
      }
      if constexpr (MODE == MODE_ADVANCE) {
        if constexpr (S::ALIAS_F) {
          // the new state goes over the flux once every point has read it
          __syncthreads();
          if (active) {
#pragma unroll
            for (int k = 0; k < NF; k++) s_u[(el * NF + k) * NU + p] = unew_r[k];
          }
        }
        __syncthreads();
        // ---- traces of the new state for the next stage ----
#pragma unroll 1
        for (int task = tid; task < EPB * NFP; task += BLOCK) {
          const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
          if (te >= N) continue;
          int dir, side, base;
          face_line<P, ND>(fp, dir, side, base);
          R uo[NF];
          interp_line<n, NF, NU>(s_u + tl * NF * NU, side ? ops.lR : ops.lL, base, T::stride(dir), uo);
#pragma unroll
          for (int k = 0; k < NF; k++) a.uf_out[(fp + NFP * (long)te) * NF + k] = uo[k];
        }
      }
    }
  }
}

}  // namespace hifr
