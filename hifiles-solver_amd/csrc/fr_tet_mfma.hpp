// Flux / divergence / update pass of the FR residual for tetrahedra on gfx950 with
// the dense element operators applied as MFMA tiles (RESID and ADVANCE stages).
//
// A workgroup is one wavefront holding EPB = 64 / Nu elements.  Every operator
// product of the pass is one small GEMM  C (M x NC) = A (M x K) * B (K x NC)  whose
// columns are the block's (element, field) pairs, n = 5 el + k, padded to NC = 16 NT:
//   uo   = I                      * u          traces for the LDG common solution
//   gr   = [D_c | L o tnorm_c]    * [u; delta] corrected reference gradient, row c Nu + p
//   fn   = [tnorm_c o I]          * Ft         normal transformed flux at the flux points
//   div  = [D_0 D_1 D_2 | L]      * [Ft; jump] divergence plus correction lift
//   uf   = I                      * u_new      traces of the new state
// (the reference's opp_0, opp_4 / opp_2, opp_3 / opp_5 / opp_6 products,
// src/eles.cpp:3620-4052, with tnorm the reference tet's face normals,
// src/eles_tets.cpp:589-620: face 0 (1,1,1)/sqrt(3), faces 1-3 -e_r, -e_s, -e_t).
// A is streamed from global memory in the MFMA operand order (16 x 4 tiles, 64
// values per tile, prepared once by hifr_set_operators), so the operators take no
// LDS and every load is one coalesced 512 B (FP64) access; B comes from LDS as
// [row][column] matrices.  The pointwise physics between the products (fluxes, LES,
// Rusanov and LDG interface fluxes, RK update) runs one thread per solution point
// or per flux point, as in the tensor-product kernels.
#pragma once
#include "fr_tensor_kernels.hpp"

namespace hifr {

// Tet of order P: Nu solution points, 4 faces of Nfpf flux points.  A workgroup
// of WAVES wavefronts holds EPB = 64 WAVES / Nu elements (16, 6, 3 at P = 1, 2, 3
// with one).  Two wavefronts, splitting the row tiles of every product so that an
// operator tile feeds twice the elements, measured 12-22 % slower at P = 3 (FP64):
// twice the LDS per block, barriers across waves, 3 row tiles split 2 + 1.
template <int P>
struct TetP {
  static constexpr int NU = (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NFPF = (P + 1) * (P + 2) / 2;
  static constexpr int NFP = 4 * NFPF;
  static constexpr int WAVES = 1;
  static constexpr int BLOCK = 64 * WAVES;
  static constexpr int EPB = BLOCK / NU;
};

// 16x16x4 MFMA per precision.  Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; element i of its result is C[row(l, i)][l % 16].
template <class R>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int i) { return (l >> 4) + 4 * i; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int i) { return 4 * (l >> 4) + i; }
};

// Sizes of the tiled products for order P (shared with the host, which prepares
// the operator tiles in hifr_set_operators).  G column groups: a wavefront's
// MFMA column block holds G groups of 64 / Nu elements, so that each operator
// tile loaded from L2 feeds G column tiles.
template <int P, int G = 1>
struct TetMfma {
  using T = TetP<P>;
  static constexpr int NU = T::NU, NFP = T::NFP, EPB = G * T::EPB;
  static constexpr int KU = (NU + 3) / 4 * 4;          // rows of u (zero-padded)
  static constexpr int KF = (3 * NU + 3) / 4 * 4;      // rows of the flux / gradient
  static constexpr int NT = (5 * EPB + 15) / 16, NC = 16 * NT;
  // row pitch of the LDS matrices: one more than NC, so that the per-point and
  // per-flux-point passes (one row per lane) spread over the LDS banks
  static constexpr int LD = NC + 1;
  static constexpr int MT_I = (NFP + 15) / 16, KS_I = KU / 4;              // I
  static constexpr int MT_G = (3 * NU + 15) / 16, KS_G = (KU + NFP) / 4;  // [D_c | L tnorm_c]
  static constexpr int MT_D = (NU + 15) / 16, KS_D = (KF + NFP) / 4;       // [D | L]
  static constexpr int SZ_I = MT_I * KS_I * 64, SZ_G = MT_G * KS_G * 64, SZ_D = MT_D * KS_D * 64;
};

// Combined operators of the tet products (host tile order in hifr_set_operators).
enum TetOp { TET_OP_I = 0, TET_OP_G = 1, TET_OP_D = 3 };

// Whether tile (mt, ks) of operator OP can hold a nonzero at order P.  The
// reference tet's face normals give [L o tnorm_c] (in G) zero blocks for the faces
// other than 0 and c + 1; I and [D' | L] are dense.
template <int P, int OP>
__host__ __device__ constexpr bool tet_tile_nz(int mt, int ks) {
  constexpr int NU = TetP<P>::NU, NFPF = TetP<P>::NFPF, KU = (NU + 3) / 4 * 4;
  if (OP != TET_OP_G) return true;
  for (int r = 16 * mt; r < 16 * mt + 16; r++)
    for (int col = 4 * ks; col < 4 * ks + 4; col++) {
      if (r >= 3 * NU) continue;
      if (col < KU) {
        if (col < NU) return true;
        continue;
      }
      const int face = (col - KU) / NFPF, c = r / NU;
      if (face == 0 || face - 1 == c) return true;
    }
  return false;
}

// acc[j][nt] += A (row tiles MT0 + j, j < CNT; k-steps KS0 .. KS0 + KS - 1 of
// operator OP with KSA k-steps) * B (KS * 4 rows of an LDS matrix with NC = 16 NT
// columns, pitch NC + 1).  Tiles that tet_tile_nz rules out are skipped at compile
// time (the loops unroll to straight-line code without them).
template <class R, int P, int OP, int MT0, int CNT, int H, int NT, int KSA, int KS0, int KS>
__device__ __forceinline__ void mfma_seg(const R* __restrict__ A, const R* B, int lane,
                                         typename Mfma<R>::v4 (&acc)[H][NT]) {
  constexpr int LD = 16 * NT + 1;
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    R b[NT];
#pragma unroll
    for (int nt = 0; nt < NT; nt++) b[nt] = B[(ks * 4 + (lane >> 4)) * LD + nt * 16 + (lane & 15)];
#pragma unroll
    for (int j = 0; j < CNT; j++) {
      if (!tet_tile_nz<P, OP>(MT0 + j, KS0 + ks)) continue;
      const R av = A[((MT0 + j) * KSA + KS0 + ks) * 64 + lane];
#pragma unroll
      for (int nt = 0; nt < NT; nt++) acc[j][nt] = Mfma<R>::op(av, b[nt], acc[j][nt]);
    }
  }
}

// rows < M of row tiles MT0 + j (j < CNT) into an LDS matrix (pitch 16 NT + 1)
template <class R, int MT0, int CNT, int H, int NT>
__device__ __forceinline__ void mfma_store(const typename Mfma<R>::v4 (&acc)[H][NT], int M, R* C, int lane) {
  constexpr int LD = 16 * NT + 1;
#pragma unroll
  for (int j = 0; j < CNT; j++)
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int r = (MT0 + j) * 16 + Mfma<R>::row(lane, i);
      if (r < M) {
#pragma unroll
        for (int nt = 0; nt < NT; nt++) C[r * LD + nt * 16 + (lane & 15)] = acc[j][nt][i];
      }
    }
}

// One product C (rows < M) = A * [B1 (4 KS1 rows); B2 (4 KS2 rows)] of the block.
// With two wavefronts, wave 0 computes row tiles [0, H) and wave 1 [H, MT),
// H = ceil(MT / 2): one wave-uniform branch around straight-line code.  SYNC: a
// barrier between the products and the stores (C overwrites B).
template <class R, int P, int OP, int MT, int NT, int KSA, int KS1, int KS2, bool SYNC = false>
__device__ __forceinline__ void tet_product(const R* __restrict__ A, const R* B1, const R* B2, int M, R* C,
                                            int tid) {
  using V4 = typename Mfma<R>::v4;
  constexpr int W = TetP<P>::WAVES, H = (MT + W - 1) / W;
  const int lane = tid & 63;
  const bool first = W == 1 || tid < 64;
  V4 acc[H][NT];
#pragma unroll
  for (int j = 0; j < H; j++)
#pragma unroll
    for (int nt = 0; nt < NT; nt++) acc[j][nt] = V4{0, 0, 0, 0};
  if (first) {
    mfma_seg<R, P, OP, 0, H, H, NT, KSA, 0, KS1>(A, B1, lane, acc);
    if constexpr (KS2 > 0) mfma_seg<R, P, OP, 0, H, H, NT, KSA, KS1, KS2>(A, B2, lane, acc);
  } else {
    mfma_seg<R, P, OP, H, MT - H, H, NT, KSA, 0, KS1>(A, B1, lane, acc);
    if constexpr (KS2 > 0) mfma_seg<R, P, OP, H, MT - H, H, NT, KSA, KS1, KS2>(A, B2, lane, acc);
  }
  if constexpr (SYNC) __syncthreads();
  if (first) mfma_store<R, 0, H, H, NT>(acc, M, C, lane);
  else mfma_store<R, H, MT - H, H, NT>(acc, M, C, lane);
}

// Zero rows [r0, r1) of an LDS matrix with row pitch LD.  The kernels zero only the
// padding rows that an MFMA product reads as its B operand before anything writes
// them: their A columns are zero, so they must hold finite values.  Every other
// row a product reads is written first, and columns that no element owns (a
// partial last block, the 16th column) reach only output columns nothing reads.
template <class R, int LD, int BLOCK>
__device__ __forceinline__ void zero_rows(R* base, int r0, int r1, int tid) {
  for (int i = r0 * LD + tid; i < r1 * LD; i += BLOCK) base[i] = R(0);
}

// Traces staged in LDS as [f][5 el + k] written to the trace array: the block's
// elements' traces are one contiguous run, written in order.
template <int P, class R, int G = 1>
__device__ __forceinline__ void tet_write_traces(const R* s_q, R* uf_out, int N, int tid) {
  using M = TetMfma<P, G>;
  constexpr int NFP = M::NFP, EPB = M::EPB, LD = M::LD, NF = 5;
  const int nel = (N - (int)blockIdx.x * EPB) < EPB ? (N - (int)blockIdx.x * EPB) : EPB;
  R* out = uf_out + (long)blockIdx.x * EPB * NFP * NF;
  for (int i = tid; i < nel * NFP * NF; i += TetP<P>::BLOCK) {
    const int tl = i / (NFP * NF), r = i - tl * NFP * NF, f = r / NF, k = r - f * NF;
    out[i] = s_q[f * LD + tl * NF + k];
  }
}

// First-stage traces (TRACE_U): uf = I u.
template <int P, bool MIX>
__global__ void __launch_bounds__(TetP<P>::BLOCK)
fr_tet_mfma_trace_kernel(const StageArgsT<std::conditional_t<MIX, float, double>> a) {
  using R = std::conditional_t<MIX, float, double>;
  using M = TetMfma<P>;
  constexpr int NU = M::NU, NFP = M::NFP, EPB = M::EPB, KU = M::KU, NT = M::NT, LD = M::LD, NF = 5;
  __shared__ R smem[(KU + NFP) * LD];
  R* s_u = smem;
  R* s_q = smem + KU * LD;
  const int tid = threadIdx.x, el = tid / NU, p = tid - el * NU, e = blockIdx.x * EPB + el;
  const long NN = a.N;
  zero_rows<R, LD, TetP<P>::BLOCK>(smem, NU, KU, tid);  // u padding
  if (tid < EPB * NU && e < a.N) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const double v = a.u[p + NU * (e + NN * k)];
      s_u[p * LD + el * NF + k] = MIX ? R(v - (k == 0 ? a.ref.rho : (k == NF - 1 ? a.ref.E : 0.0))) : R(v);
    }
  }
  __syncthreads();
  tet_product<R, P, TET_OP_I, M::MT_I, NT, M::KS_I, M::KS_I, 0>(
      a.tI, s_u, nullptr, NFP, s_q, tid);
  __syncthreads();
  tet_write_traces<P, R>(s_q, a.uf_out, a.N, tid);
}

// RIE: inviscid interface flux, 0 Rusanov, 2 Roe (riemann_solve_type)
// G: column groups (TetMfma); the per-point passes loop over them.
template <int P, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE = 0, int G = 1>
__global__ void __launch_bounds__(TetP<P>::BLOCK, G == 1 ? 3 : 2)
fr_tet_mfma_kernel(const StageArgsT<std::conditional_t<MIX, float, double>> a) {
  using R = std::conditional_t<MIX, float, double>;
  static_assert(RIE == 0 || (RIE == 2 && !MIX), "Roe flux: FP64 kernels only");
  constexpr int ND = 3, NF = 5, NFACE = 4;
  constexpr bool VISC = VIS != 0;
  constexpr bool SUTH = VIS == 2;
  static_assert(MODE == MODE_RESID || MODE == MODE_ADVANCE, "flux pass only");
  static_assert(LES == 0 || VISC, "LES needs the viscous kernels");
  using M = TetMfma<P, G>;
  constexpr int NU = M::NU, NFP = M::NFP, NFPF = TetP<P>::NFPF, EPB = M::EPB, E1 = TetP<P>::EPB;
  constexpr int KU = M::KU, KF = M::KF, NT = M::NT, LD = M::LD;
  // Two regions.  The first holds u and the LDG delta; the gradient product writes
  // the corrected gradient over them (a barrier separates its MFMAs from its stores),
  // the point pass the transformed flux over the gradient, and the new state goes
  // there once the divergence product has read the flux.  The second holds the
  // normal flux, the jumps, the divergence and the traces of the new state.
  constexpr int RA = (KU + NFP) > KF ? (KU + NFP) : KF;
  __shared__ R smem[(RA + NFP) * LD];
  R* s_u = smem;              // [KU][LD]   u (deviations on MIX), then the new state
  R* s_d = s_u + KU * LD;     // [NFP][LD]  uo -> LDG delta
  R* s_f = smem;              // [KF][LD]   corrected reference gradient -> transformed flux
  R* s_q = smem + RA * LD;    // [NFP][LD]  fn -> jump - fn; div (rows < Nu); traces of the new state

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
  // lane -> (element within a column group, solution point); group g holds the
  // block's elements g E1 .. g E1 + E1 - 1
  const int el0 = tid / NU;
  const int p = tid - el0 * NU;
  const bool lane_on = tid < E1 * NU;
  auto vol_idx = [&](long pe, int ee) -> long { return AFF ? (long)ee : pe; };
  auto geo_idx = [&](int fpt, int ee) -> long { return AFF ? (long)(fpt / NFPF) + NFACE * (long)ee
                                                             : (long)fpt + NFP * (long)ee; };
  constexpr long GEO_PER_ELE = AFF ? NFACE : NFP;

  // ---- zero the padding rows of u and of the flux, then u ----
  zero_rows<R, LD, TetP<P>::BLOCK>(smem, NU, KU, tid);
  zero_rows<R, LD, TetP<P>::BLOCK>(smem, 3 * NU, RA, tid);
  // FP64: the point's state stays in registers for the flux pass and the RK
  // update (the LDS that holds it now takes the gradient); the mixed kernel would
  // lose a wave per SIMD to the registers and reads it again
  constexpr bool KEEP_U = !MIX;
  [[maybe_unused]] double ukeep[G][NF];
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int el = g * E1 + el0, e = blockIdx.x * EPB + el;
#pragma unroll
    for (int k = 0; k < NF; k++) ukeep[g][k] = 1.0;
    if (lane_on && e < N) {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const double v = a.u[p + NU * (e + NN * k)];
        ukeep[g][k] = v;
        s_u[p * LD + el * NF + k] = MIX ? R(v - ref_of(k)) : R(v);
      }
    }
  }
  __syncthreads();

  // FP64 viscous kernels keep each flux-point task's neighbour index from the
  // delta pass for the face pass
  constexpr int MTK = (EPB * NFP + TetP<P>::BLOCK - 1) / TetP<P>::BLOCK;
  constexpr bool KEEP_NB = VISC && !MIX;
  [[maybe_unused]] int nbk[KEEP_NB ? MTK : 1];
  // ... and, from P = 2, the first task's own trace (the second's would spill)
  constexpr bool KEEP_UO0 = KEEP_NB && P >= 2;
  [[maybe_unused]] R uo0[KEEP_UO0 ? NF : 1];
  if constexpr (VISC) {
    // ---- the LDG delta = u_c - uo at each flux point
    //      (int_inters::calculate_common_invFlux viscous part, src/int_inters.cpp:325-346).
    //      The own trace uo = I u is the one the previous pass wrote (the same
    //      product of the same operator tiles and state, so the same values) ----
#pragma unroll KEEP_NB ? MTK : 1
    for (int t = 0; t < MTK; t++) {
      const int task = tid + TetP<P>::BLOCK * t;
      const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
      if (task >= EPB * NFP || te >= N) continue;
      int nb = a.nbr[fp + NFP * te];
      if constexpr (KEEP_NB) nbk[t] = nb;
      const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
      nb &= 0x7fffffff;
      R* q = s_d + fp * LD + tl * NF;
      R uo[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] = a.uf_in[(fp + NFP * (long)te) * NF + k];
      if constexpr (KEEP_UO0) {
        if (t == 0) {
#pragma unroll
          for (int k = 0; k < NF; k++) uo0[k] = uo[k];
        }
      }
      if (nb & BDY_FLAG) {
        // wall: Dirichlet common solution u_c = (u_l + u_r)/2 (src/inters.cpp:664-668)
        R ur[NF];
        ghost(uo, ur);
#pragma unroll
        for (int k = 0; k < NF; k++) q[k] = R(0.5) * (ur[k] + uo[k]) - uo[k];
      } else {
#pragma unroll
        for (int k = 0; k < NF; k++) {
          const R un = a.uf_in[(long)nb * NF + k];
          const R uc = R(0.5) * (uo[k] + un) - pen * (uo[k] - un);
          q[k] = uc - uo[k];
        }
      }
    }
    __syncthreads();
    // ---- corrected reference gradient: D_c u + (L o tnorm_c) delta (src/eles.cpp:1667-1870),
    //      written over u and delta ----
    tet_product<R, P, TET_OP_G, M::MT_G, NT, M::KS_G, KU / 4, NFP / 4, true>(
        a.tG, s_u, s_d, 3 * NU, s_f, tid);
    __syncthreads();
  }

  // ---- per solution point: physical gradient, flux, transformed flux (written
  //      over the gradient of the same point and element) ----
#pragma unroll 1
  for (int g = 0; g < G; g++) {
    const int el = g * E1 + el0, e = blockIdx.x * EPB + el;
    if (!(lane_on && e < N)) continue;
    const long pe = p + NU * (long)e;
    const long vi = vol_idx(pe, e);
    constexpr long VOL = AFF ? 1 : NU;
    R Jg[ND][ND];
#pragma unroll
    for (int l = 0; l < ND; l++)
#pragma unroll
      for (int m = 0; m < ND; m++) Jg[l][m] = a.jg[(l * ND + m) * NN * VOL + vi];
    // the point's state: from LDS, or (viscous: the gradient is over it) from HBM
    R u_own[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) {
      if constexpr (VISC) {
        const double v = KEEP_U ? ukeep[g][k] : a.u[pe + NU * NN * k];
        u_own[k] = MIX ? R(v - ref_of(k)) : R(v);
      } else {
        u_own[k] = s_u[p * LD + el * NF + k];
      }
    }
    ViscT<ND, R> vt;
    if constexpr (VISC) {
      const R inv_detj = R(rcp(a.detj[vi]));
      R gr[NF][ND];
#pragma unroll
      for (int k = 0; k < NF; k++)
#pragma unroll
        for (int c = 0; c < ND; c++) gr[k][c] = s_f[(c * NU + p) * LD + el * NF + k];
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
      for (int k = 0; k < NF; k++) s_f[(l * NU + p) * LD + el * NF + k] = fi[k];
    }
  }

  // ---- interface fluxes at the flux points (common_invFlux / common_viscFlux);
  //      this pass reads no LDS that the point pass writes, so no barrier
  //      separates them.  The normal transformed discontinuous flux that the
  //      reference subtracts here (extrapolate_totalFlux, daxpy) is folded into
  //      the divergence operator: D Ft + L (jump - N Ft) = (D - L N) Ft + L jump ----
#pragma unroll KEEP_NB ? MTK : 1
  for (int t = 0; t < MTK; t++) {
    const int task = tid + TetP<P>::BLOCK * t;
    const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
    if (task >= EPB * NFP || te >= N) continue;
    const long fe = fp + NFP * (long)te;
    int nb;
    if constexpr (KEEP_NB) nb = nbk[t];
    else nb = a.nbr[fe];
    const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
    nb &= 0x7fffffff;
    const long gi = geo_idx(fp, te);
    R nrm[ND];
#pragma unroll
    for (int l = 0; l < ND; l++) nrm[l] = R(a.fgeo[l * NN * GEO_PER_ELE + gi]);
    const R tdA = R(a.fgeo[ND * NN * GEO_PER_ELE + gi]);
    // own state trace, as the previous pass wrote it
    R uo[NF], un[NF], fn[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) uo[k] = (KEEP_UO0 && t == 0) ? uo0[k] : a.uf_in[fe * NF + k];
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
          const R fa = a.fv_in[fe * NF + k];
          const R fb = -a.fv_in[(long)nb * NF + k];
          jump[k] += ldg_normal<R>(fa, fb, uo[k], un[k], nn, pen, tau) * tdA;
        }
      }
    }
    R* q = s_q + fp * LD + tl * NF;
#pragma unroll
    for (int k = 0; k < NF; k++) q[k] = jump[k];
  }
  __syncthreads();

  // the RK register, detJ and the FP64 state are requested before the divergence
  // product, so that their latency overlaps it
  [[maybe_unused]] double resv[G][NF], uoldv[G][NF];
  [[maybe_unused]] double detjv[G];
  if constexpr (MODE == MODE_ADVANCE) {
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int el = g * E1 + el0, e = blockIdx.x * EPB + el;
      detjv[g] = 1.0;
      if (lane_on && e < N) {
        const long pe0 = p + NU * (long)e;
        detjv[g] = a.detj[vol_idx(pe0, e)];
        if (a.adv_type == 3) {
#pragma unroll
          for (int k = 0; k < NF; k++) resv[g][k] = a.res[pe0 + NU * NN * k];
        }
#pragma unroll
        for (int k = 0; k < NF; k++) uoldv[g][k] = KEEP_U ? ukeep[g][k] : a.u[pe0 + NU * NN * k];
      }
    }
  }

  // ---- divergence plus correction lift: [D' | L] [Ft; jump], D' = D - L N ----
  tet_product<R, P, TET_OP_D, M::MT_D, NT, M::KS_D, KF / 4, NFP / 4, true>(
      a.tD, s_f, s_q, NU, s_q, tid);
  __syncthreads();

  // ---- output: div, and the RK update with the new state into LDS ----
#pragma unroll
  for (int g = 0; g < G; g++) {
    const int el = g * E1 + el0, e = blockIdx.x * EPB + el;
    if (!(lane_on && e < N)) continue;
    R div[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) div[k] = s_q[p * LD + el * NF + k];
    const long pe = p + NU * (long)e;
    if (MODE == MODE_RESID || a.div) {
#pragma unroll
      for (int k = 0; k < NF; k++) a.div[pe + NU * NN * k] = double(div[k]);
    }
    if constexpr (MODE == MODE_ADVANCE) {
      const double idetj = rcp(detjv[g]);
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const long idx = pe + NU * NN * k;
        const double uold = uoldv[g][k];
        double unew;
        if (a.adv_type == 3) {
          // low-storage RK45 stage (src/eles.cpp:1155-1178)
          const double rhs = -double(div[k]) * idetj + a.ph.const_src;
          const double r = a.rk_a * resv[g][k] + a.dt * rhs;
          a.res[idx] = r;
          unew = uold + a.rk_b * r;
        } else {
          // forward Euler (src/eles.cpp:1067-1079)
          unew = uold - a.dt * (double(div[k]) * idetj - a.ph.const_src);
        }
        a.u[idx] = unew;
        s_u[p * LD + el * NF + k] = R(unew - (MIX ? ref_of(k) : 0.0));
      }
    }
  }
  if constexpr (MODE == MODE_ADVANCE) {
    __syncthreads();
    // ---- traces of the new state for the next stage: I u_new, staged in LDS and
    //      written element by element (each element's traces are contiguous) ----
    tet_product<R, P, TET_OP_I, M::MT_I, NT, M::KS_I, M::KS_I, 0>(
        a.tI, s_u, nullptr, NFP, s_q, tid);
    __syncthreads();
    tet_write_traces<P, R, G>(s_q, a.uf_out, N, tid);
  }
}

// Gradient-trace pass (TRACE_GRAD) for tetrahedra: LDG delta, corrected physical
// gradient at the solution points, then the viscous normal flux F_v(u_f, grad u_f).n
// at the flux points (plus, with LES, the extrapolated transformed SGS flux dotted
// with the normal, src/int_inters.cpp:392-427).  The gradient reaches the flux
// points one direction m at a time: I * (d u / d x_m) is one MFMA product into a
// [Nfp][LD] LDS buffer, and each flux-point task folds that column into its
// velocity / energy gradients (registers) before the next product.
template <int P, int VIS, bool AFF, bool MIX, int LES>
__global__ void __launch_bounds__(TetP<P>::BLOCK, 3)
fr_tet_mfma_grad_kernel(const StageArgsT<std::conditional_t<MIX, float, double>> a) {
  using R = std::conditional_t<MIX, float, double>;
  constexpr int ND = 3, NF = 5, NFACE = 4;
  constexpr bool SUTH = VIS == 2;
  static_assert(VIS != 0, "viscous only");
  using M = TetMfma<P>;
  constexpr int NU = M::NU, NFP = M::NFP, NFPF = TetP<P>::NFPF, EPB = M::EPB;
  constexpr int KU = M::KU, KF = M::KF, NT = M::NT, LD = M::LD;
  constexpr int KA = KU + NFP;                       // rows of u + delta
  constexpr int MT = (EPB * NFP + TetP<P>::BLOCK - 1) / TetP<P>::BLOCK;  // flux-point tasks per thread
  constexpr int RA = KA > KF ? KA : KF;               // rows of the first region
  static_assert(3 * NU <= KF, "gradient / SGS blocks fit their region");
  // Two regions.  The first holds u and the LDG delta, and the gradient product
  // writes the corrected gradient over them (a barrier separates its MFMAs from its
  // stores); the transformed SGS flux later overwrites the gradient point by point
  // once every flux point has its three gradient columns.  The second holds one
  // column of the gradient (or SGS flux) at the flux points.
  __shared__ R smem[(RA + NFP) * LD];
  R* s_u = smem;              // [KU][LD]  u (deviations on MIX)
  R* s_q = s_u + KU * LD;     // [NFP][LD] uo -> LDG delta
  R* s_f = smem;              // [KF][LD]  corrected reference gradient -> physical gradient, row m Nu + p
  R* s_s = s_f;               // [3 Nu][LD] transformed SGS flux, row l Nu + p (LES)
  R* s_o = smem + RA * LD;    // [NFP][LD] one column of the gradient (or SGS flux) at the flux points

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
  auto geo_idx = [&](int fpt, int ee) -> long { return AFF ? (long)(fpt / NFPF) + NFACE * (long)ee
                                                             : (long)fpt + NFP * (long)ee; };
  constexpr long GEO_PER_ELE = AFF ? NFACE : NFP;

  // padding rows of u and of the gradient (the last direction product reads
  // rows up to 2 Nu + KU)
  zero_rows<R, LD, TetP<P>::BLOCK>(smem, NU, KU, tid);
  zero_rows<R, LD, TetP<P>::BLOCK>(smem, 3 * NU, RA, tid);
  if (active) {
#pragma unroll
    for (int k = 0; k < NF; k++) {
      const double v = a.u[p + NU * (e + NN * k)];
      s_u[p * LD + el * NF + k] = MIX ? R(v - ref_of(k)) : R(v);
    }
  }
  __syncthreads();

  // ---- LDG delta (src/int_inters.cpp:325-346), over the own trace the previous
  //      pass wrote (the same values as I u here) ----
#pragma unroll 1
  for (int task = tid; task < EPB * NFP; task += TetP<P>::BLOCK) {
    const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
    if (te >= N) continue;
    int nb = a.nbr[fp + NFP * te];
    const R pen = R(nb < 0 ? -a.ph.pen_fact : a.ph.pen_fact);
    nb &= 0x7fffffff;
    R* q = s_q + fp * LD + tl * NF;
    R uo[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) uo[k] = a.uf_in[(fp + NFP * (long)te) * NF + k];
    if (nb & BDY_FLAG) {
      R ur[NF];
      ghost(uo, ur);
#pragma unroll
      for (int k = 0; k < NF; k++) q[k] = R(0.5) * (ur[k] + uo[k]) - uo[k];
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) {
        const R un = a.uf_in[(long)nb * NF + k];
        const R uc = R(0.5) * (uo[k] + un) - pen * (uo[k] - un);
        q[k] = uc - uo[k];
      }
    }
  }
  __syncthreads();
  // ---- corrected reference gradient (src/eles.cpp:1667-1870) ----
  tet_product<R, P, TET_OP_G, M::MT_G, NT, M::KS_G, KU / 4, NFP / 4, true>(
      a.tG, s_u, s_q, 3 * NU, s_f, tid);
  __syncthreads();

  // ---- per solution point: physical gradient over the reference one (same point
  //      and element), LES: transformed SGS flux over u (own entries) and delta ----
  if (active) {
    const long pe = p + NU * (long)e;
    const long vi = AFF ? (long)e : pe;
    constexpr long VOL = AFF ? 1 : NU;
    R gr[NF][ND];
#pragma unroll
    for (int k = 0; k < NF; k++)
#pragma unroll
      for (int c = 0; c < ND; c++) gr[k][c] = s_f[(c * NU + p) * LD + el * NF + k];
    R Jg[ND][ND];
#pragma unroll
    for (int l = 0; l < ND; l++)
#pragma unroll
      for (int m = 0; m < ND; m++) Jg[l][m] = a.jg[(l * ND + m) * NN * VOL + vi];
    const R inv_detj = R(rcp(a.detj[vi]));
    R gf[NF][ND];
#pragma unroll
    for (int k = 0; k < NF; k++)
#pragma unroll
      for (int m = 0; m < ND; m++) {
        R v = 0;
#pragma unroll
        for (int d = 0; d < ND; d++) v += gr[k][d] * (Jg[d][m] * inv_detj);
        gf[k][m] = v;
      }
#pragma unroll
    for (int k = 0; k < NF; k++)
#pragma unroll
      for (int m = 0; m < ND; m++) s_f[(m * NU + p) * LD + el * NF + k] = gf[k][m];
  }
  __syncthreads();

  // ---- the gradient at the flux points, one direction at a time ----
  R dv[MT][ND][ND], de[MT][ND], fs[MT][NF];
#pragma unroll
  for (int t = 0; t < MT; t++)
#pragma unroll
    for (int k = 0; k < NF; k++) fs[t][k] = R(0);
  auto valid = [&](int task) { return task < EPB * NFP && blockIdx.x * EPB + task / NFP < N; };
  // each task's own state trace: in FP64 read once for the four passes that use
  // it (the mixed kernel would lose a wave per SIMD to the registers and reads
  // it again each time)
  constexpr bool KEEP_UO = !MIX;
  [[maybe_unused]] R uo_t[KEEP_UO ? MT : 1][NF];
  if constexpr (KEEP_UO) {
#pragma unroll
    for (int t = 0; t < MT; t++) {
      const int task = tid + TetP<P>::BLOCK * t;
      const int tl = task / NFP, fp = task - tl * NFP;
      const long fe = fp + NFP * (long)(blockIdx.x * EPB + tl);
#pragma unroll
      for (int k = 0; k < NF; k++) uo_t[t][k] = valid(task) ? a.uf_in[fe * NF + k] : R(1);
    }
  }
  auto own_prim = [&](int t, ViscPrim<ND, R>& q) {
    R uo[NF], uu[NF];
    if constexpr (KEEP_UO) {
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] = uo_t[t][k];
    } else {
      const int task = tid + TetP<P>::BLOCK * t;
      const int tl = task / NFP, fp = task - tl * NFP;
      const long fe = fp + NFP * (long)(blockIdx.x * EPB + tl);
#pragma unroll
      for (int k = 0; k < NF; k++) uo[k] = a.uf_in[fe * NF + k];
    }
    full(uo, uu);
    visc_prim<ND, R>(uu, q);
  };
#pragma unroll
  for (int m = 0; m < ND; m++) {
    tet_product<R, P, TET_OP_I, M::MT_I, NT, M::KS_I, M::KS_I, 0>(
        a.tI, s_f + m * NU * LD, nullptr, NFP, s_o, tid);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MT; t++) {
      const int task = tid + TetP<P>::BLOCK * t;
      if (!valid(task)) continue;
      const int tl = task / NFP, fp = task - tl * NFP;
      R g[NF];
#pragma unroll
      for (int k = 0; k < NF; k++) g[k] = s_o[fp * LD + tl * NF + k];
      if (a.gf_out) {
        const long fe = fp + NFP * (long)(blockIdx.x * EPB + tl);
#pragma unroll
        for (int k = 0; k < NF; k++) a.gf_out[fe + NFP * NN * (k + NF * m)] = g[k];
      }
      ViscPrim<ND, R> q;
      own_prim(t, q);
      visc_column<ND, R>(q, g, dv[t], m, de[t][m]);
    }
    __syncthreads();
  }
  if constexpr (LES != 0) {
    // transformed SGS flux at the solution points, over the point's own gradient
    // entries: sgsf(k,l) = sum_m JGinv(l,m) F_sgs(k,m) (src/eles.cpp:2369-2417)
    if (active) {
      const long pe = p + NU * (long)e;
      const long vi = AFF ? (long)e : pe;
      constexpr long VOL = AFF ? 1 : NU;
      R gf[NF][ND], uo[NF], uu[NF], sgM[ND][ND], sgE[ND];
#pragma unroll
      for (int k = 0; k < NF; k++) {
#pragma unroll
        for (int m = 0; m < ND; m++) gf[k][m] = s_f[(m * NU + p) * LD + el * NF + k];
        const double v = a.u[p + NU * (e + NN * k)];
        uo[k] = MIX ? R(v - ref_of(k)) : R(v);
      }
      full(uo, uu);
      sgs_flux<ND, LES, R>(uu, gf, R(a.detj[vi] * 8. / 6.), R(a.filter_ratio), P, gam, sgM, sgE);
      R Jg[ND][ND];
#pragma unroll
      for (int l = 0; l < ND; l++)
#pragma unroll
        for (int m = 0; m < ND; m++) Jg[l][m] = a.jg[(l * ND + m) * NN * VOL + vi];
#pragma unroll
      for (int l = 0; l < ND; l++) {
        R* srow = s_s + (l * NU + p) * LD + el * NF;
        srow[0] = R(0);
#pragma unroll
        for (int i = 0; i < ND; i++) {
          R v = 0;
#pragma unroll
          for (int m = 0; m < ND; m++) v += Jg[l][m] * sgM[i][m];
          srow[1 + i] = v;
        }
        R v = 0;
#pragma unroll
        for (int m = 0; m < ND; m++) v += Jg[l][m] * sgE[m];
        srow[NF - 1] = v;
      }
    }
    __syncthreads();
    // extrapolated transformed SGS flux dotted with the physical normal
#pragma unroll
    for (int l = 0; l < ND; l++) {
      tet_product<R, P, TET_OP_I, M::MT_I, NT, M::KS_I, M::KS_I, 0>(
          a.tI, s_s + l * NU * LD, nullptr, NFP, s_o, tid);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < MT; t++) {
        const int task = tid + TetP<P>::BLOCK * t;
        if (!valid(task)) continue;
        const int tl = task / NFP, fp = task - tl * NFP;
        const R nl = R(a.fgeo[l * NN * GEO_PER_ELE + geo_idx(fp, blockIdx.x * EPB + tl)]);
#pragma unroll
        for (int k = 1; k < NF; k++) fs[t][k] += nl * s_o[fp * LD + tl * NF + k];
      }
      __syncthreads();
    }
  }
  // ---- viscous normal flux traces ----
#pragma unroll
  for (int t = 0; t < MT; t++) {
    const int task = tid + TetP<P>::BLOCK * t;
    if (!valid(task)) continue;
    const int tl = task / NFP, fp = task - tl * NFP, te = blockIdx.x * EPB + tl;
    const long fe = fp + NFP * (long)te;
    const long gi = geo_idx(fp, te);
    R nrm[ND];
#pragma unroll
    for (int l = 0; l < ND; l++) nrm[l] = R(a.fgeo[l * NN * GEO_PER_ELE + gi]);
    ViscPrim<ND, R> q;
    own_prim(t, q);
    ViscT<ND, R> vt;
    visc_finish<ND, SUTH, R>(q, dv[t], de[t], a.ph, vt);
    R fa[NF];
    visc_flux_dot<ND, R>(vt, nrm, fa);
#pragma unroll
    for (int k = 0; k < NF; k++) a.fv_out[fe * NF + k] = fa[k] + fs[t][k];
  }
}

}  // namespace hifr
