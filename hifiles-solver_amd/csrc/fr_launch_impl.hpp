// Template dispatch shared by kernels_fp64.hip and kernels_mixed.hip.
#pragma once
#include "fr_launch.hpp"

namespace hifr {
namespace dispatch {

template <int P, int ND>
struct Cfg {
  // elements per workgroup: about 128 threads per block
  static constexpr int EPB = TP<P, ND, 1>::NU >= 128 ? 1 : 128 / TP<P, ND, 1>::NU;
};

template <bool MIX>
using Args = StageArgsT<std::conditional_t<MIX, float, double>>;

template <int P, int ND, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE = 0>
hipError_t launch_les(const Args<MIX>& a, hipStream_t s) {
  constexpr int EPB = Cfg<P, ND>::EPB;
  using T = TP<P, ND, EPB>;
  const int grid = (a.N + EPB - 1) / EPB;
  hipLaunchKernelGGL((fr_tensor_kernel<P, ND, EPB, VIS, MODE, AFF, MIX, LES, RIE>), dim3(grid), dim3(T::BLOCK), 0,
                     s, a);
  return hipGetLastError();
}

// LES variants exist for the constant-viscosity Navier-Stokes kernels only; the
// context rejects other combinations before launching
template <int P, int ND, int VIS, int MODE, bool AFF, bool MIX>
hipError_t launch_aff(const Args<MIX>& a, hipStream_t s, int les) {
  if constexpr (VIS == 1 && MODE != MODE_TRACE_U) {
    if (les == 1) return launch_les<P, ND, VIS, MODE, AFF, MIX, 1>(a, s);
    if (les == 2) return launch_les<P, ND, VIS, MODE, AFF, MIX, 2>(a, s);
    // learned closure: hexes P = 2, 3
    if constexpr (ND == 3 && (P == 2 || P == 3))
      if (les == 3) return launch_les<P, ND, VIS, MODE, AFF, MIX, 3>(a, s);
  }
  if (les != 0) return hipErrorInvalidValue;
  return launch_les<P, ND, VIS, MODE, AFF, MIX, 0>(a, s);
}

// the trace kernel reads no metrics: one instantiation
template <int P, int ND, int VIS, int MODE, bool MIX>
hipError_t launch_one(const Args<MIX>& a, hipStream_t s, bool aff, int les) {
  if constexpr (MODE == MODE_TRACE_U) return launch_aff<P, ND, VIS, MODE, false, MIX>(a, s, 0);
  return aff ? launch_aff<P, ND, VIS, MODE, true, MIX>(a, s, les) : launch_aff<P, ND, VIS, MODE, false, MIX>(a, s, les);
}

template <int P, int ND, bool MIX>
hipError_t launch_mode(int mode, int vis, const Args<MIX>& a, hipStream_t s, bool f, int les) {
  switch (mode) {
    case MODE_TRACE_U: return launch_one<P, ND, 0, MODE_TRACE_U, MIX>(a, s, f, les);
    case MODE_TRACE_GRAD:
      return vis == 2 ? launch_one<P, ND, 2, MODE_TRACE_GRAD, MIX>(a, s, f, les)
                      : launch_one<P, ND, 1, MODE_TRACE_GRAD, MIX>(a, s, f, les);
    case MODE_RESID:
      return vis == 2 ? launch_one<P, ND, 2, MODE_RESID, MIX>(a, s, f, les)
           : vis == 1 ? launch_one<P, ND, 1, MODE_RESID, MIX>(a, s, f, les)
                      : launch_one<P, ND, 0, MODE_RESID, MIX>(a, s, f, les);
    default:
      return vis == 2 ? launch_one<P, ND, 2, MODE_ADVANCE, MIX>(a, s, f, les)
           : vis == 1 ? launch_one<P, ND, 1, MODE_ADVANCE, MIX>(a, s, f, les)
                      : launch_one<P, ND, 0, MODE_ADVANCE, MIX>(a, s, f, les);
  }
}

// hexes P = 1..5, quads P = 1..6 (callers check the order first)
template <bool MIX>
hipError_t launch(int nd, int P, int mode, int vis, bool aff, int les, const Args<MIX>& a, hipStream_t s) {
  if (nd == 3) {
    switch (P) {
      case 1: return launch_mode<1, 3, MIX>(mode, vis, a, s, aff, les);
      case 2: return launch_mode<2, 3, MIX>(mode, vis, a, s, aff, les);
      case 3: return launch_mode<3, 3, MIX>(mode, vis, a, s, aff, les);
      case 4: return launch_mode<4, 3, MIX>(mode, vis, a, s, aff, les);
      default: return launch_mode<5, 3, MIX>(mode, vis, a, s, aff, les);
    }
  }
  switch (P) {
    case 1: return launch_mode<1, 2, MIX>(mode, vis, a, s, aff, les);
    case 2: return launch_mode<2, 2, MIX>(mode, vis, a, s, aff, les);
    case 3: return launch_mode<3, 2, MIX>(mode, vis, a, s, aff, les);
    case 4: return launch_mode<4, 2, MIX>(mode, vis, a, s, aff, les);
    case 5: return launch_mode<5, 2, MIX>(mode, vis, a, s, aff, les);
    default: return launch_mode<6, 2, MIX>(mode, vis, a, s, aff, les);
  }
}

// ---- Roe interface flux (FP64, no LES): the flux kernels only; the trace kernels
// do not evaluate the inviscid flux ----
template <int P, int ND, int VIS>
hipError_t roe_vis(int mode, bool aff, const Args<false>& a, hipStream_t s) {
  if (mode == MODE_RESID)
    return aff ? launch_les<P, ND, VIS, MODE_RESID, true, false, 0, 2>(a, s)
               : launch_les<P, ND, VIS, MODE_RESID, false, false, 0, 2>(a, s);
  return aff ? launch_les<P, ND, VIS, MODE_ADVANCE, true, false, 0, 2>(a, s)
             : launch_les<P, ND, VIS, MODE_ADVANCE, false, false, 0, 2>(a, s);
}

template <int P, int ND>
hipError_t roe_order(int mode, int vis, bool aff, const Args<false>& a, hipStream_t s) {
  return vis == 2 ? roe_vis<P, ND, 2>(mode, aff, a, s)
       : vis == 1 ? roe_vis<P, ND, 1>(mode, aff, a, s)
                  : roe_vis<P, ND, 0>(mode, aff, a, s);
}

// hexes P = 1..5, quads P = 1..6; mode RESID or ADVANCE
inline hipError_t launch_roe(int nd, int P, int mode, int vis, bool aff, const Args<false>& a, hipStream_t s) {
  if (nd == 3) {
    switch (P) {
      case 1: return roe_order<1, 3>(mode, vis, aff, a, s);
      case 2: return roe_order<2, 3>(mode, vis, aff, a, s);
      case 3: return roe_order<3, 3>(mode, vis, aff, a, s);
      case 4: return roe_order<4, 3>(mode, vis, aff, a, s);
      default: return roe_order<5, 3>(mode, vis, aff, a, s);
    }
  }
  switch (P) {
    case 1: return roe_order<1, 2>(mode, vis, aff, a, s);
    case 2: return roe_order<2, 2>(mode, vis, aff, a, s);
    case 3: return roe_order<3, 2>(mode, vis, aff, a, s);
    case 4: return roe_order<4, 2>(mode, vis, aff, a, s);
    case 5: return roe_order<5, 2>(mode, vis, aff, a, s);
    default: return roe_order<6, 2>(mode, vis, aff, a, s);
  }
}

// ---- tetrahedra ----
template <int P, int VIS, int MODE, bool AFF, bool MIX, int LES, int RIE = 0>
hipError_t tet_les(const Args<MIX>& a, hipStream_t s) {
  using T = TetP<P>;
  const int grid = (a.N + T::EPB - 1) / T::EPB;
  if constexpr (MODE == MODE_RESID || MODE == MODE_ADVANCE) {
    // flux pass: operator products on MFMA (G column groups of EPB elements per
    // block; two groups at P = 3 halve the operator-tile loads per element but
    // need twice the LDS per block, and measured 20 % slower)
    constexpr int G = 1;
    const int fgrid = (a.N + G * T::EPB - 1) / (G * T::EPB);
    hipLaunchKernelGGL((fr_tet_mfma_kernel<P, VIS, MODE, AFF, MIX, LES, RIE, G>), dim3(fgrid), dim3(T::BLOCK), 0, s,
                       a);
  } else if constexpr (MODE == MODE_TRACE_GRAD) {
    hipLaunchKernelGGL((fr_tet_mfma_grad_kernel<P, VIS, AFF, MIX, LES>), dim3(grid), dim3(T::BLOCK), 0, s, a);
  } else {
    hipLaunchKernelGGL((fr_tet_mfma_trace_kernel<P, MIX>), dim3(grid), dim3(T::BLOCK), 0, s, a);
  }
  return hipGetLastError();
}

template <int P, int VIS, int MODE, bool AFF, bool MIX>
hipError_t tet_aff(const Args<MIX>& a, hipStream_t s, int les) {
  if constexpr (VIS == 1 && MODE != MODE_TRACE_U) {
    if (les == 1) return tet_les<P, VIS, MODE, AFF, MIX, 1>(a, s);
    if (les == 2) return tet_les<P, VIS, MODE, AFF, MIX, 2>(a, s);
  }
  if (les != 0) return hipErrorInvalidValue;
  return tet_les<P, VIS, MODE, AFF, MIX, 0>(a, s);
}

template <int P, int VIS, int MODE, bool MIX>
hipError_t tet_one(const Args<MIX>& a, hipStream_t s, bool aff, int les) {
  if constexpr (MODE == MODE_TRACE_U) return tet_aff<P, VIS, MODE, false, MIX>(a, s, 0);
  return aff ? tet_aff<P, VIS, MODE, true, MIX>(a, s, les) : tet_aff<P, VIS, MODE, false, MIX>(a, s, les);
}

template <int P, bool MIX>
hipError_t tet_mode(int mode, int vis, const Args<MIX>& a, hipStream_t s, bool f, int les) {
  switch (mode) {
    case MODE_TRACE_U: return tet_one<P, 0, MODE_TRACE_U, MIX>(a, s, f, les);
    case MODE_TRACE_GRAD:
      return vis == 2 ? tet_one<P, 2, MODE_TRACE_GRAD, MIX>(a, s, f, les)
                      : tet_one<P, 1, MODE_TRACE_GRAD, MIX>(a, s, f, les);
    case MODE_RESID:
      return vis == 2 ? tet_one<P, 2, MODE_RESID, MIX>(a, s, f, les)
           : vis == 1 ? tet_one<P, 1, MODE_RESID, MIX>(a, s, f, les)
                      : tet_one<P, 0, MODE_RESID, MIX>(a, s, f, les);
    default:
      return vis == 2 ? tet_one<P, 2, MODE_ADVANCE, MIX>(a, s, f, les)
           : vis == 1 ? tet_one<P, 1, MODE_ADVANCE, MIX>(a, s, f, les)
                      : tet_one<P, 0, MODE_ADVANCE, MIX>(a, s, f, les);
  }
}

template <bool MIX>
hipError_t launch_tet(int P, int mode, int vis, bool aff, int les, const Args<MIX>& a, hipStream_t s) {
  switch (P) {
    case 1: return tet_mode<1, MIX>(mode, vis, a, s, aff, les);
    case 2: return tet_mode<2, MIX>(mode, vis, a, s, aff, les);
    default: return tet_mode<3, MIX>(mode, vis, a, s, aff, les);
  }
}

template <int P, int VIS>
hipError_t tet_roe_vis(int mode, bool aff, const Args<false>& a, hipStream_t s) {
  if (mode == MODE_RESID)
    return aff ? tet_les<P, VIS, MODE_RESID, true, false, 0, 2>(a, s)
               : tet_les<P, VIS, MODE_RESID, false, false, 0, 2>(a, s);
  return aff ? tet_les<P, VIS, MODE_ADVANCE, true, false, 0, 2>(a, s)
             : tet_les<P, VIS, MODE_ADVANCE, false, false, 0, 2>(a, s);
}

template <int P>
hipError_t tet_roe_order(int mode, int vis, bool aff, const Args<false>& a, hipStream_t s) {
  return vis == 2 ? tet_roe_vis<P, 2>(mode, aff, a, s)
       : vis == 1 ? tet_roe_vis<P, 1>(mode, aff, a, s)
                  : tet_roe_vis<P, 0>(mode, aff, a, s);
}

// tets P = 1..3; mode RESID or ADVANCE
inline hipError_t launch_tet_roe(int P, int mode, int vis, bool aff, const Args<false>& a, hipStream_t s) {
  switch (P) {
    case 1: return tet_roe_order<1>(mode, vis, aff, a, s);
    case 2: return tet_roe_order<2>(mode, vis, aff, a, s);
    default: return tet_roe_order<3>(mode, vis, aff, a, s);
  }
}

}  // namespace dispatch
}  // namespace hifr
