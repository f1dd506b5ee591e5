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

template <int P, int ND, int VIS, int MODE, bool AFF, bool MIX>
hipError_t launch_aff(const Args<MIX>& a, hipStream_t s) {
  constexpr int EPB = Cfg<P, ND>::EPB;
  using T = TP<P, ND, EPB>;
  const int grid = (a.N + EPB - 1) / EPB;
  hipLaunchKernelGGL((fr_tensor_kernel<P, ND, EPB, VIS, MODE, AFF, MIX>), dim3(grid), dim3(T::BLOCK), 0, s, a);
  return hipGetLastError();
}

// the trace kernel reads no metrics: one instantiation
template <int P, int ND, int VIS, int MODE, bool MIX>
hipError_t launch_one(const Args<MIX>& a, hipStream_t s, bool aff) {
  if constexpr (MODE == MODE_TRACE_U) return launch_aff<P, ND, VIS, MODE, false, MIX>(a, s);
  return aff ? launch_aff<P, ND, VIS, MODE, true, MIX>(a, s) : launch_aff<P, ND, VIS, MODE, false, MIX>(a, s);
}

template <int P, int ND, bool MIX>
hipError_t launch_mode(int mode, int vis, const Args<MIX>& a, hipStream_t s, bool f) {
  switch (mode) {
    case MODE_TRACE_U: return launch_one<P, ND, 0, MODE_TRACE_U, MIX>(a, s, f);
    case MODE_TRACE_GRAD:
      return vis == 2 ? launch_one<P, ND, 2, MODE_TRACE_GRAD, MIX>(a, s, f)
                      : launch_one<P, ND, 1, MODE_TRACE_GRAD, MIX>(a, s, f);
    case MODE_RESID:
      return vis == 2 ? launch_one<P, ND, 2, MODE_RESID, MIX>(a, s, f)
           : vis == 1 ? launch_one<P, ND, 1, MODE_RESID, MIX>(a, s, f)
                      : launch_one<P, ND, 0, MODE_RESID, MIX>(a, s, f);
    default:
      return vis == 2 ? launch_one<P, ND, 2, MODE_ADVANCE, MIX>(a, s, f)
           : vis == 1 ? launch_one<P, ND, 1, MODE_ADVANCE, MIX>(a, s, f)
                      : launch_one<P, ND, 0, MODE_ADVANCE, MIX>(a, s, f);
  }
}

// hexes P = 1..5, quads P = 1..6 (callers check the order first)
template <bool MIX>
hipError_t launch(int nd, int P, int mode, int vis, bool aff, const Args<MIX>& a, hipStream_t s) {
  if (nd == 3) {
    switch (P) {
      case 1: return launch_mode<1, 3, MIX>(mode, vis, a, s, aff);
      case 2: return launch_mode<2, 3, MIX>(mode, vis, a, s, aff);
      case 3: return launch_mode<3, 3, MIX>(mode, vis, a, s, aff);
      case 4: return launch_mode<4, 3, MIX>(mode, vis, a, s, aff);
      default: return launch_mode<5, 3, MIX>(mode, vis, a, s, aff);
    }
  }
  switch (P) {
    case 1: return launch_mode<1, 2, MIX>(mode, vis, a, s, aff);
    case 2: return launch_mode<2, 2, MIX>(mode, vis, a, s, aff);
    case 3: return launch_mode<3, 2, MIX>(mode, vis, a, s, aff);
    case 4: return launch_mode<4, 2, MIX>(mode, vis, a, s, aff);
    case 5: return launch_mode<5, 2, MIX>(mode, vis, a, s, aff);
    default: return launch_mode<6, 2, MIX>(mode, vis, a, s, aff);
  }
}

}  // namespace dispatch
}  // namespace hifr
