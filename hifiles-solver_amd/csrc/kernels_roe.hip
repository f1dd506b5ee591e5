// Roe-flux instantiations of the element-block FR kernels (FP64; quads, hexes, tets).
#include "fr_launch_impl.hpp"

namespace hifr {
hipError_t launch_roe_fp64(int nd, int P, int mode, int vis, bool aff, const StageArgsT<double>& a,
                           hipStream_t s) {
  return dispatch::launch_roe(nd, P, mode, vis, aff, a, s);
}
hipError_t launch_tet_roe_fp64(int P, int mode, int vis, bool aff, const StageArgsT<double>& a, hipStream_t s) {
  return dispatch::launch_tet_roe(P, mode, vis, aff, a, s);
}
}  // namespace hifr
