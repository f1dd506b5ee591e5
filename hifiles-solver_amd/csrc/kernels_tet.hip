// Tetrahedron instantiations of the element-block FR kernels (FP64 and mixed).
#include "fr_launch_impl.hpp"

namespace hifr {
hipError_t launch_tet_fp64(int P, int mode, int vis, bool aff, int les, const StageArgsT<double>& a,
                           hipStream_t s) {
  return dispatch::launch_tet<false>(P, mode, vis, aff, les, a, s);
}
hipError_t launch_tet_mixed(int P, int mode, int vis, bool aff, int les, const StageArgsT<float>& a,
                            hipStream_t s) {
  return dispatch::launch_tet<true>(P, mode, vis, aff, les, a, s);
}
}  // namespace hifr
