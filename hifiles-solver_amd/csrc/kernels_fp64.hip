// FP64 instantiations of the element-block FR kernels.
#include "fr_launch_impl.hpp"

namespace hifr {
hipError_t launch_fp64(int nd, int P, int mode, int vis, bool aff, int les, const StageArgsT<double>& a,
                       hipStream_t s) {
  return dispatch::launch<false>(nd, P, mode, vis, aff, les, a, s);
}
}  // namespace hifr
