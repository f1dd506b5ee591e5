// Mixed-precision instantiations of the element-block FR kernels.
#include "fr_launch_impl.hpp"

namespace hifr {
hipError_t launch_mixed(int nd, int P, int mode, int vis, bool aff, int les, const StageArgsT<float>& a,
                       hipStream_t s) {
  return dispatch::launch<true>(nd, P, mode, vis, aff, les, a, s);
}
}  // namespace hifr
