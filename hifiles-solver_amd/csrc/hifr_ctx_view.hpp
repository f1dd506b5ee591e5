// Read-only view of a context's device arrays for entry points compiled in other
// translation units (hifr_ctx itself is private to hifr_device.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/hifr.h"

namespace hifr {

struct CtxView {
  int ele_type, order, N;
  bool have_state, have_metrics, affine, mixed;
  hipStream_t stream;
  const double* u;     // state (Nu, N, nf)
  const double* detj;  // per element when affine, else per point
  const double* jg;    // JGinv component-major (FP64 contexts; null when mixed)
  double mu;           // constant molecular viscosity mu_inf
};

hifr_status ctx_view(const hifr_ctx* c, CtxView& v);

}  // namespace hifr
