// Host-side state of a learned eddy-viscosity network (include/hifr_sgs.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace hifr {
namespace mlp {

// Switch of the learned closure's device entry points (network training and
// inference, training samples, sgs_model HIFR_SGS_LEARNED).  While false, they
// return HIFR_ERR_UNSUPPORTED before touching the device.
constexpr bool kLearnedEnabled = true;

// Input standardisation, training range and output scale of the closure.
struct Norm {
  float mean[10], scale[10], lo[10], hi[10];
  float out_scale;
};

}  // namespace mlp
}  // namespace hifr

struct hifr_mlp {
  hipStream_t stream = nullptr;
  float* params = nullptr;            // FP32 master parameters (9,089)
  float *adam_m = nullptr, *adam_v = nullptr;
  unsigned short *w1 = nullptr, *w2 = nullptr, *w3 = nullptr;  // BF16 weights, forward layouts
  float* w4 = nullptr;                // BF16-rounded output weights as FP32
  float* scalar = nullptr;            // 64 floats: loss accumulators
  float* part = nullptr;              // per-block partial gradients
  float* part_loss = nullptr;
  size_t part_bytes = 0, part_loss_bytes = 0;
  float *x = nullptr, *t = nullptr;   // training set
  size_t x_bytes = 0, t_bytes = 0;
  long n = 0;
  float *bx = nullptr, *bt = nullptr, *by = nullptr;  // staging for loss_grad / predict
  size_t bx_bytes = 0, bt_bytes = 0, by_bytes = 0;
  long step = 0;                      // Adam step count
  hifr::mlp::Norm norm;
};
