// The learned eddy-viscosity network on gfx950: BF16 MFMA tiles of 16 samples per
// wavefront (v_mfma_f32_16x16x16_bf16), FP32 accumulation.  Shared by the training
// kernels (sgs_mlp.hip) and the LES flux pass (fr_tensor_kernels.hpp).
//
// MFMA 16x16x16 operand layout (lane l = 0..63, r = l % 16, g = l / 16):
//   A (16 x 16, M x K): lane l holds A[r][4g + i], i = 0..3
//   B (16 x 16, K x N): lane l holds B[4g + i][r]
//   C (16 x 16, M x N): lane l holds C[4g + i][r]
// With features as rows and samples as columns, a layer's output tile in C layout
// is, lane for lane, the B operand of the next layer for the same 16 features, so
// the forward pass needs no data movement between layers.
#pragma once
#include <hip/hip_runtime.h>

namespace hifr {
namespace mlp {

constexpr int NIN = 10, NPAD = 16, H = 64;
constexpr int OFF_W1 = 0, OFF_B1 = 640, OFF_W2 = 704, OFF_B2 = 4800, OFF_W3 = 4864, OFF_B3 = 8960,
              OFF_W4 = 9024, OFF_B4 = 9088, NPARAM = 9089;

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even: the hardware conversion (v_cvt_pk_bf16_f32,
// two values per instruction)
__device__ __forceinline__ unsigned short to_bf16(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float from_bf16(unsigned short h) { return __uint_as_float((unsigned)h << 16); }
// logistic function from the hardware exponential and reciprocal estimates (each
// within an ulp or two in FP32, far below the BF16 rounding of the activations)
__device__ __forceinline__ float sigmoidf(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float siluf(float z) { return z * sigmoidf(z); }
__device__ __forceinline__ float d_siluf(float z) {
  const float s = sigmoidf(z);
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ float softplusf(float z) { return z > 20.0f ? z : log1pf(__expf(z)); }

__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
  struct {
    bf2 lo, hi;
  } v = {__builtin_convertvector((f2){a, b}, bf2), __builtin_convertvector((f2){c, d}, bf2)};
  return __builtin_bit_cast(s4, v);
}
__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// Network weights staged in LDS: BF16 W1 [64][16] (inputs 10..15 zero), W2, W3
// [64][64] ([out][in]), w4 as the FP32 value of its BF16 rounding, FP32 biases.
struct alignas(16) Lds {
  unsigned short w1[H * NPAD];
  unsigned short w2[H * H];
  unsigned short w3[H * H];
  float w4[H];
  float b[3 * H + 1];
};

// Device-side copy of the weights in the layouts above (global memory).
struct Weights {
  const unsigned short* w1;  // [64][16]
  const unsigned short* w2;  // [64][64]
  const unsigned short* w3;  // [64][64]
  const float* w4;           // [64] (BF16-rounded values)
  const float* params;       // FP32 master vector: biases are read from it
};

// cooperative copy of the weights into LDS by nthreads threads (tid < nthreads)
__device__ __forceinline__ void stage(const Weights& w, Lds& s, int tid, int nthreads) {
  for (int i = tid; i < H * NPAD / 2; i += nthreads)
    reinterpret_cast<unsigned*>(s.w1)[i] = reinterpret_cast<const unsigned*>(w.w1)[i];
  for (int i = tid; i < H * H / 2; i += nthreads) {
    reinterpret_cast<unsigned*>(s.w2)[i] = reinterpret_cast<const unsigned*>(w.w2)[i];
    reinterpret_cast<unsigned*>(s.w3)[i] = reinterpret_cast<const unsigned*>(w.w3)[i];
  }
  for (int i = tid; i < H; i += nthreads) {
    s.w4[i] = w.w4[i];
    s.b[i] = w.params[OFF_B1 + i];
    s.b[H + i] = w.params[OFF_B2 + i];
    s.b[2 * H + i] = w.params[OFF_B3 + i];
  }
  if (tid == 0) s.b[3 * H] = w.params[OFF_B4];
}

// Per-lane state of one 16-sample tile through the forward pass.
struct Tile {
  s4 a0;           // inputs (B operand, features 4g..4g+3 of 16)
  s4 a[3][4];      // BF16 activations of layers 1-3, k-tile kt (B operand layout)
  float z[3][4][4];  // pre-activations, [layer][row tile][i] for feature 16 rt + 4g + i
  float z4;        // output pre-activation of sample r
};

// Forward pass of one tile; x_of(f) gives the lane's input feature f (< 10) of
// sample r.  Every lane of the wavefront must call it (MFMA).
template <class X>
__device__ __forceinline__ void forward(const Lds& s, Tile& t, X x_of, int lane) {
  const int r = lane & 15, g = lane >> 4;
  float xin[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int f = 4 * g + i;
    xin[i] = f < NIN ? x_of(f) : 0.0f;
  }
  t.a0 = pack4(xin[0], xin[1], xin[2], xin[3]);
  // layer 1: K = 16 (one k tile)
#pragma unroll
  for (int rt = 0; rt < 4; rt++) {
    const s4 A = *reinterpret_cast<const s4*>(&s.w1[(16 * rt + r) * NPAD + 4 * g]);
    f4 acc = {0.f, 0.f, 0.f, 0.f};
    acc = mfma(A, t.a0, acc);
    float h[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const float z = acc[i] + s.b[16 * rt + 4 * g + i];
      t.z[0][rt][i] = z;
      h[i] = siluf(z);
    }
    t.a[0][rt] = pack4(h[0], h[1], h[2], h[3]);
  }
  // layers 2, 3: K = 64
#pragma unroll
  for (int l = 1; l < 3; l++) {
    const unsigned short* W = l == 1 ? s.w2 : s.w3;
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kt = 0; kt < 4; kt++) {
        const s4 A = *reinterpret_cast<const s4*>(&W[(16 * rt + r) * H + 16 * kt + 4 * g]);
        acc = mfma(A, t.a[l - 1][kt], acc);
      }
      float h[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float z = acc[i] + s.b[l * H + 16 * rt + 4 * g + i];
        t.z[l][rt][i] = z;
        h[i] = siluf(z);
      }
      t.a[l][rt] = pack4(h[0], h[1], h[2], h[3]);
    }
  }
  // output: each lane sums its 16 features of sample r, then the 4 lane groups
  float p = 0.0f;
#pragma unroll
  for (int rt = 0; rt < 4; rt++)
#pragma unroll
    for (int i = 0; i < 4; i++) p += s.w4[16 * rt + 4 * g + i] * from_bf16((unsigned short)t.a[2][rt][i]);
  p += __shfl_xor(p, 16);
  p += __shfl_xor(p, 32);
  t.z4 = p + s.b[3 * H];
}

}  // namespace mlp
}  // namespace hifr
