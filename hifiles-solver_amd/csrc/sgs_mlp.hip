// Training and inference of the learned eddy-viscosity network (include/hifr_sgs.h)
// on gfx950: forward and backward passes as BF16 MFMA tiles with FP32
// accumulation, per-workgroup partial gradients reduced by the Adam kernel.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/hifr_sgs.h"
#include "sgs_mlp.hpp"
#include "sgs_mlp_state.hpp"

namespace {

using namespace hifr::mlp;

constexpr int WAVES = 4, BLOCK = 64 * WAVES;

#define HIP_TRY(x)                                     \
  do {                                                 \
    if ((x) != hipSuccess) return HIFR_ERR_DEVICE;     \
  } while (0)

// Forward (and, when train, backward) over the samples [off, off + n) in tiles of
// 16 per wavefront.  Every wavefront of the grid runs the same number of tile
// iterations (tiles past the end are masked), so the block barriers are uniform.
// train: writes the block's partial gradient (NPARAM floats) and partial loss;
// otherwise writes y.
template <bool TRAIN>
__global__ void __launch_bounds__(BLOCK) mlp_kernel(Weights w, const float* __restrict__ x,
                                                    const float* __restrict__ t, long off, long n, float inv_b,
                                                    float* __restrict__ part, float* __restrict__ part_loss,
                                                    float* __restrict__ y) {
  __shared__ Lds s;
  __shared__ float sg[TRAIN ? NPARAM : 1];
  __shared__ alignas(16) unsigned short sa[TRAIN ? WAVES : 1][H * 16];  // activations [feature][sample]
  __shared__ alignas(16) unsigned short sd[TRAIN ? WAVES : 1][H * 16];  // output gradients [feature][sample]
  __shared__ float sloss;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r = lane & 15, g = lane >> 4;
  stage(w, s, tid, BLOCK);
  if constexpr (TRAIN) {
    for (int i = tid; i < NPARAM; i += BLOCK) sg[i] = 0.0f;
    if (tid == 0) sloss = 0.0f;
  }
  __syncthreads();

  // gradient accumulators (TRAIN)
  f4 gW1[4], gW2[4][4], gW3[4][4];
  float gb[3][4][4], gw4[4][4], gb4 = 0.0f, lsum = 0.0f;
  if constexpr (TRAIN) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      gW1[a] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int b = 0; b < 4; b++) gW2[a][b] = gW3[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < 4; i++) {
        gw4[a][i] = 0.0f;
        gb[0][a][i] = gb[1][a][i] = gb[2][a][i] = 0.0f;
      }
    }
  }

  const long ntiles = (n + 15) / 16;
  const long stride = (long)gridDim.x * WAVES;
  const long iters = (ntiles + stride - 1) / stride;
  for (long it = 0; it < iters; it++) {
    const long tile = it * stride + (long)blockIdx.x * WAVES + wv;
    const long smp = tile * 16 + r;  // this lane's sample (column r)
    const bool valid = smp < n;
    const float* xs = x + (off + (valid ? smp : 0)) * NIN;
    Tile T;
    forward(s, T, [&](int f) { return valid ? xs[f] : 0.0f; }, lane);
    const float yv = softplusf(T.z4);
    if constexpr (!TRAIN) {
      if (valid && g == 0) y[off + smp] = yv;
    } else {
      const float tv = valid ? t[off + smp] : 0.0f;
      const float dz4 = valid ? 2.0f * (yv - tv) * inv_b * sigmoidf(T.z4) : 0.0f;
      if (valid && g == 0) lsum += (yv - tv) * (yv - tv);
      if (g == 0) gb4 += dz4;
      // layer 4 -> dz3 (C layout), dW4
      float dz[4][4];
#pragma unroll
      for (int rt = 0; rt < 4; rt++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int f = 16 * rt + 4 * g + i;
          gw4[rt][i] += dz4 * from_bf16((unsigned short)T.a[2][rt][i]);
          dz[rt][i] = s.w4[f] * dz4 * d_siluf(T.z[2][rt][i]);
        }
      // layers 3, 2, 1
#pragma unroll
      for (int l = 2; l >= 0; l--) {
        s4 dzb[4];
#pragma unroll
        for (int rt = 0; rt < 4; rt++) {
#pragma unroll
          for (int i = 0; i < 4; i++) gb[l][rt][i] += dz[rt][i];
          dzb[rt] = pack4(dz[rt][0], dz[rt][1], dz[rt][2], dz[rt][3]);
        }
        // stage bf16(dz_l) and a_{l-1} as [feature][sample] for the weight gradient
        __syncthreads();
#pragma unroll
        for (int rt = 0; rt < 4; rt++)
#pragma unroll
          for (int i = 0; i < 4; i++) sd[wv][(16 * rt + 4 * g + i) * 16 + r] = (unsigned short)dzb[rt][i];
        if (l > 0) {
#pragma unroll
          for (int rt = 0; rt < 4; rt++)
#pragma unroll
            for (int i = 0; i < 4; i++) sa[wv][(16 * rt + 4 * g + i) * 16 + r] = (unsigned short)T.a[l - 1][rt][i];
        } else {
#pragma unroll
          for (int i = 0; i < 4; i++) sa[wv][(4 * g + i) * 16 + r] = (unsigned short)T.a0[i];
        }
        __syncthreads();
        // dW_l[o][in] += sum_s dz[o][s] a[in][s]: A = dz (M = o, K = s), B = a^T (K = s, N = in)
#pragma unroll
        for (int ot = 0; ot < 4; ot++) {
          const s4 A = *reinterpret_cast<const s4*>(&sd[wv][(16 * ot + r) * 16 + 4 * g]);
          if (l == 0) {
            const s4 B = *reinterpret_cast<const s4*>(&sa[wv][r * 16 + 4 * g]);
            gW1[ot] = mfma(A, B, gW1[ot]);
          } else {
#pragma unroll
            for (int itl = 0; itl < 4; itl++) {
              const s4 B = *reinterpret_cast<const s4*>(&sa[wv][(16 * itl + r) * 16 + 4 * g]);
              if (l == 1) gW2[ot][itl] = mfma(A, B, gW2[ot][itl]);
              else gW3[ot][itl] = mfma(A, B, gW3[ot][itl]);
            }
          }
        }
        if (l > 0) {
          // da_{l-1} = W_l^T dz_l: A[m = in][k = out] = W_l[out][in], B = bf16(dz_l) tiles
          const unsigned short* W = l == 2 ? s.w3 : s.w2;
#pragma unroll
          for (int mt = 0; mt < 4; mt++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; kt++) {
              s4 A;
#pragma unroll
              for (int i = 0; i < 4; i++) A[i] = (short)W[(16 * kt + 4 * g + i) * H + 16 * mt + r];
              acc = mfma(A, dzb[kt], acc);
            }
#pragma unroll
            for (int i = 0; i < 4; i++) dz[mt][i] = acc[i] * d_siluf(T.z[l - 1][mt][i]);
          }
        }
      }
    }
  }

  if constexpr (TRAIN) {
    // block reduction of the partial gradients into LDS, then one row per block
#pragma unroll
    for (int a = 0; a < 4; a++) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int o = 16 * a + 4 * g + i;
        if (r < NIN) atomicAdd(&sg[OFF_W1 + o * NIN + r], gW1[a][i]);
#pragma unroll
        for (int b = 0; b < 4; b++) {
          atomicAdd(&sg[OFF_W2 + o * H + 16 * b + r], gW2[a][b][i]);
          atomicAdd(&sg[OFF_W3 + o * H + 16 * b + r], gW3[a][b][i]);
        }
        atomicAdd(&sg[OFF_B1 + o], gb[0][a][i]);
        atomicAdd(&sg[OFF_B2 + o], gb[1][a][i]);
        atomicAdd(&sg[OFF_B3 + o], gb[2][a][i]);
        atomicAdd(&sg[OFF_W4 + o], gw4[a][i]);
      }
    }
    if (g == 0) {
      atomicAdd(&sg[OFF_B4], gb4);
      atomicAdd(&sloss, lsum);
    }
    __syncthreads();
    for (int i = tid; i < NPARAM; i += BLOCK) part[(long)blockIdx.x * NPARAM + i] = sg[i];
    if (tid == 0) part_loss[blockIdx.x] = sloss;
  }
}

// sum of the block partials (gradient check)
__global__ void mlp_reduce(const float* __restrict__ part, int nblk, float* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float acc = 0.0f;
  for (int b = 0; b < nblk; b++) acc += part[(long)b * NPARAM + p];
  out[p] = acc;
}

__global__ void mlp_loss_reduce(const float* __restrict__ part_loss, int nblk, float scale, float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float acc = 0.0f;
  for (int b = 0; b < nblk; b++) acc += part_loss[b];
  *out += acc * scale;
}

// BF16 copies of the weights in the forward layouts
__device__ __forceinline__ void write_bf16(long p, float v, unsigned short* w1, unsigned short* w2,
                                           unsigned short* w3, float* w4) {
  if (p < OFF_B1) {
    w1[(p / NIN) * NPAD + p % NIN] = to_bf16(v);
  } else if (p >= OFF_W2 && p < OFF_B2) {
    w2[p - OFF_W2] = to_bf16(v);
  } else if (p >= OFF_W3 && p < OFF_B3) {
    w3[p - OFF_W3] = to_bf16(v);
  } else if (p >= OFF_W4 && p < OFF_B4) {
    w4[p - OFF_W4] = from_bf16(to_bf16(v));
  }
}

__global__ void mlp_adam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                         const float* __restrict__ part, int nblk, float lr, float c1, float c2,
                         unsigned short* w1, unsigned short* w2, unsigned short* w3, float* w4) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  float gr = 0.0f;
  for (int b = 0; b < nblk; b++) gr += part[(long)b * NPARAM + p];
  const float mm = 0.9f * m[p] + 0.1f * gr;
  const float vv = 0.999f * v[p] + 0.001f * gr * gr;
  m[p] = mm;
  v[p] = vv;
  const float np = params[p] - lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
  params[p] = np;
  write_bf16(p, np, w1, w2, w3, w4);
}

__global__ void mlp_refresh(const float* __restrict__ params, unsigned short* w1, unsigned short* w2,
                            unsigned short* w3, float* w4) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  write_bf16(p, params[p], w1, w2, w3, w4);
}

int grid_for(long n) {
  const long tiles = (n + 15) / 16;
  long b = (tiles + WAVES - 1) / WAVES;
  return (int)(b < 1 ? 1 : (b > 256 ? 256 : b));
}

hifr_status ensure(void** p, size_t& have, size_t need) {
  if (have >= need) return HIFR_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  have = 0;
  HIP_TRY(hipMalloc(p, need));
  have = need;
  return HIFR_OK;
}

hifr_status refresh(hifr_mlp* m) {
  hipLaunchKernelGGL(mlp_refresh, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->params, m->w1, m->w2,
                     m->w3, m->w4);
  HIP_TRY(hipGetLastError());
  return HIFR_OK;
}

Weights weights_of(const hifr_mlp* m) { return Weights{m->w1, m->w2, m->w3, m->w4, m->params}; }

}  // namespace

extern "C" {

hifr_status hifr_mlp_create(hifr_mlp** out, unsigned long long seed) {
  if (!out) return HIFR_ERR_ARG;
  *out = nullptr;
  if (!kLearnedEnabled) return HIFR_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return HIFR_ERR_DEVICE;
  hifr_mlp* m = new hifr_mlp();
  // He-uniform weights from a 64-bit linear congruential generator (Knuth's MMIX
  // constants), zero biases
  std::vector<float> p(NPARAM, 0.0f);
  uint64_t st = seed * 2862933555777941757ull + 3037000493ull;
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
  };
  for (int i = 0; i < H * NIN; i++) p[OFF_W1 + i] = (float)(std::sqrt(6.0 / NIN) * uni());
  for (int i = 0; i < H * H; i++) p[OFF_W2 + i] = (float)(std::sqrt(6.0 / H) * uni());
  for (int i = 0; i < H * H; i++) p[OFF_W3 + i] = (float)(std::sqrt(6.0 / H) * uni());
  for (int i = 0; i < H; i++) p[OFF_W4 + i] = (float)(std::sqrt(3.0 / H) * uni());
  hifr_status st_ = HIFR_OK;
  struct { void** ptr; size_t bytes; } al[] = {
      {(void**)&m->params, sizeof(float) * NPARAM}, {(void**)&m->adam_m, sizeof(float) * NPARAM},
      {(void**)&m->adam_v, sizeof(float) * NPARAM}, {(void**)&m->w1, sizeof(unsigned short) * H * NPAD},
      {(void**)&m->w2, sizeof(unsigned short) * H * H}, {(void**)&m->w3, sizeof(unsigned short) * H * H},
      {(void**)&m->w4, sizeof(float) * H}, {(void**)&m->scalar, sizeof(float) * 64}};
  for (auto& a : al) {
    if (hipMalloc(a.ptr, a.bytes) != hipSuccess) st_ = HIFR_ERR_DEVICE;
  }
  if (st_ == HIFR_OK && (hipMemset(m->w1, 0, sizeof(unsigned short) * H * NPAD) != hipSuccess ||
                         hipMemset(m->adam_m, 0, sizeof(float) * NPARAM) != hipSuccess ||
                         hipMemset(m->adam_v, 0, sizeof(float) * NPARAM) != hipSuccess ||
                         hipMemcpy(m->params, p.data(), sizeof(float) * NPARAM, hipMemcpyHostToDevice) != hipSuccess))
    st_ = HIFR_ERR_DEVICE;
  if (st_ == HIFR_OK) st_ = refresh(m);
  if (st_ == HIFR_OK && hipStreamSynchronize(m->stream) != hipSuccess) st_ = HIFR_ERR_DEVICE;
  if (st_ != HIFR_OK) {
    hifr_mlp_destroy(m);
    return st_;
  }
  for (int i = 0; i < NIN; i++) {
    m->norm.mean[i] = 0.0f;
    m->norm.scale[i] = 1.0f;
    m->norm.lo[i] = -INFINITY;
    m->norm.hi[i] = INFINITY;
  }
  m->norm.out_scale = 1.0f;
  *out = m;
  return HIFR_OK;
}

hifr_status hifr_mlp_destroy(hifr_mlp* m) {
  if (!m) return HIFR_OK;
  void* ps[] = {m->params, m->adam_m, m->adam_v, m->w1, m->w2, m->w3, m->w4, m->scalar, m->part, m->part_loss,
                m->x, m->t, m->bx, m->bt, m->by};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  delete m;
  return HIFR_OK;
}

hifr_status hifr_mlp_set_stream(hifr_mlp* m, void* s) {
  if (!m) return HIFR_ERR_ARG;
  m->stream = (hipStream_t)s;
  return HIFR_OK;
}

hifr_status hifr_mlp_set_params(hifr_mlp* m, const float* params) {
  if (!m || !params) return HIFR_ERR_ARG;
  for (int i = 0; i < NPARAM; i++)
    if (!std::isfinite(params[i])) return HIFR_ERR_ARG;
  HIP_TRY(hipMemcpyAsync(m->params, params, sizeof(float) * NPARAM, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_m, 0, sizeof(float) * NPARAM, m->stream));
  HIP_TRY(hipMemsetAsync(m->adam_v, 0, sizeof(float) * NPARAM, m->stream));
  m->step = 0;
  hifr_status st = refresh(m);
  if (st != HIFR_OK) return st;
  HIP_TRY(hipStreamSynchronize(m->stream));
  return HIFR_OK;
}

hifr_status hifr_mlp_get_params(const hifr_mlp* m, float* params) {
  if (!m || !params) return HIFR_ERR_ARG;
  HIP_TRY(hipMemcpyAsync(params, m->params, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return HIFR_OK;
}

hifr_status hifr_mlp_set_normalization(hifr_mlp* m, const float* mean, const float* scale, const float* lo,
                                       const float* hi, float out_scale) {
  if (!m || !mean || !scale || !lo || !hi || !(out_scale > 0.0f)) return HIFR_ERR_ARG;
  for (int i = 0; i < NIN; i++) {
    if (!(scale[i] > 0.0f) || !std::isfinite(mean[i]) || !(lo[i] <= hi[i])) return HIFR_ERR_ARG;
    m->norm.mean[i] = mean[i];
    m->norm.scale[i] = scale[i];
    m->norm.lo[i] = lo[i];
    m->norm.hi[i] = hi[i];
  }
  m->norm.out_scale = out_scale;
  return HIFR_OK;
}

hifr_status hifr_mlp_loss_grad(hifr_mlp* m, long n, const float* x, const float* t, float* loss, float* grad) {
  if (!m || n <= 0 || !x || !t || !loss || !grad) return HIFR_ERR_ARG;
  hifr_status st;
  if ((st = ensure((void**)&m->bx, m->bx_bytes, sizeof(float) * NIN * n)) != HIFR_OK) return st;
  if ((st = ensure((void**)&m->bt, m->bt_bytes, sizeof(float) * n)) != HIFR_OK) return st;
  const int nblk = grid_for(n);
  if ((st = ensure((void**)&m->part, m->part_bytes, sizeof(float) * NPARAM * (nblk + 1))) != HIFR_OK) return st;
  if ((st = ensure((void**)&m->part_loss, m->part_loss_bytes, sizeof(float) * 256)) != HIFR_OK) return st;
  HIP_TRY(hipMemcpyAsync(m->bx, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->bt, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_kernel<true>, dim3(nblk), dim3(BLOCK), 0, m->stream, weights_of(m), m->bx, m->bt, 0L, n,
                     1.0f / (float)n, m->part, m->part_loss, (float*)nullptr);
  HIP_TRY(hipGetLastError());
  float* gout = m->part + (size_t)NPARAM * nblk;
  hipLaunchKernelGGL(mlp_reduce, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->part, nblk, gout);
  HIP_TRY(hipMemsetAsync(m->scalar, 0, sizeof(float), m->stream));
  hipLaunchKernelGGL(mlp_loss_reduce, dim3(1), dim3(64), 0, m->stream, m->part_loss, nblk, 1.0f / (float)n,
                     m->scalar);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(grad, gout, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipMemcpyAsync(loss, m->scalar, sizeof(float), hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return HIFR_OK;
}

hifr_status hifr_mlp_set_data(hifr_mlp* m, long n, const float* x, const float* t) {
  if (!m || n <= 0 || !x || !t) return HIFR_ERR_ARG;
  hifr_status st;
  if ((st = ensure((void**)&m->x, m->x_bytes, sizeof(float) * NIN * n)) != HIFR_OK) return st;
  if ((st = ensure((void**)&m->t, m->t_bytes, sizeof(float) * n)) != HIFR_OK) return st;
  HIP_TRY(hipMemcpyAsync(m->x, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipMemcpyAsync(m->t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  m->n = n;
  return HIFR_OK;
}

hifr_status hifr_mlp_train(hifr_mlp* m, int epochs, int batch, float lr, float* loss_per_epoch) {
  if (!m || epochs <= 0 || batch < 16 || !(lr > 0.0f)) return HIFR_ERR_ARG;
  if (m->n <= 0) return HIFR_ERR_STATE;
  const int nblk = grid_for(batch);
  hifr_status st;
  if ((st = ensure((void**)&m->part, m->part_bytes, sizeof(float) * NPARAM * (nblk + 1))) != HIFR_OK) return st;
  if ((st = ensure((void**)&m->part_loss, m->part_loss_bytes, sizeof(float) * 256)) != HIFR_OK) return st;
  if (epochs > 63) return HIFR_ERR_ARG;  // epoch losses live in the 64-float scratch
  HIP_TRY(hipMemsetAsync(m->scalar, 0, sizeof(float) * 64, m->stream));
  const long nb = (m->n + batch - 1) / batch;
  for (int e = 0; e < epochs; e++) {
    for (long b = 0; b < nb; b++) {
      const long off = b * batch, cnt = off + batch <= m->n ? batch : m->n - off;
      const int gb = grid_for(cnt);
      hipLaunchKernelGGL(mlp_kernel<true>, dim3(gb), dim3(BLOCK), 0, m->stream, weights_of(m), m->x, m->t, off,
                         cnt, 1.0f / (float)cnt, m->part, m->part_loss, (float*)nullptr);
      m->step++;
      const float c1 = 1.0f - std::pow(0.9f, (float)m->step), c2 = 1.0f - std::pow(0.999f, (float)m->step);
      hipLaunchKernelGGL(mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, m->stream, m->params, m->adam_m,
                         m->adam_v, m->part, gb, lr, c1, c2, m->w1, m->w2, m->w3, m->w4);
      hipLaunchKernelGGL(mlp_loss_reduce, dim3(1), dim3(64), 0, m->stream, m->part_loss, gb,
                         1.0f / (float)(cnt * nb), m->scalar + e);
    }
    HIP_TRY(hipGetLastError());
  }
  if (loss_per_epoch)
    HIP_TRY(hipMemcpyAsync(loss_per_epoch, m->scalar, sizeof(float) * epochs, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return HIFR_OK;
}

hifr_status hifr_mlp_predict(hifr_mlp* m, long n, const float* x, float* y) {
  if (!m || n <= 0 || !x || !y) return HIFR_ERR_ARG;
  hifr_status st;
  if ((st = ensure((void**)&m->bx, m->bx_bytes, sizeof(float) * NIN * n)) != HIFR_OK) return st;
  if ((st = ensure((void**)&m->by, m->by_bytes, sizeof(float) * n)) != HIFR_OK) return st;
  HIP_TRY(hipMemcpyAsync(m->bx, x, sizeof(float) * NIN * n, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(mlp_kernel<false>, dim3(grid_for(n)), dim3(BLOCK), 0, m->stream, weights_of(m), m->bx,
                     (const float*)nullptr, 0L, n, 0.0f, (float*)nullptr, (float*)nullptr, m->by);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(y, m->by, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
  HIP_TRY(hipStreamSynchronize(m->stream));
  return HIFR_OK;
}

}  // extern "C"
