// mlp.hip -- the learned search-radius model: 27 -> 64 -> 64 -> 2 (ReLU),
// BF16 weights and activations on the MFMA units, FP32 accumulation.
//
// Features of a cell: log(1 + count) of the 27 grid cells around the site's
// grid cell (0 outside the grid), padded to 32.  Outputs (standardised):
// log(final R / smallest grid-cell edge) and log(1 + candidates streamed),
// as recorded by the plain kernel.
//
// Matrix products use v_mfma_f32_16x16x32_bf16: per wavefront
//   A (16 x 32): lane l holds row l&15, k = 8(l>>4) .. 8(l>>4)+7
//   B (32 x 16): lane l holds col l&15, k = 8(l>>4) .. 8(l>>4)+7
//   D (16 x 16): lane l holds col l&15, rows 4(l>>4) .. 4(l>>4)+3
// (checked by k_mfma_probe against a host product).
//
// Training: one wavefront per 32 samples runs forward and backward; its
// weight gradients go to a per-wavefront slot and k_adam reduces the slots in
// a fixed order (deterministic) and applies Adam to the FP32 master weights,
// refreshing the BF16 copies.
#include "launch.h"
#include "tess_internal.h"

namespace tess {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// round-to-nearest-even f32 -> bf16 bits
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
  bf16x8 v;
  __builtin_memcpy(&v, p, 16);
  return v;
}
__device__ __forceinline__ bf16x8 gather8(const uint16_t* p, int stride) {
  uint16_t t[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) t[j] = p[j * stride];
  bf16x8 v;
  __builtin_memcpy(&v, t, 16);
  return v;
}

}  // namespace

// Lane-layout probe: D = A (16x32) * B (32x16), A and B given row-major.
__global__ __launch_bounds__(64) void k_mfma_probe(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                                                   float* __restrict__ D) {
  const int l = threadIdx.x, row = l & 15, kb = 8 * (l >> 4);
  const bf16x8 a = load8(A + row * 32 + kb);
  const bf16x8 b = gather8(B + kb * 16 + row, 16);
  f32x4 c = {0.f, 0.f, 0.f, 0.f};
  c = mfma(a, b, c);
#pragma unroll
  for (int i = 0; i < 4; ++i) D[(4 * (l >> 4) + i) * 16 + (l & 15)] = c[i];
}

// Features: log(1 + count) of the 3x3x3 grid cells around each point's grid
// cell (offsets -1..1, x outermost), 0 outside the grid; bf16, 32 per point
// (the last 5 are 0).  bins: grid cell of each point (original order).  The
// values of counts below kFeatTab are computed once per block into LDS by
// the same expression (so bit for bit the same) and looked up.
constexpr int kFeatTab = 1024;
__global__ __launch_bounds__(256) void k_features(const uint32_t* __restrict__ bins,
                                                  const uint32_t* __restrict__ delim, int cpd, int n,
                                                  uint16_t* __restrict__ feat) {
  __shared__ uint16_t tab[kFeatTab];
  for (int c = threadIdx.x; c < kFeatTab; c += blockDim.x) tab[c] = to_bf16((float)log1p((double)c));
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b0 = (int)bins[i];
  const int iz = b0 % cpd, iy = (b0 / cpd) % cpd, ix = b0 / (cpd * cpd);
  uint16_t f[kMlpIn];
#pragma unroll
  for (int j = 0; j < kMlpIn; ++j) {
    uint32_t c = 0;
    if (j < kMlpFeatures) {
      const int bx = ix + j / 9 - 1, by = iy + (j / 3) % 3 - 1, bz = iz + j % 3 - 1;
      if (bx >= 0 && bx < cpd && by >= 0 && by < cpd && bz >= 0 && bz < cpd) {
        const int b = (bx * cpd + by) * cpd + bz;
        c = delim[b + 1] - delim[b];
      }
      f[j] = c < (uint32_t)kFeatTab ? tab[c] : to_bf16((float)log1p((double)c));
    } else {
      f[j] = 0;
    }
  }
  uint4* dst = reinterpret_cast<uint4*>(feat + (size_t)i * kMlpIn);
  const uint4* src = reinterpret_cast<const uint4*>(f);
#pragma unroll
  for (int q = 0; q < kMlpIn / 8; ++q) dst[q] = src[q];
}

// Forward of 16 samples per wavefront (inference).  W: bf16 weights laid out
// as MlpWeights; out: 2 floats per sample (standardised).
__global__ __launch_bounds__(64) void k_mlp_forward(const uint16_t* __restrict__ X, int n,
                                                    const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                                    float* __restrict__ out) {
  __shared__ __attribute__((aligned(16))) uint16_t h[16 * kMlpH];
  const int l = threadIdx.x, row = l & 15, kb = 8 * (l >> 4), r0 = 4 * (l >> 4);
  const int s0 = blockIdx.x * 16;
  const uint16_t *W1 = W, *W2 = W + kMlpIn * kMlpH, *W3 = W2 + kMlpH * kMlpH;
  const float *b1 = bias, *b2 = bias + kMlpH, *b3 = b2 + kMlpH;
  // layer 1: [16 x 32] x [32 x 64]
  bf16x8 a = {};
  if (s0 + row < n) a = load8(X + (size_t)(s0 + row) * kMlpIn + kb);
  f32x4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
    acc[t] = mfma(a, gather8(W1 + kb * kMlpH + 16 * t + row, kMlpH), acc[t]);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float z = acc[t][i] + b1[16 * t + row];
      h[(r0 + i) * kMlpH + 16 * t + row] = to_bf16(z > 0.f ? z : 0.f);
    }
  __syncthreads();
  // layer 2: [16 x 64] x [64 x 64]
#pragma unroll
  for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 2; ++kk) {
    const bf16x8 ah = load8(h + row * kMlpH + 32 * kk + kb);
#pragma unroll
    for (int t = 0; t < 4; ++t)
      acc[t] = mfma(ah, gather8(W2 + (32 * kk + kb) * kMlpH + 16 * t + row, kMlpH), acc[t]);
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float z = acc[t][i] + b2[16 * t + row];
      h[(r0 + i) * kMlpH + 16 * t + row] = to_bf16(z > 0.f ? z : 0.f);
    }
  __syncthreads();
  // layer 3: [16 x 64] x [64 x 16] (2 of 16 columns used)
  f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < 2; ++kk)
    y = mfma(load8(h + row * kMlpH + 32 * kk + kb), gather8(W3 + (32 * kk + kb) * kMlpOutPad + row, kMlpOutPad), y);
  if (row < kMlpOut)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (s0 + r0 + i < n) out[(size_t)(s0 + r0 + i) * kMlpOut + row] = y[i] + b3[row];
}

// Forward + backward of 32 samples per wavefront; mean-squared error over
// the whole batch (scale = 1 / batch).  Gradients of this wavefront's samples
// go to grad + blockIdx.x * kMlpParams.
__global__ __launch_bounds__(64) void k_mlp_train_step(const uint16_t* __restrict__ X, const float* __restrict__ T,
                                                       const int* __restrict__ idx, int nsamp, float scale,
                                                       const uint16_t* __restrict__ W,
                                                       const float* __restrict__ bias, float* __restrict__ grad,
                                                       float* __restrict__ loss) {
  __shared__ __attribute__((aligned(16))) uint16_t xs[32 * kMlpIn];
  __shared__ __attribute__((aligned(16))) uint16_t h1[32 * kMlpH], h2[32 * kMlpH], dz[32 * kMlpH];
  __shared__ float z1[32 * kMlpH], z2[32 * kMlpH], dy[32 * kMlpOut];
  const int l = threadIdx.x, row = l & 15, kb = 8 * (l >> 4), r0 = 4 * (l >> 4);
  const int s0 = blockIdx.x * 32;
  const uint16_t *W1 = W, *W2 = W + kMlpIn * kMlpH, *W3 = W2 + kMlpH * kMlpH;
  const float *b1 = bias, *b2 = bias + kMlpH, *b3 = b2 + kMlpH;
  float* g = grad + (size_t)blockIdx.x * kMlpParams;
  float *gW1 = g, *gW2 = g + kMlpIn * kMlpH, *gW3 = gW2 + kMlpH * kMlpH, *gb1 = gW3 + kMlpH * kMlpOutPad,
        *gb2 = gb1 + kMlpH, *gb3 = gb2 + kMlpH;
  // stage the 32 samples' features (zero rows past the end)
  for (int e = l; e < 32 * kMlpIn; e += 64) {
    const int s = e / kMlpIn, c = e % kMlpIn;
    xs[e] = (s0 + s < nsamp) ? X[(size_t)idx[s0 + s] * kMlpIn + c] : (uint16_t)0;
  }
  __syncthreads();
  // forward, two halves of 16 rows
  for (int half = 0; half < 2; ++half) {
    const int rb = 16 * half;
    f32x4 acc[4];
    const bf16x8 a = load8(xs + (rb + row) * kMlpIn + kb);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
      acc[t] = mfma(a, gather8(W1 + kb * kMlpH + 16 * t + row, kMlpH), acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = (rb + r0 + i) * kMlpH + 16 * t + row;
        const float z = acc[t][i] + b1[16 * t + row];
        z1[e] = z;
        h1[e] = to_bf16(z > 0.f ? z : 0.f);
      }
  }
  __syncthreads();
  for (int half = 0; half < 2; ++half) {
    const int rb = 16 * half;
    f32x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
      const bf16x8 ah = load8(h1 + (rb + row) * kMlpH + 32 * kk + kb);
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[t] = mfma(ah, gather8(W2 + (32 * kk + kb) * kMlpH + 16 * t + row, kMlpH), acc[t]);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = (rb + r0 + i) * kMlpH + 16 * t + row;
        const float z = acc[t][i] + b2[16 * t + row];
        z2[e] = z;
        h2[e] = to_bf16(z > 0.f ? z : 0.f);
      }
  }
  __syncthreads();
  // output layer and dL/dy
  float lsum = 0.f;
  for (int half = 0; half < 2; ++half) {
    const int rb = 16 * half;
    f32x4 y = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
      y = mfma(load8(h2 + (rb + row) * kMlpH + 32 * kk + kb),
               gather8(W3 + (32 * kk + kb) * kMlpOutPad + row, kMlpOutPad), y);
    if (row < kMlpOut)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int s = rb + r0 + i;
        float d = 0.f;
        if (s0 + s < nsamp) {
          const float diff = y[i] + b3[row] - T[(size_t)idx[s0 + s] * kMlpOut + row];
          d = diff * scale;
          lsum += 0.5f * diff * diff;
        }
        dy[s * kMlpOut + row] = d;
      }
  }
  __syncthreads();
  // dW3 = h2^T dy, db3 = sum dy (VALU: 64 x 2 outputs over 32 samples)
  for (int e = l; e < kMlpH * kMlpOutPad; e += 64) {
    const int hi = e / kMlpOutPad, o = e % kMlpOutPad;
    float s = 0.f;
    if (o < kMlpOut)
      for (int q = 0; q < 32; ++q) s += from_bf16(h2[q * kMlpH + hi]) * dy[q * kMlpOut + o];
    gW3[e] = s;
  }
  if (l < kMlpOutPad) {
    float s = 0.f;
    if (l < kMlpOut)
      for (int q = 0; q < 32; ++q) s += dy[q * kMlpOut + l];
    gb3[l] = s;
  }
  // dz2 = (dy W3^T) * [z2 > 0]
  for (int e = l; e < 32 * kMlpH; e += 64) {
    const int s = e / kMlpH, hi = e % kMlpH;
    float d = 0.f;
#pragma unroll
    for (int o = 0; o < kMlpOut; ++o) d += dy[s * kMlpOut + o] * from_bf16(W3[hi * kMlpOutPad + o]);
    dz[e] = to_bf16(z2[e] > 0.f ? d : 0.f);
  }
  __syncthreads();
  // dW2 = h1^T dz2 (K = 32 samples), db2 = sum dz2
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
      c = mfma(gather8(h1 + kb * kMlpH + 16 * m + row, kMlpH), gather8(dz + kb * kMlpH + 16 * t + row, kMlpH), c);
#pragma unroll
      for (int i = 0; i < 4; ++i) gW2[(16 * m + r0 + i) * kMlpH + 16 * t + row] = c[i];
    }
  if (l < kMlpH) {
    float s = 0.f;
    for (int q = 0; q < 32; ++q) s += from_bf16(dz[q * kMlpH + l]);
    gb2[l] = s;
  }
  // dz1 = (dz2 W2^T) * [z1 > 0]  (K = 64)
  f32x4 dh[2][4];
#pragma unroll
  for (int half = 0; half < 2; ++half)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      dh[half][t] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < 2; ++kk)
        dh[half][t] = mfma(load8(dz + (16 * half + row) * kMlpH + 32 * kk + kb),
                           load8(W2 + (16 * t + row) * kMlpH + 32 * kk + kb), dh[half][t]);
    }
  __syncthreads();
#pragma unroll
  for (int half = 0; half < 2; ++half)
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int e = (16 * half + r0 + i) * kMlpH + 16 * t + row;
        dz[e] = to_bf16(z1[e] > 0.f ? dh[half][t][i] : 0.f);
      }
  __syncthreads();
  // dW1 = x^T dz1 (K = 32 samples), db1 = sum dz1
#pragma unroll
  for (int m = 0; m < 2; ++m)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      f32x4 c = {0.f, 0.f, 0.f, 0.f};
      c = mfma(gather8(xs + kb * kMlpIn + 16 * m + row, kMlpIn), gather8(dz + kb * kMlpH + 16 * t + row, kMlpH), c);
#pragma unroll
      for (int i = 0; i < 4; ++i) gW1[(16 * m + r0 + i) * kMlpH + 16 * t + row] = c[i];
    }
  if (l < kMlpH) {
    float s = 0.f;
    for (int q = 0; q < 32; ++q) s += from_bf16(dz[q * kMlpH + l]);
    gb1[l] = s;
  }
  // per-wavefront loss
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  if (l == 0) loss[blockIdx.x] = lsum;
}

// Adam on the FP32 master parameters; gradients are the sum of nslots
// per-wavefront slots, added in slot order.  Refreshes the bf16 weights.
__global__ void k_adam(const float* __restrict__ grad, int nslots, float* __restrict__ master,
                       float* __restrict__ m1, float* __restrict__ m2, uint16_t* __restrict__ wbf,
                       float* __restrict__ bias, float lr, float b1c, float b2c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  float gsum = 0.f;
  for (int s = 0; s < nslots; ++s) gsum += grad[(size_t)s * kMlpParams + i];
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  m1[i] = beta1 * m1[i] + (1.f - beta1) * gsum;
  m2[i] = beta2 * m2[i] + (1.f - beta2) * gsum * gsum;
  const float mh = m1[i] / b1c, vh = m2[i] / b2c;
  const float w = master[i] - lr * mh / (sqrtf(vh) + eps);
  master[i] = w;
  if (i < kMlpWeights)
    wbf[i] = to_bf16(w);
  else
    bias[i - kMlpWeights] = w;
}

// Training targets from the plain search's counters: log(R / s_min) with R
// the final security radius, and log(1 + classified candidates); vmax: the
// largest vertex count the cell reached (at least 65 when the cell had to be
// rerun at a larger capacity: its cuts ran out of the 64 slots).
__global__ void k_targets(const tess_cell_stats* __restrict__ st, int n, double smin, float* __restrict__ T,
                          uint16_t* __restrict__ vmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const tess_cell_stats c = st[i];
  T[2 * i] = (float)log(sqrt(c.r2) / smin);
  T[2 * i + 1] = (float)log1p((double)c.classified);
  uint32_t v = c.max_vertices < 65535u ? c.max_vertices : 65535u;
  if ((c.status & TESS_CELL_RERUN) && v < 65u) v = 65u;
  vmax[i] = (uint16_t)v;
}

// Sum of nslots gradient slots in slot order.
__global__ void k_grad_sum(const float* __restrict__ grad, int nslots, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  float g = 0.f;
  for (int s = 0; s < nslots; ++s) g += grad[(size_t)s * kMlpParams + i];
  out[i] = g;
}

// Standardise targets: Tn = (T - mean) / std per column.
__global__ void k_standardise(const float* __restrict__ T, size_t n, float m0, float s0, float m1, float s1,
                              float* __restrict__ Tn) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  Tn[2 * i] = (T[2 * i] - m0) / s0;
  Tn[2 * i + 1] = (T[2 * i + 1] - m1) / s1;
}

// bf16 copy of the weight part of the master parameters; biases to fp32.
__global__ void k_weights_from_master(const float* __restrict__ master, uint16_t* __restrict__ wbf,
                                      float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  if (i < kMlpWeights)
    wbf[i] = to_bf16(master[i]);
  else
    bias[i - kMlpWeights] = master[i];
}

hipError_t launch_targets(const tess_cell_stats* st, int n, double smin, float* T, uint16_t* vmax, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_targets, dim3((n + 255) / 256), dim3(256), 0, s, st, n, smin, T, vmax);
  return hipGetLastError();
}

hipError_t launch_grad_sum(const float* grad, int nslots, float* out, hipStream_t s) {
  hipLaunchKernelGGL(k_grad_sum, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, grad, nslots, out);
  return hipGetLastError();
}

hipError_t launch_standardise(const float* T, size_t n, const float norm[4], float* Tn, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_standardise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, T, n, norm[0], norm[1],
                     norm[2], norm[3], Tn);
  return hipGetLastError();
}

hipError_t launch_weights_from_master(const float* master, uint16_t* wbf, float* bias, hipStream_t s) {
  hipLaunchKernelGGL(k_weights_from_master, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, master, wbf, bias);
  return hipGetLastError();
}

hipError_t launch_mfma_probe(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s) {
  hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D);
  return hipGetLastError();
}

hipError_t launch_features(const uint32_t* bins, const uint32_t* delim, int cpd, int n, uint16_t* feat,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_features, dim3((n + 255) / 256), dim3(256), 0, s, bins, delim, cpd, n, feat);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const uint16_t* X, int n, const uint16_t* W, const float* bias, float* out,
                              hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 15) / 16), dim3(64), 0, s, X, n, W, bias, out);
  return hipGetLastError();
}

hipError_t launch_mlp_train_step(const uint16_t* X, const float* T, const int* idx, int nsamp, float scale,
                                 const uint16_t* W, const float* bias, float* grad, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_train_step, dim3((nsamp + 31) / 32), dim3(64), 0, s, X, T, idx, nsamp, scale, W, bias,
                     grad, loss);
  return hipGetLastError();
}

hipError_t launch_adam(const float* grad, int nslots, float* master, float* m1, float* m2, uint16_t* wbf,
                       float* bias, float lr, float b1c, float b2c, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, grad, nslots, master, m1, m2, wbf, bias,
                     lr, b1c, b2c);
  return hipGetLastError();
}

}  // namespace tess
