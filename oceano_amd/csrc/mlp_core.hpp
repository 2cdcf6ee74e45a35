// mlp_core.hpp -- the closure network on BF16 matrix cores: shapes, parameter layout and the
// wave-level building blocks shared by the training kernels (mlp.hip) and the in-stage
// inference (dg_closure.hpp).
//
// Network: x (IN, zero padded) -> 3 x [64, GELU] -> y (OUT, zero padded), all products on
// the BF16 matrix cores (BF16 operands, FP32 accumulation), biases, GELU, loss and Adam in
// FP32.  Inputs are normalised, (x - mu) / sigma, and outputs de-normalised,
// y * scale + shift, in FP32 by the kernels.  Products whose contraction is 32 or more deep
// (every forward layer, the backward products dY W^T) use gfx950's
// v_mfma_f32_16x16x32_bf16; the weight gradients, contracted over a 16-row tile of
// samples, use v_mfma_f32_16x16x16_bf16.
//
// MFMA operand layouts (one wave, lane l, r = l % 16, g = l / 16):
//   16x16x16: A (16 x 16, i x k) lane l supplies A[r][4 g + j], j = 0..3; B (k x n) supplies
//             B[4 g + j][r]
//   16x16x32: A (16 x 32) lane l supplies A[r][8 g + j], j = 0..7; B (32 x 16) supplies
//             B[8 g + j][r]
//   D (16 x 16, FP32, both):            lane l holds D[4 g + j][r], j = 0..3
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace oceano {
namespace mlp {

constexpr int IN = 32, HID = 64, OUT = 32;
// FP32 master parameters, W_l row-major [fan_in][fan_out]
constexpr int W1 = 0, B1 = W1 + IN * HID, W2 = B1 + HID, B2 = W2 + HID * HID, W3 = B2 + HID,
              B3 = W3 + HID * HID, W4 = B3 + HID, B4 = W4 + HID * OUT, NPARAM = B4 + OUT;
// BF16 transposed weights [fan_out][fan_in]: the B operand of the forward products
constexpr int T1 = 0, T2 = T1 + HID * IN, T3 = T2 + HID * HID, T4 = T3 + HID * HID,
              NT = T4 + OUT * HID;
// normalisation block: mu[IN], inv_sigma[IN], scale[OUT], shift[OUT], lo[OUT], hi[OUT]
// (inference saturates the normalised output to [lo, hi]; training does not)
constexpr int NNORM = 2 * IN + 4 * OUT;
constexpr int NLO = 2 * IN + 2 * OUT, NHI = 2 * IN + 3 * OUT;

// de-normalised network output of column o at inference
__device__ inline float output_value(float yn, const float *norm, int o)
{
  yn = fminf(fmaxf(yn, norm[NLO + o]), norm[NHI + o]);
  return yn * norm[2 * IN + o] + norm[2 * IN + OUT + o];
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ inline bf16 to_bf16(float f)
{
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(u >> 16);
}

__host__ __device__ inline float from_bf16(bf16 h)
{
  const uint32_t u = (uint32_t)h << 16;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// GELU, erf form, and its derivative
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ inline float gelu_prime(float x)
{
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
         x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ inline f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// the 8 consecutive BF16 values at p (16-byte aligned) as an MFMA operand
__device__ inline bf16x8 load8(const bf16 *p) { return *(const bf16x8 *)p; }

// acc[ct] = A[16 x K] * W for the N / 16 column tiles (K a multiple of 32); A row-major with
// leading dimension K in LDS, W given transposed (WT[n][k], leading dimension K).  The k
// order is fixed, so a row's result does not depend on the other rows of the tile.
template <int K, int N>
__device__ inline void dense_tile(const bf16 *A, const bf16 *WT, f32x4 (&acc)[N / 16])
{
  static_assert(K % 32 == 0, "contraction depth must be a multiple of 32");
  const int l = threadIdx.x & 63, r = l & 15, k8 = (l >> 4) * 8;
#pragma unroll
  for (int ct = 0; ct < N / 16; ++ct)
    acc[ct] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kk = 0; kk < K / 32; ++kk)
    {
      const bf16x8 a = load8(A + r * K + 32 * kk + k8);
#pragma unroll
      for (int ct = 0; ct < N / 16; ++ct)
        {
          const bf16x8 b = load8(WT + (16 * ct + r) * K + 32 * kk + k8);
          acc[ct] = mfma32(a, b, acc[ct]);
        }
    }
}

// hidden layer epilogue: z = acc + bias, H = bf16(gelu(z)) (and z kept when Z != nullptr)
template <int N>
__device__ inline void hidden_epilogue(const f32x4 (&acc)[N / 16], const float *bias, bf16 *H,
                                       float *Z)
{
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
#pragma unroll
  for (int ct = 0; ct < N / 16; ++ct)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      {
        const int row = kq + j, col = 16 * ct + r;
        const float z = acc[ct][j] + bias[col];
        if (Z)
          Z[row * N + col] = z;
        H[row * N + col] = to_bf16(gelu(z));
      }
}

// forward of one 16-row tile whose normalised BF16 inputs are in X (16 x IN): Y (16 x OUT)
// gets the network output (before de-normalisation); H0..H2 (16 x HID) the BF16 hidden
// activations, Z (3 x 16 x HID, may be null) the pre-activations.  wt: transposed BF16
// weights (LDS), p: FP32 parameters (for the biases).
__device__ inline void forward_tile(const bf16 *X, const bf16 *wt, const float *p, bf16 *H0,
                                    bf16 *H1, bf16 *H2, float *Z, float *Y)
{
  f32x4 acc[HID / 16];
  dense_tile<IN, HID>(X, wt + T1, acc);
  hidden_epilogue<HID>(acc, p + B1, H0, Z);
  __syncthreads();
  dense_tile<HID, HID>(H0, wt + T2, acc);
  hidden_epilogue<HID>(acc, p + B2, H1, Z ? Z + 16 * HID : nullptr);
  __syncthreads();
  dense_tile<HID, HID>(H1, wt + T3, acc);
  hidden_epilogue<HID>(acc, p + B3, H2, Z ? Z + 32 * HID : nullptr);
  __syncthreads();
  f32x4 out[OUT / 16];
  dense_tile<HID, OUT>(H2, wt + T4, out);
  const int l = threadIdx.x & 63, r = l & 15, kq = (l >> 4) * 4;
#pragma unroll
  for (int ct = 0; ct < OUT / 16; ++ct)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      Y[(kq + j) * OUT + 16 * ct + r] = out[ct][j] + p[B4 + 16 * ct + r];
  __syncthreads();
}

}  // namespace mlp
}  // namespace oceano
