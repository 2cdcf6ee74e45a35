// dg_mfma.hpp -- dense per-cell matrix products of the FP32 stage on the matrix cores.
//
// Two operators of the mixed-precision stage are a fixed dense matrix applied to every cell:
// the closed-form depth increment delta = J r / area (J = sum_{k<15} (I - D^-1 M)^k D^-1 of
// the unit cell, the 15 capped Jacobi sweeps of OceanDG.h:1443-1504 while the depth floor
// cannot bind) and the momentum inverse mass M^-1 = K / area (K = B1 (x) B1, the basis
// change of CellwiseInverseMassMatrix, OceanDG.h:1541-1553, written as one matrix).  Both
// are nd x nd with nd = (p+1)^2 and neither factors further (J is dense); at p >= 3 the
// product runs on v_mfma_f32_16x16x4_f32: the 64 cells of the wave (times the number of
// vectors per cell) are the N dimension, the matrix is the A operand, and the vectors travel
// between the one-cell-per-lane layout and the MFMA operand layouts through LDS.
//
// v_mfma_f32_16x16x4_f32 (one block): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16]
// of a 16x4 by 4x16 product and holds D[4 (l / 16) + e][l % 16], e = 0..3, of the 16x16
// accumulator.
#pragma once

#include "dg_kernels.hpp"

namespace oceano {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// LDS floats wave_matvec needs for NV vectors of ND entries per cell
template <int ND, int NV>
__host__ __device__ constexpr int matvec_lds_words()
{
  return (((ND + 15) / 16) * 16 > ((ND + 3) / 4) * 4 ? ((ND + 15) / 16) * 16 : ((ND + 3) / 4) * 4) *
         (CB * NV + 16);
}

// y[v] = A x[v] for v < NV, for every cell of the wave (lane = cell); A is ND x ND,
// row-major, in global memory.  `lds` holds matvec_lds_words<ND, NV>() floats that no other
// wave uses.  Rows of the staged vectors are padded by 16 words so that the four row groups
// a B-operand read touches fall on different banks.
template <int ND, int NV>
__device__ __forceinline__ void wave_matvec(const float *__restrict__ A, const float (&x)[NV][ND],
                                            float (&y)[NV][ND], float *lds)
{
  constexpr int MT = (ND + 15) / 16;  // 16-row tiles of A
  constexpr int KS = (ND + 3) / 4;    // 4-deep steps of the contraction
  constexpr int NT = CB * NV / 16;    // 16-column tiles: cells x vectors
  constexpr int LD = CB * NV + 16;    // LDS row stride (floats)
  const int lane = threadIdx.x & (CB - 1);
  const int r16 = lane & 15, g4 = lane >> 4;
  // the vectors as rows k, columns v * 64 + cell; rows past ND are zero
#pragma unroll
  for (int v = 0; v < NV; ++v)
    {
#pragma unroll
      for (int k = 0; k < ND; ++k)
        lds[k * LD + v * CB + lane] = x[v][k];
#pragma unroll
      for (int k = ND; k < KS * 4; ++k)
        lds[k * LD + v * CB + lane] = 0.f;
    }
  // A operands of every (row tile, step), zero outside the matrix
  float a[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int ks = 0; ks < KS; ++ks)
      {
        const int row = 16 * mt + r16, col = 4 * ks + g4;
        a[mt][ks] = (row < ND && col < ND) ? A[row * ND + col] : 0.f;
      }
  f32x4 acc[MT][NT];
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
      {
        const float b = lds[(4 * ks + g4) * LD + 16 * nt + r16];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
          acc[mt][nt] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mt][ks], b, acc[mt][nt], 0, 0, 0);
      }
  // accumulators back as rows i, columns v * 64 + cell (rows past ND are not written)
#pragma unroll
  for (int mt = 0; mt < MT; ++mt)
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        {
          const int i = 16 * mt + 4 * g4 + e;
          if (i < ND)
            lds[i * LD + 16 * nt + r16] = acc[mt][nt][e];
        }
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int i = 0; i < ND; ++i)
      y[v][i] = lds[i * LD + v * CB + lane];
}

}  // namespace oceano
