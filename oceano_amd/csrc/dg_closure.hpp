// dg_closure.hpp -- the learned momentum-source correction inside the stage.
//
// Per cell (one lane), from the stage input: the modal coefficients (orthonormal tensor
// Legendre modes) of zeta, hu, hv and the sub-grid statistics (mean, variance, skewness) of
// the bed form the network input; the output holds the modal coefficients of the
// correction tau of hu and hv, which is brought back to nodal values and added to
// M^-1 L_q by the discharge update (a source s in L_q enters as M s, so adding tau after
// the inverse mass is the source term of ShallowWater.h:425-429 with s = tau).  A cell
// that is not active (no sub-grid bathymetry nearby) or whose network output is not finite
// gets tau = 0: the plain sub-grid model.
#pragma once

#include "dg_internal.hpp"
#include "dg_stage.hpp"
#include "mlp_core.hpp"

extern "C" void oc_mlp_device_state(const oc_mlp *m, const float **p, const uint16_t **wt,
                                    const float **norm);

namespace oceano {

struct ClosureArgs
{
  const double *zeta, *q;
  const double *stats;  // [batch][4][lane]: mean, variance, skewness, active flag
  const double *vi;     // [nd][nd] modal <- nodal
  const double *v;      // [nd][nd] nodal <- modal
  const float *p;       // network parameters (biases)
  const uint16_t *wt;   // transposed BF16 weights
  const float *norm;    // normalisation block
  int n_in, n_out;
  double *tau;          // [batch][2][nd][lane]
  float *features;      // [n_cells][n_in] in the caller's cell order (features kernel)
  const int *order;     // slot -> cell
};

template <int P>
__device__ inline void closure_features(const ClosureArgs &C, int batch, int lane,
                                        float (&f)[mlp::IN])
{
  constexpr int ND = (P + 1) * (P + 1);
  static_assert(3 * ND + 3 <= mlp::IN, "closure inputs exceed the network width");
  double z[ND], qx[ND], qy[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      z[k] = C.zeta[vix(batch, 1, 0, ND, k, lane)];
      qx[k] = C.q[vix(batch, 2, 0, ND, k, lane)];
      qy[k] = C.q[vix(batch, 2, 1, ND, k, lane)];
    }
#pragma unroll
  for (int m = 0; m < ND; ++m)
    {
      double a = 0, b = 0, c = 0;
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          const double w = C.vi[m * ND + k];
          a += w * z[k];
          b += w * qx[k];
          c += w * qy[k];
        }
      f[m] = float(a);
      f[ND + m] = float(b);
      f[2 * ND + m] = float(c);
    }
#pragma unroll
  for (int s = 0; s < 3; ++s)
    f[3 * ND + s] = float(C.stats[((size_t)batch * 4 + s) * CB + lane]);
#pragma unroll
  for (int i = 3 * ND + 3; i < mlp::IN; ++i)
    f[i] = 0.f;
}

template <int P>
__global__ __launch_bounds__(CB) void k_closure_features(OpArgs O, ClosureArgs C)
{
  const int batch = blockIdx.x, lane = threadIdx.x;
  const int slot = batch * CB + lane;
  if (slot >= O.n_cells)
    return;
  float f[mlp::IN];
  closure_features<P>(C, batch, lane, f);
  const size_t row = (size_t)C.order[slot] * C.n_in;
#pragma unroll
  for (int i = 0; i < mlp::IN; ++i)
    if (i < C.n_in)
      C.features[row + i] = f[i];
}

template <int P>
__global__ __launch_bounds__(CB) void k_closure(OpArgs O, ClosureArgs C)
{
  using namespace mlp;
  constexpr int ND = (P + 1) * (P + 1);
  __shared__ __attribute__((aligned(16))) bf16 s_wt[NT];
  __shared__ __attribute__((aligned(16))) bf16 s_x[CB * IN];
  __shared__ __attribute__((aligned(16))) bf16 s_h[3][16 * HID];
  __shared__ float s_y[CB * OUT];
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x;
  for (int i = lane; i < NT; i += CB)
    s_wt[i] = C.wt[i];
  const bool active = batch * CB + lane < O.n_cells;
  float f[IN];
  if (active)
    closure_features<P>(C, batch, lane, f);
#pragma unroll
  for (int k = 0; k < IN; ++k)
    {
      const float x = (active && k < C.n_in) ? (f[k] - C.norm[k]) * C.norm[IN + k] : 0.f;
      s_x[lane * IN + k] = to_bf16(x);
    }
  __syncthreads();
  for (int rt = 0; rt < CB / 16; ++rt)
    forward_tile(s_x + rt * 16 * IN, s_wt, C.p, s_h[0], s_h[1], s_h[2], nullptr,
                 s_y + rt * 16 * OUT);
  double tm[2 * ND];
  bool finite = active && C.stats[((size_t)batch * 4 + 3) * CB + lane] != 0.0;
#pragma unroll
  for (int o = 0; o < 2 * ND; ++o)
    {
      const float y = output_value(s_y[lane * OUT + o], C.norm, o);
      finite = finite && isfinite(y);
      tm[o] = double(y);
    }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < ND; ++k)
      {
        double t = 0;
#pragma unroll
        for (int m = 0; m < ND; ++m)
          t += C.v[k * ND + m] * tm[c * ND + m];
        C.tau[vix(batch, 2, c, ND, k, lane)] = finite ? t : 0.0;
      }
}

// features != nullptr: write the network inputs of every cell (k_closure_features);
// otherwise evaluate the attached network into op->d_closure (k_closure)
template <int P>
int launch_closure(oc_dg *op, const OpArgs &O, const double *zeta, const double *q,
                   float *features, hipStream_t s)
{
  if constexpr (3 * (P + 1) * (P + 1) + 3 > mlp::IN)
    {
      set_error("closure: degree above 2 exceeds the network input width");
      return OC_ERR_ARGUMENT;
    }
  else
    {
      constexpr int ND = (P + 1) * (P + 1);
      ClosureArgs C{};
      C.zeta = zeta;
      C.q = q;
      C.stats = op->d_cstats;
      C.vi = op->d_cvi;
      C.v = op->d_cv;
      C.order = op->d_order;
      C.features = features;
      C.tau = op->d_closure;
      C.n_in = 3 * ND + 3;
      C.n_out = 2 * ND;
      if (features)
        hipLaunchKernelGGL((k_closure_features<P>), dim3(op->NB), dim3(CB), 0, s, O, C);
      else
        {
          oc_mlp_device_state(op->closure_mlp, &C.p, &C.wt, &C.norm);
          hipLaunchKernelGGL((k_closure<P>), dim3(op->NB), dim3(CB), 0, s, O, C);
        }
      HIP_TRY(hipGetLastError());
      return OC_OK;
    }
}

}  // namespace oceano
