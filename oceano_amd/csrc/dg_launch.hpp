// dg_launch.hpp -- launch sequences of one stage / rhs / apply, per degree.
#pragma once

#include "dg_internal.hpp"
#include "dg_stage_fp64.hpp"

namespace oceano {
namespace {

__global__ void k_zero_u64(unsigned long long *p) { *p = 0ull; }

struct RkArgs
{
  size_t n;
  const double *k[MAXKI];
  double factor[MAXKI];
  int nk;
  const double *w;
  double *dst;
};

// dst = w + sum_j factor_j k_j, in the reference's order (OceanDG.h:1878-1909)
__global__ void k_rk_update(RkArgs A)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < A.n;
       i += (size_t)gridDim.x * blockDim.x)
    {
      double acc = A.w[i] + A.factor[0] * A.k[0][i];
      for (int j = 1; j < A.nk; ++j)
        acc += A.factor[j] * A.k[j][i];
      A.dst[i] = acc;
    }
}


template <int P>
int launch_stage(oc_dg *op, const OpArgs &O, uint32_t stage, const double *factor,
                 const double *ri_h, const double *ri_q, const double *ri_v,
                 double *const *ki_h, double *const *ki_q, uint32_t n_ki, double *w_h,
                 double *w_q, double *r_h, double *r_q, int write_scratch, hipStream_t s)
{
  const bool last = (stage == n_ki - 2);
  const dim3 grid(op->NB), block(CB);
  hipEvent_t ev = nullptr;
  // 1) L_zeta -> ki_h[stage+1]
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_rhs<P, false>), grid, block, 0, s, O, ri_h, ri_q,
                     ki_h[stage + 1]);
  timed_end(op, s, TK_HEIGHT_RHS, ev);
  // 2) L_q -> M^-1 L_q -> ki_q[stage+1], and r_q (or w_q) = w_q + sum_j a_j ki_q[j+1]
  double *q_dst = last ? w_q : r_q;
  const bool alias = (q_dst == ri_q);  // only for single-stage schemes (forward Euler)
  StageArgs S{};
  S.src_h = ri_h;
  S.src_q = ri_q;
  S.src_v = ri_v;
  S.dst = ki_q[stage + 1];
  S.scratch = write_scratch ? ki_q[0] : nullptr;
  for (uint32_t j = 0; j < stage; ++j)
    S.kprev[j] = ki_q[j + 1];
  for (uint32_t j = 0; j <= stage; ++j)
    S.factor[j] = factor[j];
  S.stage = (int)stage;
  S.wsol = w_q;
  S.rk_dst = q_dst;
  S.sweeps = nullptr;
  timed_begin(op, s, &ev);
  if (alias)
    {
      // single-stage schemes: ri_q is the updated vector itself, so the inverse mass must
      // finish on every cell (neighbour traces) before the update overwrites it
      StageArgs A = S;
      if (O.nu != 0.0)
        hipLaunchKernelGGL((k_discharge<P, true, 1>), grid, block, 0, s, O, A);
      else
        hipLaunchKernelGGL((k_discharge<P, false, 1>), grid, block, 0, s, O, A);
      RkArgs U{};
      U.n = (size_t)op->NB * CB * 2 * op->nd;
      for (uint32_t j = 0; j <= stage; ++j)
        {
          U.k[j] = ki_q[j + 1];
          U.factor[j] = factor[j];
        }
      U.nk = (int)stage + 1;
      U.w = w_q;
      U.dst = q_dst;
      hipLaunchKernelGGL(k_rk_update, dim3((U.n + 255) / 256), dim3(256), 0, s, U);
    }
  else if (O.nu != 0.0)
    hipLaunchKernelGGL((k_discharge<P, true, 0>), grid, block, 0, s, O, S);
  else
    hipLaunchKernelGGL((k_discharge<P, false, 0>), grid, block, 0, s, O, S);
  timed_end(op, s, TK_DISCHARGE, ev);
  // 3) front and Jacobi depth solve -> r_h (or w_h in place at the last stage)
  StageArgs H{};
  for (uint32_t j = 0; j <= stage; ++j)
    {
      H.kprev[j] = ki_h[j + 1];
      H.factor[j] = factor[j];
    }
  H.stage = (int)stage;
  H.wsol = w_h;
  H.rk_dst = last ? w_h : r_h;
  H.scratch = write_scratch ? ki_h[0] : nullptr;
  H.sweeps = op->d_sweeps;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, s, op->d_sweeps);
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_jacobi<P>), grid, block, 0, s, O, H);
  timed_end(op, s, TK_HEIGHT_JACOBI, ev);
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

template <int P>
int launch_rhs_apply(oc_dg *op, const OpArgs &O, bool invert, const double *src_h,
                     const double *src_q, const double *src_v, double *dst_h, double *dst_q,
                     hipStream_t s)
{
  const dim3 grid(op->NB), block(CB);
  if (invert)
    hipLaunchKernelGGL((k_height_rhs<P, true>), grid, block, 0, s, O, src_h, src_q, dst_h);
  else
    hipLaunchKernelGGL((k_height_rhs<P, false>), grid, block, 0, s, O, src_h, src_q, dst_h);
  StageArgs S{};
  S.src_h = src_h;
  S.src_q = src_q;
  S.src_v = src_v;
  S.dst = dst_q;
  const bool visc = O.nu != 0.0;
  if (invert)
    {
      if (visc)
        hipLaunchKernelGGL((k_discharge<P, true, 1>), grid, block, 0, s, O, S);
      else
        hipLaunchKernelGGL((k_discharge<P, false, 1>), grid, block, 0, s, O, S);
    }
  else
    {
      if (visc)
        hipLaunchKernelGGL((k_discharge<P, true, 2>), grid, block, 0, s, O, S);
      else
        hipLaunchKernelGGL((k_discharge<P, false, 2>), grid, block, 0, s, O, S);
    }
  HIP_TRY(hipGetLastError());
  return OC_OK;
}


}  // namespace
}  // namespace oceano
