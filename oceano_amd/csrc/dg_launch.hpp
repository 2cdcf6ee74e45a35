// dg_launch.hpp -- launch sequences of one stage / rhs / apply, per degree.
#pragma once

#include "dg_ark.hpp"
#include "dg_closure.hpp"
#include "dg_internal.hpp"
#include "dg_post.hpp"
#include "dg_stage.hpp"

namespace oceano {
namespace {

struct RkArgs
{
  size_t n;
  const double *k[MAXKI];
  double factor[MAXKI];
  int nk;
  const double *w;
  double *dst;
};

// dst = w + sum_j factor_j k_j, in the reference's order (OceanDG.h:1878-1909); the k
// vectors are stored in KT
template <typename KT>
__global__ void k_rk_update(RkArgs A)
{
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < A.n;
       i += (size_t)gridDim.x * blockDim.x)
    {
      double acc = A.w[i] + A.factor[0] * double(reinterpret_cast<const KT *>(A.k[0])[i]);
      for (int j = 1; j < A.nk; ++j)
        acc += A.factor[j] * double(reinterpret_cast<const KT *>(A.k[j])[i]);
      A.dst[i] = acc;
    }
}


// the arguments of the height and discharge work of one explicit stage
inline void stage_args(const OpArgs &O, uint32_t stage, const double *factor,
                       const double *ri_h, const double *ri_q, const double *ri_v,
                       double *const *ki_h, double *const *ki_q, uint32_t n_ki, double *w_h,
                       double *w_q, double *r_h, double *r_q, int write_scratch, StageArgs &H,
                       StageArgs &Q)
{
  const bool last = (stage == n_ki - 2);
  double *h_dst = last ? w_h : r_h;
  double *q_dst = last ? w_q : r_q;
  // discharge: L_q -> M^-1 L_q -> ki_q[stage+1]; q_dst = w_q + sum_j a_j ki_q[j+1]
  Q = StageArgs{};
  Q.src_h = ri_h;
  Q.src_q = ri_q;
  Q.src_v = ri_v;
  Q.dst = ki_q[stage + 1];
  Q.scratch = write_scratch ? ki_q[0] : nullptr;
  for (uint32_t j = 0; j < stage; ++j)
    Q.kprev[j] = ki_q[j + 1];
  for (uint32_t j = 0; j <= stage; ++j)
    Q.factor[j] = factor[j];
  Q.stage = (int)stage;
  Q.wsol = w_q;
  Q.rk_dst = q_dst;
  // height: L_zeta -> ki_h[stage+1]; front; Jacobi -> h_dst
  H = StageArgs{};
  H.src_h = ri_h;
  H.src_q = ri_q;
  H.dst = ki_h[stage + 1];
  for (uint32_t j = 0; j < stage; ++j)
    H.kprev[j] = ki_h[j + 1];
  for (uint32_t j = 0; j <= stage; ++j)
    H.factor[j] = factor[j];
  H.stage = (int)stage;
  H.wsol = w_h;
  H.rk_dst = h_dst;
  H.scratch = write_scratch ? ki_h[0] : nullptr;
  H.sweeps_blk = O.sweeps;
}

template <int P, typename R>
int launch_stage(oc_dg *op, const OpArgs &O, uint32_t stage, const double *factor,
                 const double *ri_h, const double *ri_q, const double *ri_v,
                 double *const *ki_h, double *const *ki_q, uint32_t n_ki, double *w_h,
                 double *w_q, double *r_h, double *r_q, int write_scratch, hipStream_t s)
{
  const bool last = (stage == n_ki - 2);
  const dim3 grid(op->NB), block(CB);
  double *h_dst = last ? w_h : r_h;
  double *q_dst = last ? w_q : r_q;
  const bool visc = O.nu != 0.0;
  hipEvent_t ev = nullptr;
  StageArgs H, Q;
  stage_args(O, stage, factor, ri_h, ri_q, ri_v, ki_h, ki_q, n_ki, w_h, w_q, r_h, r_q,
             write_scratch, H, Q);
  // the one-pass stage reads neighbour traces of ri while other cells write the outputs:
  // it needs outputs distinct from the inputs (single-stage schemes write w in place)
  const bool alias = (h_dst == ri_h) || (q_dst == ri_q) || (h_dst == ri_q) || (q_dst == ri_h);
  if (!alias)
    {
      timed_begin(op, s, &ev);
      const int items = 2 * op->NB;
      if (visc)
        hipLaunchKernelGGL((k_stage<P, true, R, 1>), dim3(items), block, 0, s, O, H, Q);
      else if (op->paired_items)
        {
          if (std::is_same<R, float>::value && P <= 3)
            H.out32 = Q.out32 = op->rk_out32 && !last;
          if (std::is_same<R, float>::value && P <= 3 && op->rk_in32)
            hipLaunchKernelGGL((k_stage_pair<P, R, std::is_same<R, float>::value && P <= 3>),
                               grid, dim3(2 * CB), 0, s, O, H, Q);
          else
            hipLaunchKernelGGL((k_stage_pair<P, R>), grid, dim3(2 * CB), 0, s, O, H, Q);
        }
      else if (op->waves_per_block == 4)
        hipLaunchKernelGGL((k_stage<P, false, R, 4>), dim3((items + 3) / 4), dim3(4 * CB), 0,
                           s, O, H, Q);
      else if (op->waves_per_block == 2)
        hipLaunchKernelGGL((k_stage<P, false, R, 2>), dim3((items + 1) / 2), dim3(2 * CB), 0,
                           s, O, H, Q);
      else
        hipLaunchKernelGGL((k_stage<P, false, R, 1>), dim3(items), block, 0, s, O, H, Q);
      timed_end(op, s, TK_STAGE, ev);
      HIP_TRY(hipGetLastError());
      if (std::is_same<R, float>::value && defer_depth_solve(P))
        {
          // FP64 depth solve of the cells the FP32 stage left near the depth floor; a wave
          // without such a cell returns at once
          StageArgs D = H;
          for (uint32_t j = 0; j <= stage; ++j)
            D.kprev[j] = ki_h[j + 1];
          timed_begin(op, s, &ev);
          const dim3 dgrid((op->NB + DEFER_BATCHES - 1) / DEFER_BATCHES);
          hipLaunchKernelGGL((k_depth_deferred<P>), dgrid, block, 0, s, O, D);
          timed_end(op, s, TK_HEIGHT_JACOBI, ev);
          HIP_TRY(hipGetLastError());
        }
      return OC_OK;
    }
  // split path: every kernel finishes reading ri before the next one overwrites it
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_rhs<P, false, R, R>), grid, block, 0, s, O, ri_h, ri_q,
                     ki_h[stage + 1]);
  timed_end(op, s, TK_HEIGHT_RHS, ev);
  timed_begin(op, s, &ev);
  StageArgs A = Q;
  if (visc)
    hipLaunchKernelGGL((k_discharge<P, true, 3, R>), grid, block, 0, s, O, A);
  else
    hipLaunchKernelGGL((k_discharge<P, false, 3, R>), grid, block, 0, s, O, A);
  timed_end(op, s, TK_DISCHARGE, ev);
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
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL(k_rk_update<R>, dim3((U.n + 255) / 256), dim3(256), 0, s, U);
  timed_end(op, s, TK_RK_UPDATE, ev);
  StageArgs J = H;
  for (uint32_t j = 0; j <= stage; ++j)
    J.kprev[j] = ki_h[j + 1];
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_jacobi<P, R, R>), grid, block, 0, s, O, J);
  timed_end(op, s, TK_HEIGHT_JACOBI, ev);
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

template <int P, typename R>
int launch_rhs_apply(oc_dg *op, const OpArgs &O, bool invert, const double *src_h,
                     const double *src_q, const double *src_v, double *dst_h, double *dst_q,
                     hipStream_t s)
{
  const dim3 grid(op->NB), block(CB);
  if (invert)
    hipLaunchKernelGGL((k_height_rhs<P, true, R>), grid, block, 0, s, O, src_h, src_q, dst_h);
  else
    hipLaunchKernelGGL((k_height_rhs<P, false, R>), grid, block, 0, s, O, src_h, src_q, dst_h);
  StageArgs S{};
  S.src_h = src_h;
  S.src_q = src_q;
  S.src_v = src_v;
  S.dst = dst_q;
  const bool visc = O.nu != 0.0;
  if (invert)
    {
      if (visc)
        hipLaunchKernelGGL((k_discharge<P, true, 1, R>), grid, block, 0, s, O, S);
      else
        hipLaunchKernelGGL((k_discharge<P, false, 1, R>), grid, block, 0, s, O, S);
    }
  else
    {
      if (visc)
        hipLaunchKernelGGL((k_discharge<P, true, 2, R>), grid, block, 0, s, O, S);
      else
        hipLaunchKernelGGL((k_discharge<P, false, 2, R>), grid, block, 0, s, O, S);
    }
  HIP_TRY(hipGetLastError());
  return OC_OK;
}


}  // namespace
}  // namespace oceano
