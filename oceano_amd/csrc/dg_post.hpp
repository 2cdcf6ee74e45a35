// dg_post.hpp -- per-cell work around the stage that the time loop needs on the GPU:
//
//   k_velocity         evaluate_velocity_field (OceanDG.h:2600-2662): velocity at the
//                      Gauss(p+1) points (z_b from the cell_2 store), mapped back to the
//                      nodal basis by the inverse collocation matrix in each direction
//   k_transport_speed  compute_cell_transport_speed (OceanDG.h:2506-2580): per batch the
//                      largest |J^-1 v|_inf + lambda_max(J^-T J^-1)^(1/2) sqrt(g h) at the
//                      GL(p+2) points (z_b from cell_1), with the reference's five power
//                      iterations for lambda_max
//   k_volume           check_mass (OceanDG.h:2138-2184), cell part: per batch the integral
//                      of the depth at the GL(p+2) points
//
// FP64 throughout (these run once per step or every few steps).  Per-batch results are
// reduced on the host in batch order, so the sums do not depend on the launch.
#pragma once

#include "dg_stage.hpp"

namespace oceano {

// values of one component of a cell at the m x m points of a 1-D table S[m][n1]
template <int P, int M>
__device__ __forceinline__ void cell_values(const double (&S)[MAXNG][MAXN1], const double *U,
                                            double *out)
{
  constexpr int N1 = P + 1;
  double T[M * N1];
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int a = 0; a < M; ++a)
      {
        double s = 0;
#pragma unroll
        for (int i = 0; i < N1; ++i)
          s += S[a][i] * U[i + N1 * j];
        T[a + M * j] = s;
      }
#pragma unroll
  for (int b = 0; b < M; ++b)
#pragma unroll
    for (int a = 0; a < M; ++a)
      {
        double s = 0;
#pragma unroll
        for (int j = 0; j < N1; ++j)
          s += S[b][j] * T[a + M * j];
        out[a + M * b] = s;
      }
}

template <int P>
__device__ __forceinline__ void load_cell(const double *__restrict__ v, int ncomp, int comp,
                                          int batch, int lane, double *U)
{
  constexpr int ND = (P + 1) * (P + 1);
#pragma unroll
  for (int k = 0; k < ND; ++k)
    U[k] = v[vix(batch, ncomp, comp, ND, k, lane)];
}

template <int P>
__global__ __launch_bounds__(CB) void k_velocity(OpArgs O, const double *__restrict__ zeta,
                                                 const double *__restrict__ q,
                                                 double *__restrict__ vel)
{
  constexpr int N1 = P + 1, ND = N1 * N1;
  const DegreeTables &T = basis<double>(P);
  const int batch = blockIdx.x, lane = threadIdx.x;
  if (batch * CB + lane >= O.n_cells)
    return;
  const Phys<double> M = physics<double>(O);
  double U[ND], z[ND], qx[ND], qy[ND];
  // Sn is [n1][n1]; widen to the [MAXNG][MAXN1] shape of cell_values
  double Sn[MAXNG][MAXN1] = {};
#pragma unroll
  for (int a = 0; a < N1; ++a)
#pragma unroll
    for (int i = 0; i < N1; ++i)
      Sn[a][i] = T.Sn[a][i];
  load_cell<P>(zeta, 1, 0, batch, lane, U);
  cell_values<P, N1>(Sn, U, z);
  load_cell<P>(q, 2, 0, batch, lane, U);
  cell_values<P, N1>(Sn, U, qx);
  load_cell<P>(q, 2, 1, batch, lane, U);
  cell_values<P, N1>(Sn, U, qy);
  const size_t zb0 = (size_t)batch * ND * CB + lane;
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      const double zb = fld(O.cell2, zb0 + (size_t)k * CB);
      const double s = M.inv_depth(M.depth(z[k], zb));
      qx[k] *= s;
      qy[k] *= s;
    }
  // transform_from_q_points_to_basis: U = (Ti x Ti) V
  for (int c = 0; c < 2; ++c)
    {
      const double *V = c == 0 ? qx : qy;
      double X[ND];
#pragma unroll
      for (int b = 0; b < N1; ++b)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            double s = 0;
#pragma unroll
            for (int a = 0; a < N1; ++a)
              s += T.Ti[i][a] * V[a + N1 * b];
            X[i + N1 * b] = s;
          }
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            double s = 0;
#pragma unroll
            for (int b = 0; b < N1; ++b)
              s += T.Ti[j][b] * X[i + N1 * b];
            vel[vix(batch, 2, c, ND, i + N1 * j, lane)] = s;
          }
    }
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v = fmax(v, __shfl_xor(v, off));
  return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off);
  return v;
}

template <int P>
__global__ __launch_bounds__(CB) void k_transport_speed(OpArgs O,
                                                        const double *__restrict__ zeta,
                                                        const double *__restrict__ q,
                                                        double *__restrict__ out_blk)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NG = P + 2, NP = NG * NG;
  const DegreeTables &T = basis<double>(P);
  const int batch = blockIdx.x, lane = threadIdx.x;
  double local_max = 0.0;
  if (batch * CB + lane < O.n_cells)
    {
      const Phys<double> M = physics<double>(O);
      const double ij[2] = {1.0 / O.hxy[((size_t)batch * 2) * CB + lane],
                            1.0 / O.hxy[((size_t)batch * 2 + 1) * CB + lane]};
      // the power iteration depends on the cell only
      double ev[2] = {1.0, 1.0};
      for (int i = 0; i < 5; ++i)
        {
          ev[0] = ij[0] * (ij[0] * ev[0]);
          ev[1] = ij[1] * (ij[1] * ev[1]);
          const double nrm = fmax(fabs(ev[0]), fabs(ev[1]));
          ev[0] /= nrm;
          ev[1] /= nrm;
        }
      const double je0 = ij[0] * ev[0], je1 = ij[1] * ev[1];
      const double lam = sqrt((je0 * je0 + je1 * je1) / (ev[0] * ev[0] + ev[1] * ev[1]));
      double U[ND], z[NP], qx[NP], qy[NP];
      load_cell<P>(zeta, 1, 0, batch, lane, U);
      cell_values<P, NG>(T.Sg, U, z);
      load_cell<P>(q, 2, 0, batch, lane, U);
      cell_values<P, NG>(T.Sg, U, qx);
      load_cell<P>(q, 2, 1, batch, lane, U);
      cell_values<P, NG>(T.Sg, U, qy);
      const size_t zb0 = (size_t)batch * NP * CB + lane;
#pragma unroll
      for (int k = 0; k < NP; ++k)
        {
          const double zb = fld(O.cell1, zb0 + (size_t)k * CB);
          const double s = M.inv_depth(M.depth(z[k], zb));
          const double conv = fmax(fabs(ij[0] * (qx[k] * s)), fabs(ij[1] * (qy[k] * s)));
          const double sos = sqrt(M.g * M.depth(z[k], zb));
          local_max = fmax(local_max, lam * sos + conv);
        }
    }
  local_max = wave_max(local_max);
  if (lane == 0)
    out_blk[batch] = local_max;
}

template <int P>
__global__ __launch_bounds__(CB) void k_volume(OpArgs O, const double *__restrict__ zeta,
                                               double *__restrict__ out_blk)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NG = P + 2, NP = NG * NG;
  const DegreeTables &T = basis<double>(P);
  const int batch = blockIdx.x, lane = threadIdx.x;
  double sum = 0.0;
  if (batch * CB + lane < O.n_cells)
    {
      const Phys<double> M = physics<double>(O);
      const double area = O.hxy[((size_t)batch * 2) * CB + lane] *
                          O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      double U[ND], z[NP];
      load_cell<P>(zeta, 1, 0, batch, lane, U);
      cell_values<P, NG>(T.Sg, U, z);
      const size_t zb0 = (size_t)batch * NP * CB + lane;
#pragma unroll
      for (int b = 0; b < NG; ++b)
#pragma unroll
        for (int a = 0; a < NG; ++a)
          {
            const int k = a + NG * b;
            const double zb = fld(O.cell1, zb0 + (size_t)k * CB);
            sum += M.depth(z[k], zb) * (T.wg[a] * T.wg[b] * area);
          }
    }
  sum = wave_sum(sum);
  if (lane == 0)
    out_blk[batch] = sum;
}

template <int P>
int launch_post(oc_dg *op, const OpArgs &O, int what, const double *zeta, const double *q,
                double *out, hipStream_t s)
{
  const dim3 grid(op->NB), block(CB);
  switch (what)
    {
      case POST_VELOCITY:
        hipLaunchKernelGGL((k_velocity<P>), grid, block, 0, s, O, zeta, q, out);
        break;
      case POST_SPEED:
        hipLaunchKernelGGL((k_transport_speed<P>), grid, block, 0, s, O, zeta, q, out);
        break;
      case POST_VOLUME:
        hipLaunchKernelGGL((k_volume<P>), grid, block, 0, s, O, zeta, out);
        break;
      default:
        return OC_ERR_ARGUMENT;
    }
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

}  // namespace oceano
