// dg_ark.hpp -- the additive (IMEX) Runge-Kutta stage: perform_stage_hydro of
// OceanDG.h:1936-2119 with the explicit/implicit split of the momentum source.
//
//   k_stiff        local_apply_cell_stiff_discharge (OceanDG.h:794-826): the Manning
//                  friction source -F(v, n, h) tested against the cell basis at Gauss(nq)
//   k_ark_update   the discharge update after the integrals.  Intermediate stage:
//                  front = sum_j (a_j kex_j + a~_j kim_j) + M w_q (local_apply_cell_mass_
//                  discharge, OceanDG.h:858-879), r_q = (S^T W' S)^-1 front with the
//                  collocation weights W' = JxW (1 + fm J(v, n, h)) at Gauss(p+1) from the
//                  stage input (local_apply_inverse_modified_mass_matrix_discharge,
//                  OceanDG.h:1574-1645).  Last stage: front = b_0 kex_0 + sum_{j>=1}
//                  (b_j kex_j + b_{j-1} kim_j), r_q = M^-1 front, w_q += r_q (2040-2071).
// The height part of the stage and the explicit momentum integrals reuse the ERK kernels
// (k_height_rhs, k_discharge with the friction switched off, k_height_jacobi).  FP64 for
// the update; the integrals follow the operator's precision.  Every k vector of the stage
// (height, explicit and stiff momentum terms) is stored in the integral precision R: on
// the FP32 path their values are FP32 numbers, so the storage is exact and halves their
// traffic; the modified-mass update reads them into FP64.
#pragma once

#include "dg_stage.hpp"

namespace oceano {

struct ArkArgs
{
  const double *ri_h, *ri_q;
  const double *kex[MAXKI], *kim[MAXKI];  // per earlier stage j: ki_q[2j+1], ki_q[2j+2]
  double factor[MAXKI], factor_tilde[MAXKI];
  double fm;  // d_tilde[stage] dt
  int stage, last;
  const double *wsol;  // w_q
  double *rdst;        // r_q
  double *wq_out;      // w_q at the last stage
  double *kim_out;     // this stage's stiff k vector (ki_q[2 stage + 2]; paired stage only)
};

// the stiff friction source of one cell tested against the basis: Rx, Ry (Gauss(nq)); own
// data Z (deviation from zcr), QX, QY through accessors (registers or an LDS copy)
template <int P, typename R, typename ZA, typename QA>
__device__ __forceinline__ void stiff_source(const OpArgs &O, const Basis<R> &T, const ZA &Z,
                                             const QA &QX, const QA &QY, R zcr, int batch,
                                             int lane, R area, R (&Rx)[(P + 1) * (P + 1)],
                                             R (&Ry)[(P + 1) * (P + 1)])
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P);
  const Phys<R> M = physics<R>(O);
#pragma unroll
  for (int k = 0; k < ND; ++k)
    Rx[k] = Ry[k] = 0;
  const size_t c0base = (size_t)batch * NQ * NQ * CB + lane;
#pragma unroll 1
  for (int b = 0; b < NQ; ++b)
    {
      R zs[N1], xs[N1], ys[N1], WX[N1], WY[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          R s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            {
              s0 += T.Sq[b][j] * Z[i + N1 * j];
              s1 += T.Sq[b][j] * QX[i + N1 * j];
              s2 += T.Sq[b][j] * QY[i + N1 * j];
            }
          zs[i] = s0;
          xs[i] = s1;
          ys[i] = s2;
          WX[i] = WY[i] = 0;
        }
#pragma unroll
      for (int a = 0; a < NQ; ++a)
        {
          R z = 0, qx = 0, qy = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              z += T.Sq[a][i] * zs[i];
              qx += T.Sq[a][i] * xs[i];
              qy += T.Sq[a][i] * ys[i];
            }
          z += zcr;
          const size_t pi = c0base + (size_t)(a + NQ * b) * CB;
          const R zb = R(fld(O.cell0[0], pi));
          const R n_m = R(fld(O.cell0[1], pi));
          // ShallowWater.h:461-472, BottomFriction.h:186-199
          const R h = M.depth(z, zb);
          const R s = M.inv_depth(h);
          const R vx = qx * s, vy = qy * s;
          const R inv_d = R(1) / (exp(log(h) * R(0.33333333333)));
          const R hcf = M.g * n_m * n_m * inv_d * sqrt(vx * vx + vy * vy);
          const R JxW = T.wq[a] * T.wq[b] * area;
          const R v0 = -(hcf * vx) * JxW, v1 = -(hcf * vy) * JxW;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              WX[i] += T.Sq[a][i] * v0;
              WY[i] += T.Sq[a][i] * v1;
            }
        }
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            Rx[i + N1 * j] += T.Sq[b][j] * WX[i];
            Ry[i + N1 * j] += T.Sq[b][j] * WY[i];
          }
    }
}

template <int P, typename R>
__global__ __launch_bounds__(CB) void k_stiff(OpArgs O, const double *__restrict__ ri_h,
                                              const double *__restrict__ ri_q,
                                              double *__restrict__ dst_buf)
{
  constexpr int ND = (P + 1) * (P + 1);
  const Basis<R> &T = basis<R>(P);
  R *const dst = reinterpret_cast<R *>(dst_buf);  // the stiff k vector, stored in R
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x;
  if (batch * CB + lane >= O.n_cells)
    return;
  const R hx = R(O.hxy[((size_t)batch * 2) * CB + lane]);
  const R hy = R(O.hxy[((size_t)batch * 2 + 1) * CB + lane]);
  R Z[ND], QX[ND], QY[ND];
  const double zc = load_own<P, R>(ri_h, ri_q, batch, lane, Z, QX, QY);
  R Rx[ND], Ry[ND];
  stiff_source<P, R>(O, T, (const R *)Z, (const R *)QX, (const R *)QY, R(zc), batch, lane,
                     hx * hy, Rx, Ry);
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      dst[vix(batch, 2, 0, ND, k, lane)] = Rx[k];
      dst[vix(batch, 2, 1, ND, k, lane)] = Ry[k];
    }
}

// X[a + n1 b] = sum_ij A[i][a] A[j][b] U[i + n1 j] (A given as A[i][a])
template <int P>
__device__ __forceinline__ void tensor_apply_T(const double (&A)[MAXN1][MAXN1], const double *U,
                                               double *X)
{
  constexpr int N1 = P + 1;
  double T[N1 * N1];
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      {
        double s = 0;
#pragma unroll
        for (int i = 0; i < N1; ++i)
          s += A[i][a] * U[i + N1 * j];
        T[a + N1 * j] = s;
      }
#pragma unroll
  for (int b = 0; b < N1; ++b)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      {
        double s = 0;
#pragma unroll
        for (int j = 0; j < N1; ++j)
          s += A[j][b] * T[a + N1 * j];
        X[a + N1 * b] = s;
      }
}

// Y[i + n1 j] = sum_ab A[i][a] A[j][b] X[a + n1 b]
template <int P>
__device__ __forceinline__ void tensor_apply(const double (&A)[MAXN1][MAXN1], const double *X,
                                             double *Y)
{
  constexpr int N1 = P + 1;
  double T[N1 * N1];
#pragma unroll
  for (int b = 0; b < N1; ++b)
#pragma unroll
    for (int i = 0; i < N1; ++i)
      {
        double s = 0;
#pragma unroll
        for (int a = 0; a < N1; ++a)
          s += A[i][a] * X[a + N1 * b];
        T[i + N1 * b] = s;
      }
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int i = 0; i < N1; ++i)
      {
        double s = 0;
#pragma unroll
        for (int b = 0; b < N1; ++b)
          s += A[j][b] * T[i + N1 * b];
        Y[i + N1 * j] = s;
      }
}

// the discharge update of one cell (lane) of an active batch slot; KT: storage type of the
// discharge k vectors (the integral precision)
template <int P, typename KT>
__device__ __forceinline__ void ark_update_cell(const OpArgs &O, const ArkArgs &A, int batch,
                                                int lane)
{
  auto kex = [&](int j, size_t i) { return double(reinterpret_cast<const KT *>(A.kex[j])[i]); };
  auto kim = [&](int j, size_t i) { return double(reinterpret_cast<const KT *>(A.kim[j])[i]); };
  constexpr int N1 = P + 1, ND = N1 * N1;
  const DegreeTables &T = basis<double>(P);
  const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
  const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
  const double area = hx * hy;
  // collocation weights 1 / JxW at Gauss(p+1), modified by the friction Jacobian
  double iw[ND];
#pragma unroll
  for (int b = 0; b < N1; ++b)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      iw[a + N1 * b] = 1. / (T.wn[a] * T.wn[b] * area);
  if (!A.last)
    {
      const Phys<double> M = physics<double>(O);
      double U[ND], z[ND], qx[ND], qy[ND];
#pragma unroll
      for (int k = 0; k < ND; ++k)
        U[k] = A.ri_h[vix(batch, 1, 0, ND, k, lane)];
      // values at the Gauss(p+1) points: Sn[a][i] = l_i(x_a); as A[i][a] that is Sn^T
      double SnT[MAXN1][MAXN1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
#pragma unroll
        for (int a = 0; a < N1; ++a)
          SnT[i][a] = T.Sn[a][i];
      tensor_apply_T<P>(SnT, U, z);
#pragma unroll
      for (int k = 0; k < ND; ++k)
        U[k] = A.ri_q[vix(batch, 2, 0, ND, k, lane)];
      tensor_apply_T<P>(SnT, U, qx);
#pragma unroll
      for (int k = 0; k < ND; ++k)
        U[k] = A.ri_q[vix(batch, 2, 1, ND, k, lane)];
      tensor_apply_T<P>(SnT, U, qy);
      const size_t c2 = (size_t)batch * ND * CB + lane;
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          const double zb = fld(O.cell2, c2 + (size_t)k * CB);
          const double cf = fld(O.cell2n, c2 + (size_t)k * CB);
          const double h = M.depth(z[k], zb);
          const double s = M.inv_depth(h);
          const double vx = qx[k] * s, vy = qy[k] * s;
          // BottomFrictionManning::jacobian (BottomFriction.h:201-212)
          const double jac = M.g * cf * cf * (1. / (exp(log(h) * 1.33333333333))) *
                             sqrt(vx * vx + vy * vy);
          iw[k] *= 1. / (1. + A.fm * jac);
        }
    }
  for (int c = 0; c < 2; ++c)
    {
      double F[ND];
      const size_t base = ((size_t)batch * 2 + c) * ND * CB + lane;
      if (A.last)
        {
#pragma unroll
          for (int k = 0; k < ND; ++k)
            F[k] = A.factor[0] * kex(0, base + (size_t)k * CB);
          for (int j = 1; j <= A.stage; ++j)
            {
#pragma unroll
              for (int k = 0; k < ND; ++k)
                F[k] += A.factor[j] * kex(j, base + (size_t)k * CB) +
                        A.factor[j - 1] * kim(j - 1, base + (size_t)k * CB);
            }
        }
      else
        {
#pragma unroll
          for (int k = 0; k < ND; ++k)
            F[k] = A.factor[0] * kex(0, base + (size_t)k * CB) +
                   A.factor_tilde[0] * kim(0, base + (size_t)k * CB);
          for (int j = 1; j <= A.stage; ++j)
            {
#pragma unroll
              for (int k = 0; k < ND; ++k)
                F[k] += A.factor[j] * kex(j, base + (size_t)k * CB) +
                        A.factor_tilde[j] * kim(j, base + (size_t)k * CB);
            }
          // + M w_q at Gauss(p+1): S^T W S w
          double U[ND], V[ND], MW[ND];
#pragma unroll
          for (int k = 0; k < ND; ++k)
            U[k] = A.wsol[base + (size_t)k * CB];
          double SnT[MAXN1][MAXN1], Sn[MAXN1][MAXN1];
#pragma unroll
          for (int i = 0; i < N1; ++i)
#pragma unroll
            for (int a = 0; a < N1; ++a)
              {
                SnT[i][a] = T.Sn[a][i];
                Sn[a][i] = T.Sn[a][i];
              }
          tensor_apply_T<P>(SnT, U, V);
#pragma unroll
          for (int b = 0; b < N1; ++b)
#pragma unroll
            for (int a = 0; a < N1; ++a)
              V[a + N1 * b] *= T.wn[a] * T.wn[b] * area;
          // integrate: MW[i + n1 j] = sum_ab Sn[a][i] Sn[b][j] V[a + n1 b]
          tensor_apply_T<P>(Sn, V, MW);
#pragma unroll
          for (int k = 0; k < ND; ++k)
            F[k] += MW[k];
        }
      // (S^T W' S)^-1 F = S^-1 W'^-1 S^-T F with Ti = S^-1 stored [i][a]
      double X[ND], Y[ND];
      tensor_apply_T<P>(T.Ti, F, X);
#pragma unroll
      for (int k = 0; k < ND; ++k)
        X[k] *= iw[k];
      tensor_apply<P>(T.Ti, X, Y);
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          A.rdst[base + (size_t)k * CB] = Y[k];
          if (A.last)
            A.wq_out[base + (size_t)k * CB] += Y[k];
        }
    }
}

template <int P, typename KT>
__global__ __launch_bounds__(CB) void k_ark_update(OpArgs O, ArkArgs A)
{
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x;
  if (batch * CB + lane >= O.n_cells)
    return;
  ark_update_cell<P, KT>(O, A, batch, lane);
}

// The momentum item of the paired additive stage: the explicit integrals (k_stage_pair's
// discharge item in its rhs form: faces 2 and 3 and the cell, plus the exchanged faces 0
// and 1) stored as kex; before the last stage the stiff friction source of the same stage
// input (read from the LDS copy of the batch's own data, still intact) stored as kim; then
// the modified-mass update of each cell (ark_update_cell, which reads this stage's kex and
// kim back from the vectors this lane has just written).
template <int P, typename R>
__device__ __forceinline__ void discharge_ark_block(const OpArgs &O, const StageArgs &Q,
                                                    const ArkArgs &A, int batch, R *lds,
                                                    R *xch)
{
  constexpr int ND = (P + 1) * (P + 1);
  discharge_block<P, false, 4, R, true>(O, Q, batch, lds, xch);
  const int lane = threadIdx.x & (CB - 1);
  if (batch * CB + lane >= O.n_cells)
    return;
  if (!A.last)
    {
      const R hx = R(O.hxy[((size_t)batch * 2) * CB + lane]);
      const R hy = R(O.hxy[((size_t)batch * 2 + 1) * CB + lane]);
      const double zc = centred<R>() ? Q.src_h[vix(batch, 1, 0, ND, 0, lane)] : 0.0;
      const LdsVec<R> Z{lds + lane}, QX{lds + ND * CB + lane}, QY{lds + 2 * ND * CB + lane};
      R Rx[ND], Ry[ND];
      stiff_source<P, R>(O, basis<R>(P), Z, QX, QY, R(zc), batch, lane, hx * hy, Rx, Ry);
      R *kim = reinterpret_cast<R *>(A.kim_out);
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          kim[vix(batch, 2, 0, ND, k, lane)] = Rx[k];
          kim[vix(batch, 2, 1, ND, k, lane)] = Ry[k];
        }
    }
  ark_update_cell<P, R>(O, A, batch, lane);
}

// The additive stage in one launch with paired work items (inviscid): workgroup = one batch,
// wave 0 the height item (L_zeta, front and depth solve: the explicit stage's height item
// with the additive stage's factors), wave 1 the momentum item (discharge_ark_block).  The
// operator arguments carry the friction switched off before the last stage (the stiff term
// takes it; local_apply_cell_nonstiff_discharge).
// waves per SIMD of the paired additive stage: the momentum item's FP64 modified-mass update
// needs more registers than the explicit stage's momentum item.  FP64 at p <= 3: two waves
// (at p = 3, 256 VGPRs and 20 spill slots; measured 173 us per C2 stage against 204 us at
// one wave per SIMD with no spills; the workgroup's 36 KiB of LDS lets four share a CU).
// FP32 at p = 4: one wave (two would spill 279 slots)
template <int P, typename R>
constexpr int ark_min_waves()
{
  return std::is_same<R, float>::value ? (P <= 2 ? 3 : P == 3 ? 2 : 1) : (P <= 3 ? 2 : 1);
}

template <int P, typename R>
__global__ __launch_bounds__(2 * CB, (ark_min_waves<P, R>())) void k_stage_ark_pair(
  OpArgs O, StageArgs H, StageArgs Q, ArkArgs A)
{
  __shared__ __align__(16) R lds[pair_lds_words<P, R>()];
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  R *xch = lds + pair_xch_offset<P, R>();
  R *dlds = lds + pair_height_words<P, R>();
  // the height item on wave 0 of every workgroup: unlike the explicit stage (pair_height_wave),
  // alternating the roles measured slower here (mixed C2 stage 119-122 against 106-112 us)
  if (wave_in_block() == 0)
    height_stage_block<P, R, true>(O, H, batch, reinterpret_cast<float *>(lds), xch, dlds);
  else
    discharge_ark_block<P, R>(O, Q, A, batch, dlds, xch);
}

template <int P, typename R>
int launch_stage_ark(oc_dg *op, const OpArgs &O, uint32_t stage, uint32_t n_stages,
                     const double *factor, const double *factor_tilde, double fm,
                     const double *ri_h, const double *ri_q, const double *ri_v,
                     double *const *ki_h, double *const *ki_q, double *w_h, double *w_q,
                     double *r_h, double *r_q, hipStream_t s)
{
  const bool last = stage == n_stages - 1;
  const dim3 grid(op->NB), block(CB);
  hipEvent_t ev = nullptr;
  if (O.nu == 0.0 && op->paired_items)
    {
      // one launch: k_stage_ark_pair
      OpArgs Oe = O;
      if (!last)
        Oe.has_friction = 0;
      StageArgs H{}, Q{};
      H.src_h = ri_h;
      H.src_q = ri_q;
      H.dst = ki_h[stage + 1];
      for (uint32_t j = 0; j < stage; ++j)
        H.kprev[j] = ki_h[j + 1];
      for (uint32_t j = 0; j <= stage; ++j)
        H.factor[j] = factor[j];
      H.stage = (int)stage;
      H.wsol = w_h;
      H.rk_dst = last ? w_h : r_h;
      H.sweeps_blk = O.sweeps;
      Q.src_h = ri_h;
      Q.src_q = ri_q;
      Q.src_v = ri_v;
      Q.dst = ki_q[2 * stage + 1];
      ArkArgs A{};
      A.ri_h = ri_h;
      A.ri_q = ri_q;
      for (uint32_t j = 0; j <= stage; ++j)
        {
          A.kex[j] = ki_q[2 * j + 1];
          A.kim[j] = ki_q[2 * j + 2 < 2 * n_stages ? 2 * j + 2 : 0];
          A.factor[j] = factor[j];
          A.factor_tilde[j] = factor_tilde ? factor_tilde[j] : 0.0;
        }
      A.fm = fm;
      A.stage = (int)stage;
      A.last = last;
      A.wsol = w_q;
      A.rdst = r_q;
      A.wq_out = w_q;
      A.kim_out = last ? nullptr : ki_q[2 * stage + 2];
      timed_begin(op, s, &ev);
      hipLaunchKernelGGL((k_stage_ark_pair<P, R>), grid, dim3(2 * CB), 0, s, Oe, H, Q, A);
      timed_end(op, s, TK_STAGE, ev);
      HIP_TRY(hipGetLastError());
      if (std::is_same<R, float>::value && defer_depth_solve(P))
        {
          // FP64 depth solve of the cells the FP32 stage left near the depth floor
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
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_rhs<P, false, R, R>), grid, block, 0, s, O, ri_h, ri_q,
                     ki_h[stage + 1]);
  timed_end(op, s, TK_HEIGHT_RHS, ev);
  // explicit momentum integrals: the full operator at the last stage, without the friction
  // source before (local_apply_cell_nonstiff_discharge)
  StageArgs D{};
  D.src_h = ri_h;
  D.src_q = ri_q;
  D.src_v = ri_v;
  D.dst = ki_q[2 * stage + 1];
  OpArgs Oe = O;
  if (!last)
    Oe.has_friction = 0;
  timed_begin(op, s, &ev);
  if (O.nu != 0.0)
    hipLaunchKernelGGL((k_discharge<P, true, 4, R>), grid, block, 0, s, Oe, D);
  else
    hipLaunchKernelGGL((k_discharge<P, false, 4, R>), grid, block, 0, s, Oe, D);
  if (!last)
    hipLaunchKernelGGL((k_stiff<P, R>), grid, block, 0, s, O, ri_h, ri_q, ki_q[2 * stage + 2]);
  timed_end(op, s, TK_DISCHARGE, ev);
  // height: front and depth solve (as in the explicit stage)
  StageArgs J{};
  for (uint32_t j = 0; j <= stage; ++j)
    {
      J.kprev[j] = ki_h[j + 1];
      J.factor[j] = factor[j];
    }
  J.stage = (int)stage;
  J.wsol = w_h;
  J.rk_dst = last ? w_h : r_h;
  J.sweeps_blk = O.sweeps;
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_height_jacobi<P, R, R>), grid, block, 0, s, O, J);
  timed_end(op, s, TK_HEIGHT_JACOBI, ev);
  // momentum update
  ArkArgs A{};
  A.ri_h = ri_h;
  A.ri_q = ri_q;
  for (uint32_t j = 0; j <= stage; ++j)
    {
      A.kex[j] = ki_q[2 * j + 1];
      A.kim[j] = ki_q[2 * j + 2 < 2 * n_stages ? 2 * j + 2 : 0];
      A.factor[j] = factor[j];
      A.factor_tilde[j] = factor_tilde ? factor_tilde[j] : 0.0;
    }
  A.fm = fm;
  A.stage = (int)stage;
  A.last = last;
  A.wsol = w_q;
  A.rdst = r_q;
  A.wq_out = w_q;
  timed_begin(op, s, &ev);
  hipLaunchKernelGGL((k_ark_update<P, R>), grid, block, 0, s, O, A);
  timed_end(op, s, TK_RK_UPDATE, ev);
  HIP_TRY(hipGetLastError());
  return OC_OK;
}

}  // namespace oceano
