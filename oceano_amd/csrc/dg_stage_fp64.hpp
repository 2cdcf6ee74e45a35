// dg_stage_fp64.hpp -- FP64 kernels of the stage (the reference arithmetic).
//
//   k_height_rhs    L_zeta: local_apply_cell_height + face_height + boundary_face_height
//                   (OceanDG.h:670-693, 937-982, 1106-1214)
//   k_discharge     L_q: local_apply_cell_discharge + face_discharge + boundary_face_discharge
//                   (OceanDG.h:696-743, 985-1050, 1217-1339), fused with the tensor-product
//                   inverse mass (1533-1571) and the post-loop RK update (1878-1909)
//   k_height_jacobi front = sum_j a_j L_zeta,j + M h(zeta^n) (829-855, 1835-1847) fused
//                   with the capped lumped-Jacobi depth-mass solve (1374-1507)
#pragma once

#include <cfloat>

#include "dg_kernels.hpp"

namespace oceano {



// exterior state on a boundary face (OceanDG.h:1130-1203 / 1246-1317); returns at_outflow
template <typename T>
__device__ __forceinline__ bool boundary_state(const Phys<T> &M, int kind, T zm, Vec2<T> qm,
                                               Vec2<T> n, T zb, const double *bv, T &zp,
                                               Vec2<T> &qp)
{
  const T un = qm.x * n.x + qm.y * n.y;
  switch (kind)
    {
      case OC_BC_WALL:
        zp = zm;
        qp = {qm.x - T(2) * un * n.x, qm.y - T(2) * un * n.y};
        return false;
      case OC_BC_SUPERCRITICAL_INFLOW:
        {
          zp = T(bv[0]);
          const T h = M.depth(zp, zb);
          qp = {T(bv[1]) * h, T(bv[2]) * h};
          return false;
        }
      case OC_BC_SUPERCRITICAL_OUTFLOW:
        zp = zm;
        qp = qm;
        return true;
      case OC_BC_HEIGHT_INFLOW:
        zp = T(bv[0]);
        qp = qm;
        return false;
      case OC_BC_DISCHARGE_INFLOW:
        zp = zm;
        qp = {T(bv[1]) * -n.x, T(bv[1]) * -n.y};
        return false;
      default:  // OC_BC_ABSORBING_OUTFLOW (kinds are validated at set-up)
        {
          T z2 = T(bv[0]);
          const T h2 = M.depth(z2, zb);
          const Vec2<T> q2 = {T(bv[1]) * h2, T(bv[2]) * h2};
          const T r_p = M.riemann_p(zm, qm, n, zb);
          const T r_m = M.riemann_m(z2, q2, n, zb);
          const T c_b = T(0.25) * (r_p - r_m);
          const T h_b = c_b * c_b / M.g;
          const T u_b = T(0.5) * (r_p + r_m);
          zp = h_b - zb;
          const T hh = M.depth(zp, zb);
          qp = {u_b * hh * n.x, u_b * hh * n.y};
          return false;
        }
    }
}

// Face terms of one cell: for every face point, `inner` / `bdry` return NC values that are
// multiplied by JxW and integrated against this cell's face basis into R[NC][nd].
// own-cell data accessors: registers (stride 1) or the lane-interleaved LDS copy
struct LdsVec
{
  const double *p;  // &lds[component base + lane]
  __device__ __forceinline__ double operator[](int k) const { return p[k * CB]; }
};

template <int P, int NC, typename ZA, typename QA, typename InnerFn, typename BdryFn>
__device__ __forceinline__ void face_terms(const OpArgs &O, const DegreeTables &T, int batch,
                                           int lane, double hx, double hy, const ZA &Z,
                                           const QA &QX, const QA &QY,
                                           const double *__restrict__ src_h,
                                           const double *__restrict__ src_q,
                                           double (&R)[NC][(P + 1) * (P + 1)], InnerFn inner,
                                           BdryFn bdry)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P), NG = P + 2;
#pragma unroll
  for (int f = 0; f < 4; ++f)
    {
      phase_barrier();
      const size_t fo = ((size_t)batch * 4 + f);
      const int info = O.finfo[fo * CB + lane];
      const int kind = info & 3;
      const int nsub = (info >> 2) & 1;
      const double flen = (f < 2) ? hy : hx;
      double zo_l[N1], xo_l[N1], yo_l[N1];
#pragma unroll
      for (int k = 0; k < N1; ++k)
        {
          zo_l[k] = Z[face_dof<N1>(f, k)];
          xo_l[k] = QX[face_dof<N1>(f, k)];
          yo_l[k] = QY[face_dof<N1>(f, k)];
        }
      if (kind == FK_BDRY)
        {
          const int bf = O.fid[(fo * 2) * CB + lane];
          const int bk = O.bkind[bf];
          const Vec2<double> n = {f == 0 ? -1.0 : f == 1 ? 1.0 : 0.0,
                                  f == 2 ? -1.0 : f == 3 ? 1.0 : 0.0};
#pragma unroll 1
          for (int a = 0; a < NQ; ++a)
            {
              double zo = 0, xo = 0, yo = 0;
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  zo += T.Sq[a][k] * zo_l[k];
                  xo += T.Sq[a][k] * xo_l[k];
                  yo += T.Sq[a][k] * yo_l[k];
                }
              const double zb = fld(O.bdryz, (size_t)bf * NQ + a);
              const double *bv = O.bvals + ((size_t)bf * NQ + a) * 3;
              double val[NC];
              bdry(bk, zo, Vec2<double>{xo, yo}, n, zb, bv, val);
              const double JxW = T.wq[a] * flen;
#pragma unroll
              for (int cc = 0; cc < NC; ++cc)
                {
                  const double v = val[cc] * JxW;
#pragma unroll
                  for (int k = 0; k < N1; ++k)
                    R[cc][face_dof<N1>(f, k)] += T.Sq[a][k] * v;
                }
            }
        }
      else
        {
          const bool is_m = (f & 1);  // canonical orientation: minus side owns faces 1, 3
          const Vec2<double> n = {f < 2 ? 1.0 : 0.0, f < 2 ? 0.0 : 1.0};
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              if (s == 1 && kind != FK_COARSE)
                break;
              phase_barrier();
              const int c2 = O.nbr[(fo * 2 + s) * CB + lane];
              const int fi = O.fid[(fo * 2 + s) * CB + lane];
              const int b2 = c2 / CB, l2 = c2 % CB;
              double zn_l[N1], xn_l[N1], yn_l[N1];
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  const int d = face_dof<N1>(f ^ 1, k);
                  zn_l[k] = src_h[vix(b2, 1, 0, ND, d, l2)];
                  xn_l[k] = src_q[vix(b2, 2, 0, ND, d, l2)];
                  yn_l[k] = src_q[vix(b2, 2, 1, ND, d, l2)];
                }
              const double len = (kind == FK_COARSE) ? 0.5 * flen : flen;
#pragma unroll 1
              for (int a = 0; a < NG; ++a)
                {
                  double mo[N1];
                  double zo = 0, xo = 0, yo = 0, zn = 0, xn = 0, yn = 0;
#pragma unroll
                  for (int k = 0; k < N1; ++k)
                    {
                      mo[k] = (kind == FK_COARSE) ? T.Sgs[s][a][k] : T.Sg[a][k];
                      const double mn =
                        (kind == FK_FINE) ? (nsub ? T.Sgs[1][a][k] : T.Sgs[0][a][k]) : T.Sg[a][k];
                      zo += mo[k] * zo_l[k];
                      xo += mo[k] * xo_l[k];
                      yo += mo[k] * yo_l[k];
                      zn += mn * zn_l[k];
                      xn += mn * xn_l[k];
                      yn += mn * yn_l[k];
                    }
                  const double zbm = fld(O.facez, ((size_t)fi * NG + a) * 2 + 0);
                  const double zbp = fld(O.facez, ((size_t)fi * NG + a) * 2 + 1);
                  double val[NC];
                  inner(is_m, zo, Vec2<double>{xo, yo}, zn, Vec2<double>{xn, yn}, n, zbm, zbp,
                        val);
                  const double JxW = T.wg[a] * len;
#pragma unroll
                  for (int cc = 0; cc < NC; ++cc)
                    {
                      const double v = val[cc] * JxW;
#pragma unroll
                      for (int k = 0; k < N1; ++k)
                        R[cc][face_dof<N1>(f, k)] += mo[k] * v;
                    }
                }
            }
        }
    }
}

// in place: U <- M^-1 U for one component on an affine cell of area `area`
template <int P>
__device__ __forceinline__ void tensor_inverse(const DegreeTables &T, double *U, double inv_area)
{
  constexpr int N1 = P + 1;
  double X[N1 * N1], Y[N1 * N1];
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      {
        double s = 0;
#pragma unroll
        for (int i = 0; i < N1; ++i)
          s += T.Ti[i][a] * U[i + N1 * j];
        X[a + N1 * j] = s;
      }
#pragma unroll
  for (int b = 0; b < N1; ++b)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      {
        double s = 0;
#pragma unroll
        for (int j = 0; j < N1; ++j)
          s += T.Ti[j][b] * X[a + N1 * j];
        Y[a + N1 * b] = s * (inv_area / (T.wn[a] * T.wn[b]));
      }
#pragma unroll
  for (int b = 0; b < N1; ++b)
#pragma unroll
    for (int i = 0; i < N1; ++i)
      {
        double s = 0;
#pragma unroll
        for (int a = 0; a < N1; ++a)
          s += T.Ti[i][a] * Y[a + N1 * b];
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
        U[i + N1 * j] = s;
      }
}

// ---------------------------------------------------------------------------------------
// L_zeta of one cell (cell term + faces) into R; Z, QX, QY: the cell's own data
// ---------------------------------------------------------------------------------------
template <int P, typename ZA, typename QA>
__device__ __forceinline__ void height_operator(const OpArgs &O, const DegreeTables &T,
                                                const Phys<double> &M, int batch, int lane,
                                                double hx, double hy, const ZA &Z, const QA &QX,
                                                const QA &QY, const double *__restrict__ src_h,
                                                const double *__restrict__ src_q,
                                                double (&R)[1][(P + 1) * (P + 1)])
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P);
  const double ihx = 1.0 / hx, ihy = 1.0 / hy, area = hx * hy;
#pragma unroll
  for (int k = 0; k < ND; ++k)
    R[0][k] = 0.0;
  // cell term: integral of grad(phi) . q  (mass_flux = q, ShallowWater.h:352-357)
#pragma unroll 1
  for (int b = 0; b < NQ; ++b)
    {
      double xs[N1], ys[N1], WS[N1], WD[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          double sx = 0, sy = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            {
              sx += T.Sq[b][j] * QX[i + N1 * j];
              sy += T.Sq[b][j] * QY[i + N1 * j];
            }
          xs[i] = sx;
          ys[i] = sy;
          WS[i] = 0;
          WD[i] = 0;
        }
#pragma unroll
      for (int a = 0; a < NQ; ++a)
        {
          double qx = 0, qy = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              qx += T.Sq[a][i] * xs[i];
              qy += T.Sq[a][i] * ys[i];
            }
          const double JxW = T.wq[a] * T.wq[b] * area;
          const double gx = qx * ihx * JxW, gy = qy * ihy * JxW;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              WS[i] += T.Dq[a][i] * gx;
              WD[i] += T.Sq[a][i] * gy;
            }
        }
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          R[0][i + N1 * j] += T.Sq[b][j] * WS[i] + T.Dq[b][j] * WD[i];
    }
  phase_barrier();
  face_terms<P, 1>(
    O, T, batch, lane, hx, hy, Z, QX, QY, src_h, src_q, R,
    [&](bool is_m, double zo, Vec2<double> qo, double zn, Vec2<double> qn, Vec2<double> n,
        double zbm, double zbp, double *val) {
      if (is_m)
        val[0] = -M.mass_flux(zo, zn, qo, qn, n, zbm, zbp);
      else
        val[0] = M.mass_flux(zn, zo, qn, qo, n, zbm, zbp);
    },
    [&](int bk, double zo, Vec2<double> qo, Vec2<double> n, double zb, const double *bv,
        double *val) {
      double zp;
      Vec2<double> qp;
      boundary_state(M, bk, zo, qo, n, zb, bv, zp, qp);
      val[0] = -M.mass_flux(zo, zp, qo, qp, n, zb, zb);
    });
}

// L_zeta (INVERT: M^-1 L_zeta, for apply())
template <int P, bool INVERT>
__global__ __launch_bounds__(CB) void k_height_rhs(OpArgs O, const double *__restrict__ src_h,
                                                   const double *__restrict__ src_q,
                                                   double *__restrict__ dst)
{
  constexpr int ND = (P + 1) * (P + 1);
  const DegreeTables &T = c_tab[P - 1];
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x;
  if (batch * CB + lane >= O.n_cells)
    return;
  const Phys<double> M{O.g, O.cu4, O.nu};
  const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
  const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
  double Z[ND], QX[ND], QY[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      Z[k] = src_h[vix(batch, 1, 0, ND, k, lane)];
      QX[k] = src_q[vix(batch, 2, 0, ND, k, lane)];
      QY[k] = src_q[vix(batch, 2, 1, ND, k, lane)];
    }
  double R[1][ND];
  height_operator<P>(O, T, M, batch, lane, hx, hy, (const double *)Z, (const double *)QX,
                     (const double *)QY, src_h, src_q, R);
  if (INVERT)
    tensor_inverse<P>(T, R[0], 1.0 / (hx * hy));
#pragma unroll
  for (int k = 0; k < ND; ++k)
    dst[vix(batch, 1, 0, ND, k, lane)] = R[0][k];
}

// ---------------------------------------------------------------------------------------
// L_q of one cell; MODE 0: stage (M^-1, RK update), 1: apply (M^-1 -> dst), 2: rhs
// The cell's own zeta, q (and velocity) are staged in LDS, lane-interleaved.
// ---------------------------------------------------------------------------------------
template <int P, bool VISC>
constexpr int discharge_lds_doubles()
{
  return (VISC ? 5 : 3) * (P + 1) * (P + 1) * CB;
}

template <int P, bool VISC, int MODE>
__device__ __forceinline__ void discharge_block(const OpArgs &O, const StageArgs &S, int batch,
                                                double *lds)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P);
  const DegreeTables &T = c_tab[P - 1];
  const int lane = threadIdx.x;
  const double *__restrict__ src_h = S.src_h;
  const double *__restrict__ src_q = S.src_q;
  // stage the batch's own data: contiguous, coalesced
  {
    const double *gh = src_h + (size_t)batch * ND * CB;
    const double *gq = src_q + (size_t)batch * 2 * ND * CB;
#pragma unroll
    for (int k = 0; k < ND; ++k)
      {
        lds[k * CB + lane] = gh[k * CB + lane];
        lds[(ND + k) * CB + lane] = gq[k * CB + lane];
        lds[(2 * ND + k) * CB + lane] = gq[(ND + k) * CB + lane];
        if (VISC)
          {
            const double *gv = S.src_v + (size_t)batch * 2 * ND * CB;
            lds[(3 * ND + k) * CB + lane] = gv[k * CB + lane];
            lds[(4 * ND + k) * CB + lane] = gv[(ND + k) * CB + lane];
          }
      }
  }
  // each lane reads only its own column of the staged data: no barrier needed beyond the
  // wave's own ordering of its LDS writes and reads
  __builtin_amdgcn_wave_barrier();
  if (batch * CB + lane >= O.n_cells)
    return;
  const LdsVec Z{lds + lane}, QX{lds + ND * CB + lane}, QY{lds + 2 * ND * CB + lane};
  const LdsVec UX{lds + 3 * ND * CB + lane}, UY{lds + 4 * ND * CB + lane};
  const Phys<double> M{O.g, O.cu4, O.nu};
  const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
  const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
  const double ihx = 1.0 / hx, ihy = 1.0 / hy, area = hx * hy;
  double R[2][ND];
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      R[0][k] = 0.0;
      R[1][k] = 0.0;
    }
  const size_t c0base = (size_t)batch * NQ * NQ * CB + lane;
#pragma unroll 1
  for (int b = 0; b < NQ; ++b)
    {
      double zs[N1], zd[N1], xs[N1], ys[N1];
      double uxs[VISC ? N1 : 1], uxd[VISC ? N1 : 1], uys[VISC ? N1 : 1], uyd[VISC ? N1 : 1];
      double WSX[N1], WDX[N1], WSY[N1], WDY[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            {
              const double zv = Z[i + N1 * j];
              s0 += T.Sq[b][j] * zv;
              s1 += T.Dq[b][j] * zv;
              s2 += T.Sq[b][j] * QX[i + N1 * j];
              s3 += T.Sq[b][j] * QY[i + N1 * j];
            }
          zs[i] = s0;
          zd[i] = s1;
          xs[i] = s2;
          ys[i] = s3;
          if (VISC)
            {
              double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
              for (int j = 0; j < N1; ++j)
                {
                  t0 += T.Sq[b][j] * UX[i + N1 * j];
                  t1 += T.Dq[b][j] * UX[i + N1 * j];
                  t2 += T.Sq[b][j] * UY[i + N1 * j];
                  t3 += T.Dq[b][j] * UY[i + N1 * j];
                }
              uxs[i] = t0;
              uxd[i] = t1;
              uys[i] = t2;
              uyd[i] = t3;
            }
          WSX[i] = WDX[i] = WSY[i] = WDY[i] = 0;
        }
#pragma unroll
      for (int a = 0; a < NQ; ++a)
        {
          double z = 0, zx = 0, zy = 0, qx = 0, qy = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              z += T.Sq[a][i] * zs[i];
              zx += T.Dq[a][i] * zs[i];
              zy += T.Sq[a][i] * zd[i];
              qx += T.Sq[a][i] * xs[i];
              qy += T.Sq[a][i] * ys[i];
            }
          zx *= ihx;
          zy *= ihy;
          const int q = a + NQ * b;
          const size_t pi = c0base + (size_t)q * CB;
          const double zb = fld(O.cell0[0], pi);
          const double h = M.depth(z, zb);
          const double s = M.inv_depth(h);
          const double vx = qx * s, vy = qy * s;
          double F00 = qx * vx, F01 = qx * vy, F10 = qy * vx, F11 = qy * vy;
          if (VISC)
            {
              double G00 = 0, G01 = 0, G10 = 0, G11 = 0;
#pragma unroll
              for (int i = 0; i < N1; ++i)
                {
                  G00 += T.Dq[a][i] * uxs[i];
                  G01 += T.Sq[a][i] * uxd[i];
                  G10 += T.Dq[a][i] * uys[i];
                  G11 += T.Sq[a][i] * uyd[i];
                }
              G00 = h * (G00 * ihx);
              G01 = h * (G01 * ihy);
              G10 = h * (G10 * ihx);
              G11 = h * (G11 * ihy);
              const double div = G00 + G11;
              F00 = F00 - O.nu * (G00 + G00) + O.nu * div;
              F01 = F01 - O.nu * (G01 + G10);
              F10 = F10 - O.nu * (G10 + G01);
              F11 = F11 - O.nu * (G11 + G11) + O.nu * div;
            }
          const double mgh = -O.g * h;
          double Sx = mgh * zx, Sy = mgh * zy;
          if (O.has_friction)
            {
              const double n_m = fld(O.cell0[1], pi);
              const double inv_d = 1. / (exp(log(h) * 0.33333333333));
              const double hcf = O.g * n_m * n_m * inv_d * sqrt(vx * vx + vy * vy);
              Sx -= hcf * vx;
              Sy -= hcf * vy;
            }
          if (O.has_wind)
            {
              Sx += fld(O.cell0[2], pi);
              Sy += fld(O.cell0[3], pi);
            }
          if (O.has_coriolis)
            {
              const double fc = fld(O.cell0[4], pi);
              Sx += fc * qy;
              Sy += -fc * qx;
            }
          const double JxW = T.wq[a] * T.wq[b] * area;
          const double g0x = F00 * ihx * JxW, g0y = F01 * ihy * JxW;
          const double g1x = F10 * ihx * JxW, g1y = F11 * ihy * JxW;
          const double v0 = Sx * JxW, v1 = Sy * JxW;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              WSX[i] += T.Sq[a][i] * v0 + T.Dq[a][i] * g0x;
              WDX[i] += T.Sq[a][i] * g0y;
              WSY[i] += T.Sq[a][i] * v1 + T.Dq[a][i] * g1x;
              WDY[i] += T.Sq[a][i] * g1y;
            }
        }
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            R[0][i + N1 * j] += T.Sq[b][j] * WSX[i] + T.Dq[b][j] * WDX[i];
            R[1][i + N1 * j] += T.Sq[b][j] * WSY[i] + T.Dq[b][j] * WDY[i];
          }
    }
  phase_barrier();
  face_terms<P, 2>(
    O, T, batch, lane, hx, hy, Z, QX, QY, src_h, src_q, R,
    [&](bool is_m, double zo, Vec2<double> qo, double zn, Vec2<double> qn, Vec2<double> n,
        double zbm, double zbp, double *val) {
      if (is_m)
        {
          const Vec2<double> F = M.adv_flux(zo, zn, qo, qn, n, zbm, zbp);
          const Vec2<double> Pm = M.pres_flux(zo, zn, n, zbm, zbp);
          val[0] = -F.x - Pm.x;
          val[1] = -F.y - Pm.y;
        }
      else
        {
          const Vec2<double> F = M.adv_flux(zn, zo, qn, qo, n, zbm, zbp);
          const Vec2<double> Pp = M.pres_flux(zo, zn, n, zbp, zbm);
          val[0] = F.x + Pp.x;
          val[1] = F.y + Pp.y;
        }
    },
    [&](int bk, double zo, Vec2<double> qo, Vec2<double> n, double zb, const double *bv,
        double *val) {
      double zp;
      Vec2<double> qp;
      const bool at_outflow = boundary_state(M, bk, zo, qo, n, zb, bv, zp, qp);
      Vec2<double> F = M.adv_flux(zo, zp, qo, qp, n, zb, zb);
      const Vec2<double> Pr = M.pres_flux(zo, zp, n, zb, zb);
      F.x += Pr.x;
      F.y += Pr.y;
      if (at_outflow && (qo.x * n.x + qo.y * n.y) < -1e-12)
        F = {0.0, 0.0};
      val[0] = -F.x;
      val[1] = -F.y;
    });
  if (MODE == 2)
    {
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          S.dst[vix(batch, 2, c, ND, k, lane)] = R[c][k];
      return;
    }
  if (MODE == 0 && S.scratch)
    {
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          S.scratch[vix(batch, 2, c, ND, k, lane)] = R[c][k];
    }
  phase_barrier();
  const double inv_area = 1.0 / area;
  tensor_inverse<P>(T, R[0], inv_area);
  tensor_inverse<P>(T, R[1], inv_area);
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < ND; ++k)
      S.dst[vix(batch, 2, c, ND, k, lane)] = R[c][k];
  if (MODE == 0)
    {
      // post-op of OceanDG.h:1878-1909: dst = w + sum_j a_j k_{j+1}, k_{stage+1} = R
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          {
            const size_t idx = vix(batch, 2, c, ND, k, lane);
            double acc = S.wsol[idx] + S.factor[0] * (S.stage == 0 ? R[c][k] : S.kprev[0][idx]);
            for (int j = 1; j <= S.stage; ++j)
              acc += S.factor[j] * (j == S.stage ? R[c][k] : S.kprev[j][idx]);
            S.rk_dst[idx] = acc;
          }
    }
}

template <int P, bool VISC, int MODE>
__global__ __launch_bounds__(CB) void k_discharge(OpArgs O, StageArgs S)
{
  __shared__ double lds[discharge_lds_doubles<P, VISC>()];
  discharge_block<P, VISC, MODE>(O, S, logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle),
                                 lds);
}

// ---------------------------------------------------------------------------------------
// front + capped lumped-Jacobi depth-mass solve
// ---------------------------------------------------------------------------------------
// R += sgn * integral(phi * depth(Z)) at GL(ng)^2; optionally counts dry points
template <int P, bool COUNT>
__device__ __forceinline__ int depth_mass(const OpArgs &O, const DegreeTables &T,
                                          const Phys<double> &M, int batch, int lane,
                                          const double *Z, double *R, double sgn, double area,
                                          unsigned long long &drybits)
{
  constexpr int N1 = P + 1, NG = P + 2;
  const size_t zbase = (size_t)batch * NG * NG * CB + lane;
  int ndry = 0;
#pragma unroll 1
  for (int b = 0; b < NG; ++b)
    {
      double zs[N1], W[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          double s = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            s += T.Sg[b][j] * Z[i + N1 * j];
          zs[i] = s;
          W[i] = 0;
        }
#pragma unroll
      for (int a = 0; a < NG; ++a)
        {
          double z = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            z += T.Sg[a][i] * zs[i];
          const int q = a + NG * b;
          const double zb = fld(O.cell1, zbase + (size_t)q * CB);
          if (COUNT && z <= -zb)
            {
              ++ndry;
              drybits |= 1ull << q;
            }
          const double v = sgn * M.depth(z, zb) * (T.wg[a] * T.wg[b] * area);
#pragma unroll
          for (int i = 0; i < N1; ++i)
            W[i] += T.Sg[a][i] * v;
        }
#pragma unroll
      for (int j = 0; j < N1; ++j)
#pragma unroll
        for (int i = 0; i < N1; ++i)
          R[i + N1 * j] += T.Sg[b][j] * W[i];
    }
  return ndry;
}

// Jacobi solve of one cell: Z holds zeta^n on entry (and the dry lift is applied to it),
// F the front; returns the number of sweeps.  OceanDG.h:1374-1507.
template <int P>
__device__ __forceinline__ int jacobi_solve(const OpArgs &O, const DegreeTables &T,
                                            const Phys<double> &M, int batch, int lane,
                                            double area, const double *F, double *Z)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NG = P + 2, NP = NG * NG;
  // dry-cell lift (OceanDG.h:1399-1422)
  {
    const size_t zbase = (size_t)batch * NP * CB + lane;
    double zbmax = -DBL_MAX;
    int ndry = 0;
#pragma unroll 1
    for (int b = 0; b < NG; ++b)
      {
        double zs[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            double s = 0;
#pragma unroll
            for (int j = 0; j < N1; ++j)
              s += T.Sg[b][j] * Z[i + N1 * j];
            zs[i] = s;
          }
#pragma unroll
        for (int a = 0; a < NG; ++a)
          {
            double z = 0;
#pragma unroll
            for (int i = 0; i < N1; ++i)
              z += T.Sg[a][i] * zs[i];
            const double zb = fld(O.cell1, zbase + (size_t)(a + NG * b) * CB);
            if (z <= -zb)
              {
                zbmax = zb > zbmax ? zb : zbmax;
                ++ndry;
              }
          }
      }
    if (ndry == NP)
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Z[i] = -zbmax + 1e-09;
  }
  double norm = 1.0;
  int k = 0;
  for (; k < O.max_it && norm > 1e-12; ++k)
    {
      double Rr[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Rr[i] = F[i];
      unsigned long long bits = 0;
      const int ndry = depth_mass<P, true>(O, T, M, batch, lane, Z, Rr, -1.0, area, bits);
      double D[ND];
      if (ndry == 0)
        {
#pragma unroll
          for (int j = 0; j < N1; ++j)
#pragma unroll
            for (int i = 0; i < N1; ++i)
              D[i + N1 * j] = T.lumped[i] * T.lumped[j] * area;
        }
      else
        {
#pragma unroll
          for (int i = 0; i < ND; ++i)
            D[i] = 0;
#pragma unroll 1
          for (int b = 0; b < NG; ++b)
            {
              double W[N1];
#pragma unroll
              for (int i = 0; i < N1; ++i)
                W[i] = 0;
#pragma unroll
              for (int a = 0; a < NG; ++a)
                {
                  const double mask = ((bits >> (a + NG * b)) & 1ull) ? 0.0 : 1.0;
                  const double v = mask * (T.wg[a] * T.wg[b] * area);
#pragma unroll
                  for (int i = 0; i < N1; ++i)
                    W[i] += T.Sg[a][i] * v;
                }
#pragma unroll
              for (int j = 0; j < N1; ++j)
#pragma unroll
                for (int i = 0; i < N1; ++i)
                  D[i + N1 * j] += T.Sg[b][j] * W[i];
            }
        }
      norm = 0.0;
#pragma unroll
      for (int i = 0; i < ND; ++i)
        {
          const double dz = (ndry < NP) ? (1. / D[i]) * Rr[i] : 0.0;
          const double old = Z[i];
          Z[i] += dz;
          norm += fabs(old - Z[i]);
        }
    }
  return k;
}

__device__ __forceinline__ void store_sweeps(unsigned int *blk, int batch, int k)
{
  if (!blk)
    return;
  unsigned int v = (unsigned int)k;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off);
  if (threadIdx.x == 0)
    blk[batch] = v;
}

// front from ki_h[1..stage+1] (all in memory) and the Jacobi solve: the path used when
// the stage's outputs alias its inputs
template <int P>
__global__ __launch_bounds__(CB) void k_height_jacobi(OpArgs O, StageArgs S)
{
  constexpr int N1 = P + 1, ND = N1 * N1;
  const DegreeTables &T = c_tab[P - 1];
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x;
  const bool active = batch * CB + lane < O.n_cells;
  const Phys<double> M{O.g, O.cu4, O.nu};
  int k = 0;
  if (active)
    {
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      const double area = hx * hy;
      double F[ND], Z[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i)
        {
          const size_t idx = vix(batch, 1, 0, ND, i, lane);
          double f = S.factor[0] * S.kprev[0][idx];
          for (int j = 1; j <= S.stage; ++j)
            f += S.factor[j] * S.kprev[j][idx];
          F[i] = f;
          Z[i] = S.wsol[idx];
        }
      unsigned long long bits = 0;
      depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits);
      if (S.scratch)
#pragma unroll
        for (int i = 0; i < ND; ++i)
          S.scratch[vix(batch, 1, 0, ND, i, lane)] = F[i];
      k = jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
#pragma unroll
      for (int i = 0; i < ND; ++i)
        S.rk_dst[vix(batch, 1, 0, ND, i, lane)] = Z[i];
    }
  (void)N1;
  store_sweeps(S.sweeps_blk, batch, k);
}

// L_zeta, front and Jacobi solve of one batch of cells in one pass: L_zeta stays in
// registers for the front (it is also written out as ki_h[stage+1] for later stages)
template <int P>
__device__ __forceinline__ void height_stage_block(const OpArgs &O, const StageArgs &S,
                                                   int batch)
{
  constexpr int ND = (P + 1) * (P + 1);
  const DegreeTables &T = c_tab[P - 1];
  const int lane = threadIdx.x;
  const bool active = batch * CB + lane < O.n_cells;
  const Phys<double> M{O.g, O.cu4, O.nu};
  int k = 0;
  if (active)
    {
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      const double area = hx * hy;
      double F[ND];
      {
        double Z[ND], QX[ND], QY[ND];
#pragma unroll
        for (int i = 0; i < ND; ++i)
          {
            Z[i] = S.src_h[vix(batch, 1, 0, ND, i, lane)];
            QX[i] = S.src_q[vix(batch, 2, 0, ND, i, lane)];
            QY[i] = S.src_q[vix(batch, 2, 1, ND, i, lane)];
          }
        double R[1][ND];
        height_operator<P>(O, T, M, batch, lane, hx, hy, (const double *)Z,
                           (const double *)QX, (const double *)QY, S.src_h, S.src_q, R);
#pragma unroll
        for (int i = 0; i < ND; ++i)
          {
            const size_t idx = vix(batch, 1, 0, ND, i, lane);
            S.dst[idx] = R[0][i];  // ki_h[stage+1]
            const double kk = R[0][i];
            double f = S.factor[0] * (S.stage == 0 ? kk : S.kprev[0][idx]);
            for (int j = 1; j <= S.stage; ++j)
              f += S.factor[j] * (j == S.stage ? kk : S.kprev[j][idx]);
            F[i] = f;
          }
      }
      phase_barrier();
      double Z[ND];
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Z[i] = S.wsol[vix(batch, 1, 0, ND, i, lane)];
      unsigned long long bits = 0;
      depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits);
      if (S.scratch)
#pragma unroll
        for (int i = 0; i < ND; ++i)
          S.scratch[vix(batch, 1, 0, ND, i, lane)] = F[i];
      k = jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
#pragma unroll
      for (int i = 0; i < ND; ++i)
        S.rk_dst[vix(batch, 1, 0, ND, i, lane)] = Z[i];
    }
  store_sweeps(S.sweeps_blk, batch, k);
}

// the whole stage in one launch: work items w = (batch, kind) are dealt so that each XCD
// holds a contiguous range of batches; kind 0 = height stage, kind 1 = discharge stage
template <int P, bool VISC>
__global__ __launch_bounds__(CB) void k_stage(OpArgs O, StageArgs H, StageArgs Q)
{
  __shared__ double lds[discharge_lds_doubles<P, VISC>()];
  const int w = logical_batch(blockIdx.x, 2 * O.n_batches, O.xcd_swizzle);
  if (w & 1)
    discharge_block<P, VISC, 0>(O, Q, w >> 1, lds);
  else
    height_stage_block<P>(O, H, w >> 1);
}

}  // namespace oceano
