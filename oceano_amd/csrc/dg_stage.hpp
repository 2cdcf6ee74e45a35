// dg_stage.hpp -- gfx950 kernels of one RK stage, FP64 or mixed precision.
//
//   height_operator   L_zeta: local_apply_cell_height + face_height + boundary_face_height
//                     (OceanDG.h:670-693, 937-982, 1106-1214)
//   discharge_block   L_q: local_apply_cell_discharge + face_discharge +
//                     boundary_face_discharge (OceanDG.h:696-743, 985-1050, 1217-1339),
//                     fused with the tensor-product inverse mass (1533-1571) and the
//                     post-loop RK update (1878-1909)
//   jacobi_solve      front = sum_j a_j L_zeta,j + M h(zeta^n) (829-855, 1835-1847) and the
//                     capped lumped-Jacobi depth-mass solve (1374-1507)
//   k_stage           the whole stage in one launch (height and discharge work items)
//
// R is the arithmetic type of the cell and face integrals: double reproduces the
// reference arithmetic; float is the mixed-precision path.  In the mixed path the state,
// the RK combinations, the bathymetry data and the depth solve of any cell that is not
// safely wet stay FP64; the free surface enters the FP32 integrals as a deviation from
// the cell's first nodal value, so a flat free surface gives an exactly zero gradient and
// an exactly zero pressure jump (the lake at rest stays exactly balanced).
//
// This header expects the including translation unit to define
//   static __constant__ Basis<double> c_tab[4];  static __constant__ Basis<float> c_tabf[4];
#pragma once

#include <cfloat>
#include <type_traits>

#include "dg_kernels.hpp"
#include "dg_mfma.hpp"

namespace oceano {

template <typename R>
__device__ __forceinline__ const Basis<R> &basis(int p);
template <>
__device__ __forceinline__ const Basis<double> &basis<double>(int p)
{
  return c_tab[p - 1];
}
template <>
__device__ __forceinline__ const Basis<float> &basis<float>(int p)
{
  return c_tabf[p - 1];
}

template <typename R>
constexpr bool centred()
{
  return std::is_same<R, float>::value;
}

// exterior state on a boundary face (OceanDG.h:1130-1203 / 1246-1317); returns at_outflow
template <typename T>
__device__ __forceinline__ bool boundary_state(const Phys<T> &M, int kind, T zm, Vec2<T> qm,
                                               Vec2<T> n, T zb, const double *bv, T &zp,
                                               Vec2<T> &qp)
{
  const T un = dot_n(qm, n);
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

// own-cell data in the lane-interleaved LDS copy
template <typename R>
struct LdsVec
{
  const R *p;  // &lds[component base + lane]
  __device__ __forceinline__ R operator[](int k) const { return p[k * CB]; }
};

// position of the j-th (0-based) set bit of m; m has more than j set bits
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int j)
{
  int pos = 0;
#pragma unroll
  for (int w = 32; w >= 1; w >>= 1)
    {
      const int c = __popcll(m & ((1ull << w) - 1));
      if (j >= c)
        {
          j -= c;
          pos += w;
          m >>= w;
        }
    }
  return pos;
}

// nodal values of a trace (N1 values per field) restricted to half `half` of the face:
// u_j <- sum_k Psub[half][j][k] u_k
template <int N1, typename R>
__device__ __forceinline__ void restrict_half(const Basis<R> &T, int half, R *z, R *x, R *y)
{
  R zt[N1], xt[N1], yt[N1];
#pragma unroll
  for (int j = 0; j < N1; ++j)
    {
      R sz = 0, sx = 0, sy = 0;
#pragma unroll
      for (int k = 0; k < N1; ++k)
        {
          sz += T.Psub[half][j][k] * z[k];
          sx += T.Psub[half][j][k] * x[k];
          sy += T.Psub[half][j][k] * y[k];
        }
      zt[j] = sz;
      xt[j] = sx;
      yt[j] = sy;
    }
#pragma unroll
  for (int j = 0; j < N1; ++j)
    {
      z[j] = zt[j];
      x[j] = xt[j];
      y[j] = yt[j];
    }
}

// integral over one (sub)face of NG Gauss-Lobatto points: both traces as nodal values on
// the (sub)face, deviations from zcr / zc2r; acc[cc][k] = sum_a Sg[a][k] inner(..)_cc JxW_a
template <int P, int NC, typename R, typename InnerFn>
__device__ __forceinline__ void subface_integral(const Basis<R> &T, const FieldV<R> &fz, int fi,
                                                 R len, bool is_m, Vec2<R> n, const R *zo_s,
                                                 const R *xo_s, const R *yo_s, const R *zn_s,
                                                 const R *xn_s, const R *yn_s, R zcr, R zc2r,
                                                 InnerFn &inner, R (&acc)[NC][P + 1])
{
  constexpr int N1 = P + 1, NG = P + 2;
#pragma unroll
  for (int cc = 0; cc < NC; ++cc)
#pragma unroll
    for (int k = 0; k < N1; ++k)
      acc[cc][k] = 0;
#pragma unroll 1
  for (int a = 0; a < NG; ++a)
    {
      R zo = 0, xo = 0, yo = 0, zn = 0, xn = 0, yn = 0;
#pragma unroll
      for (int k = 0; k < N1; ++k)
        {
          const R m = T.Sg[a][k];
          zo += m * zo_s[k];
          xo += m * xo_s[k];
          yo += m * yo_s[k];
          zn += m * zn_s[k];
          xn += m * xn_s[k];
          yn += m * yn_s[k];
        }
      zo += zcr;
      zn += zc2r;
      const R zbm = fld(fz, ((size_t)fi * NG + a) * 2 + 0);
      const R zbp = fld(fz, ((size_t)fi * NG + a) * 2 + 1);
      R val[NC];
      inner(is_m, zo, Vec2<R>{xo, yo}, zn, Vec2<R>{xn, yn}, n, zbm, zbp, val);
      const R JxW = T.wg[a] * len;
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
        {
          const R v = val[cc] * JxW;
#pragma unroll
          for (int k = 0; k < N1; ++k)
            acc[cc][k] += T.Sg[a][k] * v;
        }
    }
}

// Face terms of one cell: for every face point, `inner` / `bdry` return NC values that are
// multiplied by JxW and integrated against this cell's face basis into Racc[NC][nd].
// Own data: Z (free-surface deviation from zc), QX, QY; neighbour traces are gathered
// from the FP64 vectors and centred on the neighbour's first nodal value.
//
// The lanes of a wave are 64 cells.  A coarse face (two fine neighbours) has a second
// subface that only its own cell integrates; in the per-face loop a wave in which any cell
// has such a face runs that second subface for all 64 lanes (on a mesh with 11 % of its
// faces hanging and cells in the caller's order, every wave does, on every face).  With
// COMPACT (the FP32 path) the second subfaces of the whole wave are instead listed after
// the face loop, face-major, and dealt one per lane: a lane integrates another cell's
// second subface from the FP64 vectors (the same values that cell holds) and hands the
// result back by a cross-lane read.  Each cell adds its four faces' second-subface
// contributions in face order, so the result does not depend on which lane computed them.
// The active lanes of a wave are a prefix (cells batch*64 + lane < n_cells), which is what
// lets lane i take item i.  The FP64 path (one wave per SIMD, cells sorted by face
// signature so that waves rarely diverge) cannot hide the compacted pass's extra load
// latency and keeps the second subface in the face loop.
//
// FMASK selects the faces integrated (bit f: face f); the other faces contribute nothing.
template <int P, int NC, typename R, int FMASK = 0xF, typename ZA, typename QA, typename InnerFn,
          typename BdryFn>
__device__ __forceinline__ void face_terms(const OpArgs &O, const Basis<R> &T, int batch,
                                           int lane, R hx, R hy, const ZA &Z, const QA &QX,
                                           const QA &QY, double zc,
                                           const double *__restrict__ src_h,
                                           const double *__restrict__ src_q,
                                           R (&Racc)[NC][(P + 1) * (P + 1)], InnerFn inner,
                                           BdryFn bdry, bool in32 = false)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P), NG = P + 2;
  constexpr bool COMPACT = std::is_same<R, float>::value;
  constexpr int NS = COMPACT ? 1 : 2;  // subfaces integrated in the face loop
  const R zcr = R(zc);
  // stage input in its centred FP32 form (FP32 stage only; c32_values, c32_centres)
  const bool c32 = centred<R>() && in32;
  const float *__restrict__ fh = c32_values(src_h);
  const float *__restrict__ fq = c32_values(src_q);
  const double *__restrict__ zcen = c32_centres(src_h, O.n_batches, ND);
  // the face tables of all four faces in one group of loads: each face then waits for one
  // memory latency (its neighbour's traces) instead of three dependent ones
  int infos[4] = {}, nbrs[4][NS] = {}, fids[4][NS] = {};
#pragma unroll
  for (int f = 0; f < 4; ++f)
    {
      if (!((FMASK >> f) & 1))
        continue;
      const size_t fo = ((size_t)batch * 4 + f);
      infos[f] = O.finfo[fo * CB + lane];
#pragma unroll
      for (int s = 0; s < NS; ++s)
        {
          nbrs[f][s] = O.nbr[(fo * 2 + s) * CB + lane];
          fids[f][s] = O.fid[(fo * 2 + s) * CB + lane];
        }
    }
  issue_barrier();
  unsigned long long coarse[4];  // lanes whose face f is the coarse side of a hanging face
#pragma unroll
  for (int f = 0; f < 4; ++f)
    coarse[f] = COMPACT && ((FMASK >> f) & 1) ? __ballot((infos[f] & 3) == FK_COARSE) : 0ull;
#pragma unroll
  for (int f = 0; f < 4; ++f)
    {
      if (!((FMASK >> f) & 1))
        continue;
      phase_barrier();
      const int info = infos[f];
      const int kind = info & 3;
      const int nsub = (info >> 2) & 1;
      const R flen = (f < 2) ? hy : hx;
      R zo_l[N1], xo_l[N1], yo_l[N1];
#pragma unroll
      for (int k = 0; k < N1; ++k)
        {
          zo_l[k] = Z[face_dof<N1>(f, k)];
          xo_l[k] = QX[face_dof<N1>(f, k)];
          yo_l[k] = QY[face_dof<N1>(f, k)];
        }
      if (kind == FK_BDRY)
        {
          const int bf = fids[f][0];
          const int bk = O.bkind[bf];
          const Vec2<R> n = {f == 0 ? R(-1) : f == 1 ? R(1) : R(0),
                             f == 2 ? R(-1) : f == 3 ? R(1) : R(0)};
          const FieldV<R> bz = in_vgprs<R>(O.bdryz);
          const double *bvb = stage_bvals(O);
#pragma unroll 1
          for (int a = 0; a < NQ; ++a)
            {
              R zo = 0, xo = 0, yo = 0;
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  zo += T.Sq[a][k] * zo_l[k];
                  xo += T.Sq[a][k] * xo_l[k];
                  yo += T.Sq[a][k] * yo_l[k];
                }
              zo += zcr;
              const R zb = fld(bz, (size_t)bf * NQ + a);
              const double *bv = bvb + ((size_t)bf * NQ + a) * 3;
              R val[NC];
              bdry(bk, zo, Vec2<R>{xo, yo}, n, zb, bv, val);
              const R JxW = T.wq[a] * flen;
#pragma unroll
              for (int cc = 0; cc < NC; ++cc)
                {
                  const R v = val[cc] * JxW;
#pragma unroll
                  for (int k = 0; k < N1; ++k)
                    Racc[cc][face_dof<N1>(f, k)] += T.Sq[a][k] * v;
                }
            }
        }
      else
        {
          const bool is_m = (f & 1);  // canonical orientation: minus side owns faces 1, 3
          const Vec2<R> n = {f < 2 ? R(1) : R(0), f < 2 ? R(0) : R(1)};
          const FieldV<R> fz = in_vgprs<R>(O.facez);
          // neighbour traces of the subfaces in one group of loads (the second only on the
          // coarse side, and only where it is integrated here): one memory latency per face
          phase_barrier();
          // zsub: what is subtracted from zg to centre it (0 when it comes centred)
          double zg[NS][N1], xg[NS][N1], yg[NS][N1], zc2[NS], zsub[NS];
#pragma unroll
          for (int s = 0; s < NS; ++s)
            if (s == 0 || kind == FK_COARSE)
              {
                const int b2 = nbrs[f][s] / CB, l2 = nbrs[f][s] % CB;
                if (c32)
                  {
                    zc2[s] = zcen[(size_t)b2 * CB + l2];
                    zsub[s] = 0.0;
#pragma unroll
                    for (int k = 0; k < N1; ++k)
                      {
                        const int d = face_dof<N1>(f ^ 1, k);
                        zg[s][k] = fh[vix(b2, 1, 0, ND, d, l2)];
                        xg[s][k] = fq[vix(b2, 2, 0, ND, d, l2)];
                        yg[s][k] = fq[vix(b2, 2, 1, ND, d, l2)];
                      }
                  }
                else
                  {
                    zc2[s] = centred<R>() ? src_h[vix(b2, 1, 0, ND, 0, l2)] : 0.0;
                    zsub[s] = zc2[s];
#pragma unroll
                    for (int k = 0; k < N1; ++k)
                      {
                        const int d = face_dof<N1>(f ^ 1, k);
                        zg[s][k] = src_h[vix(b2, 1, 0, ND, d, l2)];
                        xg[s][k] = src_q[vix(b2, 2, 0, ND, d, l2)];
                        yg[s][k] = src_q[vix(b2, 2, 1, ND, d, l2)];
                      }
                  }
              }
          issue_barrier();
#pragma unroll
          for (int s = 0; s < NS; ++s)
            {
              if (s == 1 && kind != FK_COARSE)
                break;
              phase_barrier();
              // both traces as nodal values on this (sub)face: the coarse side restricts
              // its own trace to half s, the fine side the neighbour's to half nsub, so
              // that every face point uses the one (wave-uniform) table Sg
              R zo_s[N1], xo_s[N1], yo_s[N1], zn_s[N1], xn_s[N1], yn_s[N1];
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  zo_s[k] = zo_l[k];
                  xo_s[k] = xo_l[k];
                  yo_s[k] = yo_l[k];
                  zn_s[k] = R(zg[s][k] - zsub[s]);
                  xn_s[k] = R(xg[s][k]);
                  yn_s[k] = R(yg[s][k]);
                }
              if (kind == FK_COARSE)
                restrict_half<N1>(T, s, zo_s, xo_s, yo_s);
              else if (kind == FK_FINE)
                {
                  if (nsub)
                    restrict_half<N1>(T, 1, zn_s, xn_s, yn_s);
                  else
                    restrict_half<N1>(T, 0, zn_s, xn_s, yn_s);
                }
              const R len = (kind == FK_COARSE) ? R(0.5) * flen : flen;
              R acc[NC][N1];
              subface_integral<P, NC, R>(T, fz, fids[f][s], len, is_m, n, zo_s, xo_s, yo_s,
                                         zn_s, xn_s, yn_s, zcr, R(zc2[s]), inner, acc);
              // test functions of the cell's face: the transposed restriction on the
              // coarse side
              if (kind == FK_COARSE)
                {
#pragma unroll
                  for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                    for (int k = 0; k < N1; ++k)
                      {
                        R t = 0;
#pragma unroll
                        for (int j = 0; j < N1; ++j)
                          t += T.Psub[s][j][k] * acc[cc][j];
                        Racc[cc][face_dof<N1>(f, k)] += t;
                      }
                }
              else
                {
#pragma unroll
                  for (int cc = 0; cc < NC; ++cc)
#pragma unroll
                    for (int k = 0; k < N1; ++k)
                      Racc[cc][face_dof<N1>(f, k)] += acc[cc][k];
                }
            }
        }
    }
  // second subfaces of the coarse faces, compacted over the wave (face-major item order)
  const int c0 = __popcll(coarse[0]), c1 = c0 + __popcll(coarse[1]);
  const int c2 = c1 + __popcll(coarse[2]), total = c2 + __popcll(coarse[3]);
  if (!COMPACT || total == 0)
    return;
  const int n_act = __popcll(__ballot(1));
  const int first[4] = {0, c0, c1, c2};
  for (int base = 0; base < total; base += n_act)
    {
      phase_barrier();
      R res[NC][N1];
#pragma unroll
      for (int cc = 0; cc < NC; ++cc)
#pragma unroll
        for (int k = 0; k < N1; ++k)
          res[cc][k] = 0;
      const int item = base + lane;
      if (item < total)
        {
          const int f = item < c0 ? 0 : item < c1 ? 1 : item < c2 ? 2 : 3;
          const unsigned long long m =
            f == 0 ? coarse[0] : f == 1 ? coarse[1] : f == 2 ? coarse[2] : coarse[3];
          const int owner = nth_set_bit(m, item - (f == 0 ? 0 : f == 1 ? c0 : f == 2 ? c1 : c2));
          const size_t fo = (size_t)batch * 4 + f;
          const int nb = O.nbr[(fo * 2 + 1) * CB + owner];
          const int fi = O.fid[(fo * 2 + 1) * CB + owner];
          const double hsz = O.hxy[((size_t)batch * 2 + (f < 2 ? 1 : 0)) * CB + owner];
          const int b2 = nb / CB, l2 = nb % CB;
          double zt[N1], xt[N1], yt[N1], zg[N1], xg[N1], yg[N1], zco, zc2, osub, nsub2;
          if (c32)
            {
              zco = zcen[(size_t)batch * CB + owner];
              zc2 = zcen[(size_t)b2 * CB + l2];
              osub = nsub2 = 0.0;
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  const int d = face_dof<N1>(f, k), e = face_dof<N1>(f ^ 1, k);
                  zt[k] = fh[vix(batch, 1, 0, ND, d, owner)];
                  xt[k] = fq[vix(batch, 2, 0, ND, d, owner)];
                  yt[k] = fq[vix(batch, 2, 1, ND, d, owner)];
                  zg[k] = fh[vix(b2, 1, 0, ND, e, l2)];
                  xg[k] = fq[vix(b2, 2, 0, ND, e, l2)];
                  yg[k] = fq[vix(b2, 2, 1, ND, e, l2)];
                }
            }
          else
            {
              zco = centred<R>() ? src_h[vix(batch, 1, 0, ND, 0, owner)] : 0.0;
              zc2 = centred<R>() ? src_h[vix(b2, 1, 0, ND, 0, l2)] : 0.0;
              osub = zco;
              nsub2 = zc2;
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  const int d = face_dof<N1>(f, k), e = face_dof<N1>(f ^ 1, k);
                  zt[k] = src_h[vix(batch, 1, 0, ND, d, owner)];
                  xt[k] = src_q[vix(batch, 2, 0, ND, d, owner)];
                  yt[k] = src_q[vix(batch, 2, 1, ND, d, owner)];
                  zg[k] = src_h[vix(b2, 1, 0, ND, e, l2)];
                  xg[k] = src_q[vix(b2, 2, 0, ND, e, l2)];
                  yg[k] = src_q[vix(b2, 2, 1, ND, e, l2)];
                }
            }
          issue_barrier();
          R zo_s[N1], xo_s[N1], yo_s[N1], zn_s[N1], xn_s[N1], yn_s[N1];
#pragma unroll
          for (int k = 0; k < N1; ++k)
            {
              zo_s[k] = R(zt[k] - osub);
              xo_s[k] = R(xt[k]);
              yo_s[k] = R(yt[k]);
              zn_s[k] = R(zg[k] - nsub2);
              xn_s[k] = R(xg[k]);
              yn_s[k] = R(yg[k]);
            }
          restrict_half<N1>(T, 1, zo_s, xo_s, yo_s);
          const Vec2<R> n = {f < 2 ? R(1) : R(0), f < 2 ? R(0) : R(1)};
          const FieldV<R> fz = in_vgprs<R>(O.facez);
          R acc[NC][N1];
          subface_integral<P, NC, R>(T, fz, fi, R(0.5) * R(hsz), (f & 1) != 0, n, zo_s, xo_s,
                                     yo_s, zn_s, xn_s, yn_s, R(zco), R(zc2), inner, acc);
#pragma unroll
          for (int cc = 0; cc < NC; ++cc)
#pragma unroll
            for (int k = 0; k < N1; ++k)
              {
                R t = 0;
#pragma unroll
                for (int j = 0; j < N1; ++j)
                  t += T.Psub[1][j][k] * acc[cc][j];
                res[cc][k] = t;
              }
        }
      // each cell takes the results of its own items computed in this pass, face by face
#pragma unroll
      for (int f = 0; f < 4; ++f)
        if (coarse[f])
          {
            const bool mine = (coarse[f] >> lane) & 1;
            const int helper =
              first[f] + __popcll(coarse[f] & ((1ull << lane) - 1)) - base;
            const bool here = mine && helper >= 0 && helper < n_act;
            const int src = here ? helper : lane;
#pragma unroll
            for (int cc = 0; cc < NC; ++cc)
#pragma unroll
              for (int k = 0; k < N1; ++k)
                {
                  const R v = __shfl(res[cc][k], src);
                  if (here)
                    Racc[cc][face_dof<N1>(f, k)] += v;
                }
          }
    }
}

// in place: U <- M^-1 U for one component on an affine cell of area 1/inv_area
template <int P, typename R>
__device__ __forceinline__ void tensor_inverse(const Basis<R> &T, R *U, R inv_area)
{
  constexpr int N1 = P + 1;
  R X[N1 * N1], Y[N1 * N1];
#pragma unroll
  for (int j = 0; j < N1; ++j)
#pragma unroll
    for (int a = 0; a < N1; ++a)
      {
        R s = 0;
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
        R s = 0;
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
        R s = 0;
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
        R s = 0;
#pragma unroll
        for (int b = 0; b < N1; ++b)
          s += T.Ti[j][b] * X[i + N1 * b];
        U[i + N1 * j] = s;
      }
}

// ---------------------------------------------------------------------------------------
// Face fluxes of the two equations, as the `inner` / `bdry` functions of face_terms
// ---------------------------------------------------------------------------------------
// mass: minus side -F(own, neighbour); plus side F(neighbour, own).  Selects instead of
// branches: the face side is a constant in the per-face loop and a lane value in the
// compacted second-subface pass
template <typename R>
__device__ __forceinline__ R mass_face(const Phys<R> &M, bool is_m, R zo, Vec2<R> qo, R zn,
                                       Vec2<R> qn, Vec2<R> n, R zbm, R zbp)
{
  const R f = M.mass_flux(is_m ? zo : zn, is_m ? zn : zo, is_m ? qo : qn, is_m ? qn : qo, n,
                          zbm, zbp);
  return is_m ? -f : f;
}

// momentum: minus side -(F(own, nbr) + P(own, nbr; zbm, zbp)); plus side F(nbr, own) +
// P(own, nbr; zbp, zbm)
template <typename R>
__device__ __forceinline__ Vec2<R> momentum_face(const Phys<R> &M, bool is_m, R zo, Vec2<R> qo,
                                                 R zn, Vec2<R> qn, Vec2<R> n, R zbm, R zbp)
{
  const Vec2<R> F = M.adv_flux(is_m ? zo : zn, is_m ? zn : zo, is_m ? qo : qn, is_m ? qn : qo,
                               n, zbm, zbp);
  const Vec2<R> Pr = M.pres_flux(zo, zn, n, is_m ? zbm : zbp, is_m ? zbp : zbm);
  return {is_m ? -F.x - Pr.x : F.x + Pr.x, is_m ? -F.y - Pr.y : F.y + Pr.y};
}

// boundary face: exterior state of the condition (OceanDG.h:1130-1203 / 1246-1317), mass
// and momentum flux; the momentum flux is cut where an outflow boundary sees backflow
// (OceanDG.h:1326-1332)
template <typename R>
__device__ __forceinline__ void boundary_faces(const Phys<R> &M, int bk, R zo, Vec2<R> qo,
                                               Vec2<R> n, R zb, const double *bv, R &mass,
                                               Vec2<R> &mom)
{
  R zp;
  Vec2<R> qp;
  const bool at_outflow = boundary_state(M, bk, zo, qo, n, zb, bv, zp, qp);
  mass = -M.mass_flux(zo, zp, qo, qp, n, zb, zb);
  Vec2<R> F = M.adv_flux(zo, zp, qo, qp, n, zb, zb);
  const Vec2<R> Pr = M.pres_flux(zo, zp, n, zb, zb);
  F.x += Pr.x;
  F.y += Pr.y;
  if (at_outflow && dot_n(qo, n) < R(-1e-12))
    F = {R(0), R(0)};
  mom = {-F.x, -F.y};
}

template <typename R>
struct MassFaces
{
  Phys<R> M;
  __device__ __forceinline__ void operator()(bool is_m, R zo, Vec2<R> qo, R zn, Vec2<R> qn,
                                             Vec2<R> n, R zbm, R zbp, R *val) const
  {
    val[0] = mass_face(M, is_m, zo, qo, zn, qn, n, zbm, zbp);
  }
  __device__ __forceinline__ void operator()(int bk, R zo, Vec2<R> qo, Vec2<R> n, R zb,
                                             const double *bv, R *val) const
  {
    R zp;
    Vec2<R> qp;
    boundary_state(M, bk, zo, qo, n, zb, bv, zp, qp);
    val[0] = -M.mass_flux(zo, zp, qo, qp, n, zb, zb);
  }
};

template <typename R>
struct MomentumFaces
{
  Phys<R> M;
  __device__ __forceinline__ void operator()(bool is_m, R zo, Vec2<R> qo, R zn, Vec2<R> qn,
                                             Vec2<R> n, R zbm, R zbp, R *val) const
  {
    const Vec2<R> f = momentum_face(M, is_m, zo, qo, zn, qn, n, zbm, zbp);
    val[0] = f.x;
    val[1] = f.y;
  }
  __device__ __forceinline__ void operator()(int bk, R zo, Vec2<R> qo, Vec2<R> n, R zb,
                                             const double *bv, R *val) const
  {
    R mass;
    Vec2<R> mom;
    boundary_faces(M, bk, zo, qo, n, zb, bv, mass, mom);
    val[0] = mom.x;
    val[1] = mom.y;
  }
};

// both equations at once (val[0] mass, val[1], val[2] momentum): the traces, velocities,
// wave speeds and reconstructed depths of a face point are formed once for the three
template <typename R>
struct AllFaces
{
  Phys<R> M;
  __device__ __forceinline__ void operator()(bool is_m, R zo, Vec2<R> qo, R zn, Vec2<R> qn,
                                             Vec2<R> n, R zbm, R zbp, R *val) const
  {
    val[0] = mass_face(M, is_m, zo, qo, zn, qn, n, zbm, zbp);
    const Vec2<R> f = momentum_face(M, is_m, zo, qo, zn, qn, n, zbm, zbp);
    val[1] = f.x;
    val[2] = f.y;
  }
  __device__ __forceinline__ void operator()(int bk, R zo, Vec2<R> qo, Vec2<R> n, R zb,
                                             const double *bv, R *val) const
  {
    R mass;
    Vec2<R> mom;
    boundary_faces(M, bk, zo, qo, n, zb, bv, mass, mom);
    val[0] = mass;
    val[1] = mom.x;
    val[2] = mom.y;
  }
};

// cell term of L_zeta: acc += integral of grad(phi) . q (mass_flux = q,
// ShallowWater.h:352-357)
template <int P, typename R, typename QA>
__device__ __forceinline__ void height_cell_term(const Basis<R> &T, R hx, R hy, const QA &QX,
                                                 const QA &QY, R (&acc)[(P + 1) * (P + 1)])
{
  constexpr int N1 = P + 1, NQ = nq_of(P);
  const R ihx = R(1) / hx, ihy = R(1) / hy, area = hx * hy;
#pragma unroll 1
  for (int b = 0; b < NQ; ++b)
    {
      R xs[N1], ys[N1], WS[N1], WD[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          R sx = 0, sy = 0;
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
          R qx = 0, qy = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              qx += T.Sq[a][i] * xs[i];
              qy += T.Sq[a][i] * ys[i];
            }
          const R JxW = T.wq[a] * T.wq[b] * area;
          const R gx = qx * ihx * JxW, gy = qy * ihy * JxW;
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
          acc[i + N1 * j] += T.Sq[b][j] * WS[i] + T.Dq[b][j] * WD[i];
    }
}

// ---------------------------------------------------------------------------------------
// L_zeta of one cell (cell term + faces) into Racc
// ---------------------------------------------------------------------------------------
template <int P, typename R, typename ZA, typename QA>
__device__ __forceinline__ void height_operator(const OpArgs &O, const Basis<R> &T,
                                                const Phys<R> &M, int batch, int lane, R hx,
                                                R hy, const ZA &Z, const QA &QX, const QA &QY,
                                                double zc, const double *__restrict__ src_h,
                                                const double *__restrict__ src_q,
                                                R (&Racc)[1][(P + 1) * (P + 1)])
{
  constexpr int ND = (P + 1) * (P + 1);
#pragma unroll
  for (int k = 0; k < ND; ++k)
    Racc[0][k] = 0;
  height_cell_term<P, R>(T, hx, hy, QX, QY, Racc[0]);
  phase_barrier();
  const MassFaces<R> fn{M};
  face_terms<P, 1, R>(O, T, batch, lane, hx, hy, Z, QX, QY, zc, src_h, src_q, Racc, fn, fn);
}

// own data of a cell in registers: Z as deviation from zc (zc = 0 for FP64)
template <int P, typename R>
__device__ __forceinline__ double load_own(const double *__restrict__ src_h,
                                           const double *__restrict__ src_q, int batch,
                                           int lane, R *Z, R *QX, R *QY)
{
  constexpr int ND = (P + 1) * (P + 1);
  double zg[ND], xg[ND], yg[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      zg[k] = src_h[vix(batch, 1, 0, ND, k, lane)];
      xg[k] = src_q[vix(batch, 2, 0, ND, k, lane)];
      yg[k] = src_q[vix(batch, 2, 1, ND, k, lane)];
    }
  issue_barrier();
  const double zc = centred<R>() ? zg[0] : 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      Z[k] = R(zg[k] - zc);
      QX[k] = R(xg[k]);
      QY[k] = R(yg[k]);
    }
  return zc;
}

template <typename R>
__device__ __forceinline__ Phys<R> physics(const OpArgs &O)
{
  return Phys<R>{R(O.g), R(O.cu4), R(O.nu)};
}

// L_zeta (INVERT: M^-1 L_zeta, for apply())
// KT: storage type of dst (a stage's k vector is stored in R, a caller vector in FP64)
template <int P, bool INVERT, typename R, typename KT = double>
__global__ __launch_bounds__(CB) void k_height_rhs(OpArgs O, const double *__restrict__ src_h,
                                                   const double *__restrict__ src_q,
                                                   double *__restrict__ dst)
{
  constexpr int ND = (P + 1) * (P + 1);
  const Basis<R> &T = basis<R>(P);
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x & (CB - 1);
  if (batch * CB + lane >= O.n_cells)
    return;
  const Phys<R> M = physics<R>(O);
  const R hx = R(O.hxy[((size_t)batch * 2) * CB + lane]);
  const R hy = R(O.hxy[((size_t)batch * 2 + 1) * CB + lane]);
  R Z[ND], QX[ND], QY[ND];
  const double zc = load_own<P, R>(src_h, src_q, batch, lane, Z, QX, QY);
  R Racc[1][ND];
  height_operator<P, R>(O, T, M, batch, lane, hx, hy, (const R *)Z, (const R *)QX,
                        (const R *)QY, zc, src_h, src_q, Racc);
  if (INVERT)
    tensor_inverse<P, R>(T, Racc[0], R(1) / (hx * hy));
#pragma unroll
  for (int k = 0; k < ND; ++k)
    reinterpret_cast<KT *>(dst)[vix(batch, 1, 0, ND, k, lane)] = KT(Racc[0][k]);
}

// ---------------------------------------------------------------------------------------
// L_q of one batch; MODE 0: stage (M^-1, RK update), 1: apply (M^-1 -> dst), 2: rhs,
// 3: stage of the split path (M^-1 -> k vector), 4: rhs into a k vector (the explicit
// momentum integral of the additive stage).  The batch's own zeta, q (and velocity) are
// staged in LDS, lane-interleaved.
//
// k vectors of a stage (ki_h[1..], ki_q[1..]) are stored in the integral precision R: the
// FP32 path's k values are FP32 numbers, so FP32 storage is exact and halves their
// traffic; every combination with them (the RK update, the front) is FP64.
// ---------------------------------------------------------------------------------------
template <int P, bool VISC>
constexpr int discharge_lds_words()
{
  return (VISC ? 5 : 3) * (P + 1) * (P + 1) * CB;
}

// ---------------------------------------------------------------------------------------
// Paired work items: the height wave (ROLE 0) and the discharge wave (ROLE 1) of one batch
// share a workgroup and split the faces instead of each integrating all four for its own
// equation: ROLE 0 integrates faces 0 and 1 (x = const), ROLE 1 faces 2 and 3 (y = const),
// each for the mass and both momentum components, so that every face point's traces,
// velocities, wave speeds and reconstructed depths are formed once per cell instead of
// twice.  pair_exchange then hands over what belongs to the other wave.  A[0] (mass) and
// A[1], A[2] (momentum) hold, on entry, this wave's cell term and its two faces'
// integrals; on return the wave's own equation holds all four faces.  A face integral
// touches only the cell's N1 DoFs on that face, so the swap moves 4 N1 momentum values
// (ROLE 0 -> 1) and 2 N1 mass values (ROLE 1 -> 0) per cell through `xch`
// (pair_xch_words() words of LDS).  Both waves call it once, every lane (one workgroup
// barrier).
// ---------------------------------------------------------------------------------------
template <int P>
__host__ __device__ constexpr int pair_xch_words()
{
  return 6 * (P + 1) * CB;
}

template <int P, typename R, int ROLE>
__device__ __forceinline__ void pair_exchange(R (&A)[3][(P + 1) * (P + 1)], R *xch)
{
  constexpr int N1 = P + 1;
  const int lane = threadIdx.x & (CB - 1);
  // slots: momentum component c of face f < 2 at ((c - 1) 2 + f) N1 + k; mass of face
  // f >= 2 at 4 N1 + (f - 2) N1 + k
  if constexpr (ROLE == 0)
    {
#pragma unroll
      for (int c = 1; c < 3; ++c)
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
          for (int k = 0; k < N1; ++k)
            xch[(((c - 1) * 2 + f) * N1 + k) * CB + lane] = A[c][face_dof<N1>(f, k)];
    }
  else
    {
#pragma unroll
      for (int f = 2; f < 4; ++f)
#pragma unroll
        for (int k = 0; k < N1; ++k)
          xch[(4 * N1 + (f - 2) * N1 + k) * CB + lane] = A[0][face_dof<N1>(f, k)];
    }
  __syncthreads();
  if constexpr (ROLE == 0)
    {
#pragma unroll
      for (int f = 2; f < 4; ++f)
#pragma unroll
        for (int k = 0; k < N1; ++k)
          A[0][face_dof<N1>(f, k)] += xch[(4 * N1 + (f - 2) * N1 + k) * CB + lane];
    }
  else
    {
#pragma unroll
      for (int c = 1; c < 3; ++c)
#pragma unroll
        for (int f = 0; f < 2; ++f)
#pragma unroll
          for (int k = 0; k < N1; ++k)
            A[c][face_dof<N1>(f, k)] += xch[(((c - 1) * 2 + f) * N1 + k) * CB + lane];
    }
}

// cell term of L_q (advective and diffusive flux, gravity, Manning friction, wind stress,
// Coriolis; OceanDG.h:696-743, ShallowWater.h:381-432): AX, AY += the integrals against
// grad(phi) and phi.  The batch's own zeta (deviation from zcr), q (and velocity) are read
// from the lane-interleaved LDS copy `lds`.
template <int P, bool VISC, typename R>
__device__ __forceinline__ void discharge_cell_term(const OpArgs &O, const Basis<R> &T,
                                                    const R *lds, int batch, int lane, R hx,
                                                    R hy, R zcr, R (&AX)[(P + 1) * (P + 1)],
                                                    R (&AY)[(P + 1) * (P + 1)])
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P);
  const LdsVec<R> Z{lds + lane}, QX{lds + ND * CB + lane}, QY{lds + 2 * ND * CB + lane};
  const LdsVec<R> UX{lds + 3 * ND * CB + lane}, UY{lds + 4 * ND * CB + lane};
  const Phys<R> M = physics<R>(O);
  const R g = R(O.g), nu = R(O.nu);
  const R ihx = R(1) / hx, ihy = R(1) / hy, area = hx * hy;
  const size_t c0base = (size_t)batch * NQ * NQ * CB + lane;
  const FieldV<R> bed = in_vgprs<R>(O.cell0[0]);
#pragma unroll 1
  for (int b = 0; b < NQ; ++b)
    {
      // the own nodal values are read from LDS again in every row: hoisted out of the loop
      // they would stay live in 3 nd registers (FP64: 6 nd) across the whole cell term,
      // which held the FP64 kernels at p = 3 to one wave per SIMD
      asm volatile("" ::: "memory");
      R zs[N1], zd[N1], xs[N1], ys[N1];
      R uxs[VISC ? N1 : 1], uxd[VISC ? N1 : 1], uys[VISC ? N1 : 1], uyd[VISC ? N1 : 1];
      R WSX[N1], WDX[N1], WSY[N1], WDY[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          R s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            {
              const R zv = Z[i + N1 * j];
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
              R t0 = 0, t1 = 0, t2 = 0, t3 = 0;
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
          R z = 0, zx = 0, zy = 0, qx = 0, qy = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            {
              z += T.Sq[a][i] * zs[i];
              zx += T.Dq[a][i] * zs[i];
              zy += T.Sq[a][i] * zd[i];
              qx += T.Sq[a][i] * xs[i];
              qy += T.Sq[a][i] * ys[i];
            }
          z += zcr;
          zx *= ihx;
          zy *= ihy;
          const int q = a + NQ * b;
          const size_t pi = c0base + (size_t)q * CB;
          const R zb = fld(bed, pi);
          const R h = M.depth(z, zb);
          const R s = M.inv_depth(h);
          const R vx = qx * s, vy = qy * s;
          R F00 = qx * vx, F01 = qx * vy, F10 = qy * vx, F11 = qy * vy;
          if (VISC)
            {
              R G00 = 0, G01 = 0, G10 = 0, G11 = 0;
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
              const R div = G00 + G11;
              F00 = F00 - nu * (G00 + G00) + nu * div;
              F01 = F01 - nu * (G01 + G10);
              F10 = F10 - nu * (G10 + G01);
              F11 = F11 - nu * (G11 + G11) + nu * div;
            }
          const R mgh = -g * h;
          R Sx = mgh * zx, Sy = mgh * zy;
          if (O.has_friction)
            {
              const R n_m = R(fld(O.cell0[1], pi));
              const R inv_d = R(1) / (exp(log(h) * R(0.33333333333)));
              const R hcf = g * n_m * n_m * inv_d * sqrt(vx * vx + vy * vy);
              Sx -= hcf * vx;
              Sy -= hcf * vy;
            }
          if (O.has_wind)
            {
              Sx += R(fld(O.cell0[2], pi));
              Sy += R(fld(O.cell0[3], pi));
            }
          if (O.has_coriolis)
            {
              const R fc = R(fld(O.cell0[4], pi));
              Sx += fc * qy;
              Sy += -fc * qx;
            }
          const R JxW = T.wq[a] * T.wq[b] * area;
          const R g0x = F00 * ihx * JxW, g0y = F01 * ihy * JxW;
          const R g1x = F10 * ihx * JxW, g1y = F11 * ihy * JxW;
          const R v0 = Sx * JxW, v1 = Sy * JxW;
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
            AX[i + N1 * j] += T.Sq[b][j] * WSX[i] + T.Dq[b][j] * WDX[i];
            AY[i + N1 * j] += T.Sq[b][j] * WSY[i] + T.Dq[b][j] * WDY[i];
          }
    }
}

// IN32 (FP32 paired stage): the stage input S.src_h, S.src_q comes in its centred FP32 form
// (c32_values, c32_centres); S.out32: the RK update writes the next stage input so
template <int P, bool VISC, int MODE, typename R, bool PAIR = false, bool IN32 = false>
__device__ __forceinline__ void discharge_block(const OpArgs &O, const StageArgs &S, int batch,
                                                R *lds, R *xch = nullptr)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NQ = nq_of(P);
  // dst of a stage is a k vector (stored in R); of apply / rhs a caller vector (FP64)
  using DT = typename std::conditional<MODE == 0 || MODE == 3 || MODE == 4, R, double>::type;
  DT *const dst = reinterpret_cast<DT *>(S.dst);
  const Basis<R> &T = basis<R>(P);
  const int lane = threadIdx.x & (CB - 1);
  const double *__restrict__ src_h = S.src_h;
  const double *__restrict__ src_q = S.src_q;
  const double *gh = src_h + (size_t)batch * ND * CB;
  const double *gq = src_q + (size_t)batch * 2 * ND * CB;
  // stage the batch's own data: contiguous, coalesced; each lane later reads only its own
  // column, so no barrier is needed
  double zc;
  if constexpr (PAIR && centred<R>() && IN32)
    {
      // the stage input in its centred FP32 form: the batch's zeta and q blocks are, word for
      // word, the LDS copy's layout, so they go there by direct-to-LDS loads (16 bytes per
      // lane, 1 KiB per instruction; a 256-byte tail by 4-byte loads), no registers held
      typedef __attribute__((address_space(1))) void gvoid;
      typedef __attribute__((address_space(3))) void lvoid;
      const char *gz = reinterpret_cast<const char *>(c32_values(src_h) + (size_t)batch * ND * CB);
      const char *gqq =
        reinterpret_cast<const char *>(c32_values(src_q) + (size_t)batch * 2 * ND * CB);
      char *lz = reinterpret_cast<char *>(lds);
      constexpr int ZB = ND * CB * 4;  // bytes of the zeta block; q's is twice that
#pragma unroll
      for (int o = 0; o + 1024 <= ZB; o += 1024)
        __builtin_amdgcn_global_load_lds((gvoid *)(gz + o + lane * 16), (lvoid *)(lz + o), 16, 0,
                                         0);
#pragma unroll
      for (int o = ZB / 1024 * 1024; o < ZB; o += 256)
        __builtin_amdgcn_global_load_lds((gvoid *)(gz + o + lane * 4), (lvoid *)(lz + o), 4, 0, 0);
#pragma unroll
      for (int o = 0; o + 1024 <= 2 * ZB; o += 1024)
        __builtin_amdgcn_global_load_lds((gvoid *)(gqq + o + lane * 16), (lvoid *)(lz + ZB + o),
                                         16, 0, 0);
#pragma unroll
      for (int o = 2 * ZB / 1024 * 1024; o < 2 * ZB; o += 256)
        __builtin_amdgcn_global_load_lds((gvoid *)(gqq + o + lane * 4), (lvoid *)(lz + ZB + o), 4,
                                         0, 0);
      zc = c32_centres(src_h, O.n_batches, ND)[(size_t)batch * CB + lane];
      // the copy is complete before the workgroup barrier that follows the staging
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
  else
    {
      double zg[ND], xg[ND], yg[ND];
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          zg[k] = gh[k * CB + lane];
          xg[k] = gq[k * CB + lane];
          yg[k] = gq[(ND + k) * CB + lane];
        }
      issue_barrier();
      zc = centred<R>() ? zg[0] : 0.0;
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          lds[k * CB + lane] = R(zg[k] - zc);
          lds[(ND + k) * CB + lane] = R(xg[k]);
          lds[(2 * ND + k) * CB + lane] = R(yg[k]);
        }
    }
  if (VISC)
    {
      const double *gv = S.src_v + (size_t)batch * 2 * ND * CB;
      double ug[ND], vg[ND];
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          ug[k] = gv[k * CB + lane];
          vg[k] = gv[(ND + k) * CB + lane];
        }
      issue_barrier();
#pragma unroll
      for (int k = 0; k < ND; ++k)
        {
          lds[(3 * ND + k) * CB + lane] = R(ug[k]);
          lds[(4 * ND + k) * CB + lane] = R(vg[k]);
        }
    }
  // cells outside the mesh (the tail of the last batch) take part in the wave-wide
  // matrix-core product of the inverse mass with zeros and store nothing
  const bool active = batch * CB + lane < O.n_cells;
  R Racc[2][ND];
#pragma unroll
  for (int k = 0; k < ND; ++k)
    {
      Racc[0][k] = 0;
      Racc[1][k] = 0;
    }
  R area = 1;
  if constexpr (PAIR)
    {
      // faces 2 and 3 for both equations (pair_exchange); the LDS copy of the batch's own
      // data staged above is the height wave's too (read after this barrier, until the
      // exchange)
      static_assert(!VISC && (MODE == 0 || MODE == 4),
                    "paired items: inviscid stage (explicit, or the additive stage's explicit "
                    "momentum integrals)");
      __syncthreads();
      R A[3][ND];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          A[c][k] = 0;
      if (active)
        {
          const LdsVec<R> Z{lds + lane}, QX{lds + ND * CB + lane},
            QY{lds + 2 * ND * CB + lane};
          const R hx = R(O.hxy[((size_t)batch * 2) * CB + lane]);
          const R hy = R(O.hxy[((size_t)batch * 2 + 1) * CB + lane]);
          area = hx * hy;
          discharge_cell_term<P, VISC, R>(O, T, lds, batch, lane, hx, hy, R(zc), A[1], A[2]);
          phase_barrier();
          const AllFaces<R> fn{physics<R>(O)};
          face_terms<P, 3, R, 0xC>(O, T, batch, lane, hx, hy, Z, QX, QY, zc, src_h, src_q, A,
                                   fn, fn, IN32);
        }
      pair_exchange<P, R, 1>(A, xch);
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          Racc[c][k] = A[1 + c][k];
    }
  else if (active)
    {
      const LdsVec<R> Z{lds + lane}, QX{lds + ND * CB + lane}, QY{lds + 2 * ND * CB + lane};
      const Phys<R> M = physics<R>(O);
      const R hx = R(O.hxy[((size_t)batch * 2) * CB + lane]);
      const R hy = R(O.hxy[((size_t)batch * 2 + 1) * CB + lane]);
      area = hx * hy;
      discharge_cell_term<P, VISC, R>(O, T, lds, batch, lane, hx, hy, R(zc), Racc[0], Racc[1]);
      phase_barrier();
      const MomentumFaces<R> fn{M};
      face_terms<P, 2, R>(O, T, batch, lane, hx, hy, Z, QX, QY, zc, src_h, src_q, Racc, fn, fn);
    }
  if (MODE == 2 || MODE == 4)
    {
      if (active)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
          for (int k = 0; k < ND; ++k)
            dst[vix(batch, 2, c, ND, k, lane)] = DT(Racc[c][k]);
      return;
    }
  if (MODE == 0 && S.scratch && active)
    {
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          S.scratch[vix(batch, 2, c, ND, k, lane)] = double(Racc[c][k]);
    }
  phase_barrier();
  const R inv_area = R(1) / area;
  if constexpr (std::is_same<R, float>::value && P >= 3)
    {
      // M^-1 = K / area on the matrix cores, every lane of the wave taking part (the
      // batch's own data in LDS is dead by now)
      static_assert(discharge_lds_words<P, VISC>() >= matvec_lds_words<ND, 2>(),
                    "LDS of the discharge item too small for the inverse mass");
      float y[2][ND];
      wave_matvec<ND, 2>(O.kinv, Racc, y, lds);
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          Racc[c][k] = y[c][k] * inv_area;
    }
  else
    {
      tensor_inverse<P, R>(T, Racc[0], inv_area);
      tensor_inverse<P, R>(T, Racc[1], inv_area);
    }
  if (!active)
    return;
  if (O.closure)
    {
      // learned momentum-source correction, already in M^-1 form (k_closure)
#pragma unroll
      for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int k = 0; k < ND; ++k)
          Racc[c][k] += R(O.closure[vix(batch, 2, c, ND, k, lane)]);
    }
#pragma unroll
  for (int c = 0; c < 2; ++c)
#pragma unroll
    for (int k = 0; k < ND; ++k)
      dst[vix(batch, 2, c, ND, k, lane)] = DT(Racc[c][k]);
  if (MODE == 0)
    {
      // post-op of OceanDG.h:1878-1909: dst = w + sum_j a_j k_{j+1}, k_{stage+1} = Racc,
      // summed in the reference's order; the loop over the earlier k vectors is outside
      // the loop over the DoFs so that the 2 nd loads of one vector are in flight together
      // FP32 integrals: one component per pass keeps the kernel within 3 waves per SIMD
      constexpr int NCH = std::is_same<R, float>::value ? 2 : 1, CH = 2 * ND / NCH;
#pragma unroll
      for (int ch = 0; ch < NCH; ++ch)
        {
          const size_t base = ((size_t)batch * 2 * ND + ch * CH) * CB + lane;
          double acc[CH];
#pragma unroll
          for (int i = 0; i < CH; ++i)
            acc[i] = S.wsol[base + (size_t)i * CB];
          issue_barrier();
          for (int j = 0; j < S.stage; ++j)
            {
              const R *kp = reinterpret_cast<const R *>(S.kprev[j]) + base;
              const double fj = S.factor[j];
              double t[CH];
#pragma unroll
              for (int i = 0; i < CH; ++i)
                t[i] = double(kp[(size_t)i * CB]);
              issue_barrier();
#pragma unroll
              for (int i = 0; i < CH; ++i)
                acc[i] += fj * t[i];
            }
          const double fs = S.factor[S.stage];
          if (centred<R>() && S.out32)
            {
              // the next stage input in its centred FP32 form (q: the values as they are)
              float *r32 = reinterpret_cast<float *>(S.rk_dst);
#pragma unroll
              for (int i = 0; i < CH; ++i)
                {
                  const int c = (ch * CH + i) / ND, k = (ch * CH + i) % ND;
                  r32[base + (size_t)i * CB] = float(acc[i] + fs * double(Racc[c][k]));
                }
            }
          else
#pragma unroll
            for (int i = 0; i < CH; ++i)
              {
                const int c = (ch * CH + i) / ND, k = (ch * CH + i) % ND;
                S.rk_dst[base + (size_t)i * CB] = acc[i] + fs * double(Racc[c][k]);
              }
        }
    }
}

template <int P, bool VISC, int MODE, typename R>
__global__ __launch_bounds__(CB) void k_discharge(OpArgs O, StageArgs S)
{
  __shared__ R lds[discharge_lds_words<P, VISC>()];
  discharge_block<P, VISC, MODE, R>(
    O, S, logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle), lds);
}

// ---------------------------------------------------------------------------------------
// front + capped lumped-Jacobi depth-mass solve (FP64, the reference arithmetic)
// ---------------------------------------------------------------------------------------
// R += sgn * integral(phi * depth(Z)) at GL(ng)^2; optionally counts dry points
template <int P, bool COUNT>
__device__ __forceinline__ int depth_mass(const OpArgs &O, const DegreeTables &T,
                                          const Phys<double> &M, int batch, int lane,
                                          const double *Z, double *R, double sgn, double area,
                                          unsigned long long &drybits, double &hmin)
{
  constexpr int N1 = P + 1, NG = P + 2;
  const size_t zbase = (size_t)batch * NG * NG * CB + lane;
  int ndry = 0;
  hmin = DBL_MAX;
  const FieldV<double> bed = in_vgprs<double>(O.cell1);
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
          const double zb = fld(bed, zbase + (size_t)q * CB);
          if (COUNT)
            {
              if (z <= -zb)
                {
                  ++ndry;
                  drybits |= 1ull << q;
                }
              hmin = fmin(hmin, z + zb);
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
    const FieldV<double> bed = in_vgprs<double>(O.cell1);
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
            const double zb = fld(bed, zbase + (size_t)(a + NG * b) * CB);
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
  // inverse lumped mass of a wet cell: 1/(l_i l_j) * (1/area) instead of 1/(l_i l_j area),
  // a rounding-level difference that saves one FP64 division per DoF and sweep.
  //
  // Sweeps in the linear regime.  While every GL point stays wetter than the depth floor
  // (z + zb > 1e-12: no dry point, floor not binding) the residual F - M h(Z) is affine in
  // Z with the exact tensor-product mass area (M1 x M1), so after an exact sweep the next
  // residuals follow from r <- r - area (M1 x M1) dZ (2 (p+1)^3 FMAs instead of two passes
  // over the (p+2)^2 points).  The regime is certified per cell: |z_k - z_b| <= lam C at
  // every GL point, C the l1 sum of the updates since the last exact sweep b, so
  // hmin_b - lam C > 2e-12 keeps every point strictly wet; otherwise the sweep is exact.
  const double inv_area = 1.0 / area;
  double norm = 1.0, hbase = -1.0, C = 0.0;
  bool linear = false;
  double Rr[ND];
  // Z += Dinv Rr and its l1 norm; the residual is advanced to the new Z by the linear
  // formula (used only if the next sweep is certified linear, recomputed otherwise)
  auto update = [&](auto dinv) {
    norm = 0.0;
    double dZ[ND];
#pragma unroll
    for (int i = 0; i < ND; ++i)
      {
        const double old = Z[i];
        Z[i] += dinv(i) * Rr[i];
        dZ[i] = Z[i] - old;
        norm += fabs(dZ[i]);
      }
    C += norm;
#pragma unroll
    for (int l = 0; l < N1; ++l)
      {
        double x[N1];
#pragma unroll
        for (int i = 0; i < N1; ++i)
          {
            double s = 0;
#pragma unroll
            for (int m = 0; m < N1; ++m)
              s += T.M1[i][m] * dZ[m + N1 * l];
            x[i] = s;
          }
#pragma unroll
        for (int i = 0; i < N1; ++i)
          dZ[i + N1 * l] = x[i];  // dZ now holds (M1 x I) dZ
      }
#pragma unroll
    for (int j = 0; j < N1; ++j)
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          double s = 0;
#pragma unroll
          for (int l = 0; l < N1; ++l)
            s += T.M1[j][l] * dZ[i + N1 * l];
          Rr[i + N1 * j] -= area * s;
        }
  };
  auto wet_dinv = [&](int i) { return T.ilumped2[i % N1][i / N1] * inv_area; };
  int k = 0;
  for (; k < O.max_it && norm > 1e-12; ++k)
    {
      if (linear && hbase - T.lam * C > 2e-12)
        {
          update(wet_dinv);
          continue;
        }
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Rr[i] = F[i];
      unsigned long long bits = 0;
      double hmin;
      const int ndry = depth_mass<P, true>(O, T, M, batch, lane, Z, Rr, -1.0, area, bits, hmin);
      linear = (ndry == 0);
      hbase = hmin;
      C = 0.0;
      if (ndry == 0)
        update(wet_dinv);
      else if (ndry == NP)
        update([](int) { return 0.0; });
      else
        {
          // lumped mass of the wet part (OceanDG.h:1470-1489); 1/D_i = inf where no wet
          // point supports phi_i, as in the reference
          double D[ND];
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
#pragma unroll
          for (int i = 0; i < ND; ++i)
            D[i] = 1. / D[i];
          update([&](int i) { return D[i]; });
        }
    }
  return k;
}

// Mixed precision depth solve.  When the cell is wet at every GL point with a margin the
// depth floor cannot bind and the capped Jacobi iteration is linear in the increment:
// delta^K = sum_{k<K} (I - D^-1 M)^k D^-1 r with r = sum_j a_j L_zeta,j, a fixed matrix
// (on the unit cell) times r / area.  Cells near the floor take the FP64 sweep.
// Returns the sweep count reported for the cell (max_it on the closed form).  DEFER: a
// cell that needs the FP64 sweep is left unchanged and flagged in the batch's mask
// O.defer[batch] for k_depth_deferred (keeps the FP64 sweep's registers out of the fused
// FP32 stage kernel).  Called by the active lanes; lane 0 of a batch is always active.
// The FP32 stage leaves the FP64 depth solve of cells near the floor to k_depth_deferred
// at p = 4, where its registers would spill heavily (256 VGPRs, 268 B per lane); at p <= 3
// the stage kernel does it itself.  At p = 3 that costs 47 spill slots but keeps 3 waves
// per SIMD, and saves the second launch: measured 2 % faster per stage on the 97k-cell
// hanging vortex.
__host__ __device__ constexpr bool defer_depth_solve(int p) { return p >= 4; }

// The FP32 stage forms J r on the matrix cores at p >= 3 (J the closed form of the capped
// sweeps, OpArgs::jmat): every lane of the wave takes part, in uniform control flow, a cell
// outside the mesh with r = 0, so that a cell's arithmetic does not depend on which other
// cells share its wave.  Returns false (nothing computed) below degree 3.
template <int P>
__device__ __forceinline__ bool depth_increment_mfma(const OpArgs &O, bool active,
                                                     const double *r, float *lds,
                                                     float (&jr)[(P + 1) * (P + 1)])
{
  constexpr int ND = (P + 1) * (P + 1);
  if constexpr (P >= 3)
    {
      float x[1][ND], y[1][ND];
#pragma unroll
      for (int i = 0; i < ND; ++i)
        x[0][i] = active ? float(r[i]) : 0.f;
      wave_matvec<ND, 1>(O.jmat, x, y, lds);
#pragma unroll
      for (int i = 0; i < ND; ++i)
        jr[i] = y[0][i];
      return true;
    }
  return false;
}

// jr: J r of this cell computed on the matrix cores by the whole wave (depth_increment_mfma),
// or null to form it here on the VALU
template <int P, bool DEFER>
__device__ __forceinline__ int mixed_depth_solve(const OpArgs &O, const DegreeTables &T,
                                                 const Phys<double> &M, int batch, int lane,
                                                 double area, const double *r, double *Z,
                                                 const float *jr)
{
  constexpr int N1 = P + 1, ND = N1 * N1, NG = P + 2;
  const Basis<float> &Tf = basis<float>(P);
  // minimum depth of zeta^n over the GL points (FP32 on the deviation from node 0)
  const double zc = Z[0];
  float dev[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
    dev[i] = float(Z[i] - zc);
  float hmin = FLT_MAX;
  const size_t zbase = (size_t)batch * NG * NG * CB + lane;
  const FieldV<double> bed = in_vgprs<double>(O.cell1);
#pragma unroll 1
  for (int b = 0; b < NG; ++b)
    {
      float zs[N1];
#pragma unroll
      for (int i = 0; i < N1; ++i)
        {
          float s = 0;
#pragma unroll
          for (int j = 0; j < N1; ++j)
            s += Tf.Sg[b][j] * dev[i + N1 * j];
          zs[i] = s;
        }
#pragma unroll
      for (int a = 0; a < NG; ++a)
        {
          float z = 0;
#pragma unroll
          for (int i = 0; i < N1; ++i)
            z += Tf.Sg[a][i] * zs[i];
          const double zb = fld(bed, zbase + (size_t)(a + NG * b) * CB);
          const float h = float(zc + zb) + z;
          hmin = h < hmin ? h : hmin;
        }
    }
  float rf[1][ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
    rf[0][i] = float(r[i]);
  const float ia = float(1.0 / area);
  float delta[ND];
  float dmax = 0.f;
  if (jr)
#pragma unroll
    for (int i = 0; i < ND; ++i)
      {
        delta[i] = jr[i] * ia;
        dmax = fmaxf(dmax, fabsf(delta[i]));
      }
  else
#pragma unroll
    for (int i = 0; i < ND; ++i)
      {
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < ND; ++j)
          s += O.jmat[i * ND + j] * rf[0][j];
        delta[i] = s * ia;
        dmax = fmaxf(dmax, fabsf(delta[i]));
      }
  const bool closed = hmin > float(O.wet_margin) * dmax && hmin > 1e-6f;
  if (DEFER)
    {
      const unsigned long long m = __ballot(!closed);
      if (lane == 0)
        O.defer[batch] = m;
    }
  if (closed)
    {
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Z[i] += double(delta[i]);
      return O.max_it;
    }
  if (DEFER)
    return 0;
  // FP64 reference path: front = r + M h(zeta^n), then the capped sweeps
  double F[ND];
#pragma unroll
  for (int i = 0; i < ND; ++i)
    F[i] = r[i];
  unsigned long long bits = 0;
  double hdummy;
  depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits, hdummy);
  return jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
}

__device__ __forceinline__ void store_sweeps(unsigned int *blk, int batch, int k)
{
  if (!blk)
    return;
  unsigned int v = (unsigned int)k;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_xor(v, off);
  if ((threadIdx.x & (CB - 1)) == 0)
    blk[batch] = v;
}

// front from ki_h[1..stage+1] (all in memory, stored in KT) and the Jacobi solve: the
// path used when the stage's outputs alias its inputs, and by the additive stage
template <int P, typename R, typename KT>
__global__ __launch_bounds__(CB) void k_height_jacobi(OpArgs O, StageArgs S)
{
  constexpr int ND = (P + 1) * (P + 1);
  __shared__ float lds[centred<R>() && P >= 3 ? matvec_lds_words<ND, 1>() : 1];
  const DegreeTables &T = basis<double>(P);
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  const int lane = threadIdx.x & (CB - 1);
  const bool active = batch * CB + lane < O.n_cells;
  const Phys<double> M = physics<double>(O);
  int k = 0;
  double F[ND], Z[ND], area = 1.0;
  if (active)
    {
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      area = hx * hy;
      const size_t base = (size_t)batch * ND * CB + lane;
      const KT *k0 = reinterpret_cast<const KT *>(S.kprev[0]) + base;
#pragma unroll
      for (int i = 0; i < ND; ++i)
        {
          F[i] = S.factor[0] * double(k0[(size_t)i * CB]);
          Z[i] = S.wsol[base + (size_t)i * CB];
        }
      for (int j = 1; j <= S.stage; ++j)
        {
          const KT *kp = reinterpret_cast<const KT *>(S.kprev[j]) + base;
          const double fj = S.factor[j];
#pragma unroll
          for (int i = 0; i < ND; ++i)
            F[i] += fj * double(kp[(size_t)i * CB]);
        }
    }
  float jr[ND];
  bool have_jr = false;
  if constexpr (centred<R>())
    have_jr = depth_increment_mfma<P>(O, active, F, lds, jr);
  if (active)
    {
      if (centred<R>())
        k = mixed_depth_solve<P, false>(O, T, M, batch, lane, area, F, Z,
                                        have_jr ? jr : nullptr);
      else
        {
          unsigned long long bits = 0;
          double hdummy;
          depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits, hdummy);
          if (S.scratch)
#pragma unroll
            for (int i = 0; i < ND; ++i)
              S.scratch[vix(batch, 1, 0, ND, i, lane)] = F[i];
          k = jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
        }
#pragma unroll
      for (int i = 0; i < ND; ++i)
        S.rk_dst[vix(batch, 1, 0, ND, i, lane)] = Z[i];
    }
  store_sweeps(S.sweeps_blk, batch, k);
}

// The FP64 depth solve of the cells the fused mixed-precision stage flagged (near the
// depth floor): front from ki_h[1..stage+1] (all written by then) and zeta^n = w_h (the
// stage leaves these cells' w_h untouched); adds their sweeps to the batch count.
template <int P>
__device__ __forceinline__ void depth_deferred_batch(const OpArgs &O, const StageArgs &S,
                                                     int batch, bool mine)
{
  constexpr int ND = (P + 1) * (P + 1);
  const DegreeTables &T = basis<double>(P);
  const int lane = threadIdx.x & (CB - 1);
  if (!__any(mine))
    return;
  int k = 0;
  if (mine)
    {
      const Phys<double> M = physics<double>(O);
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      const double area = hx * hy;
      double F[ND], Z[ND];
      const size_t base = (size_t)batch * ND * CB + lane;
      // k vectors of the FP32 stage are stored in FP32
      const float *k0 = reinterpret_cast<const float *>(S.kprev[0]) + base;
#pragma unroll
      for (int i = 0; i < ND; ++i)
        {
          F[i] = S.factor[0] * double(k0[(size_t)i * CB]);
          Z[i] = S.wsol[base + (size_t)i * CB];
        }
      for (int j = 1; j <= S.stage; ++j)
        {
          const float *kp = reinterpret_cast<const float *>(S.kprev[j]) + base;
          const double fj = S.factor[j];
#pragma unroll
          for (int i = 0; i < ND; ++i)
            F[i] += fj * double(kp[(size_t)i * CB]);
        }
      unsigned long long bits = 0;
      double hdummy;
      depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits, hdummy);
      k = jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
#pragma unroll
      for (int i = 0; i < ND; ++i)
        S.rk_dst[base + (size_t)i * CB] = Z[i];
    }
  if (S.sweeps_blk)
    {
      unsigned int v = (unsigned int)k;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
      if (lane == 0)
        S.sweeps_blk[batch] += v;
    }
}

// The FP64 depth solve of the cells the fused mixed-precision stage flagged (near the
// depth floor): front from ki_h[1..stage+1] (all written by then) and zeta^n = w_h (the
// stage leaves these cells' w_h untouched); adds their sweeps to the batch count.  One
// wave gathers the masks of 64 batches in one load and visits only the flagged ones, so
// a stage without flagged cells (the common case) costs a launch of n_batches / 64 waves.
constexpr int DEFER_BATCHES = 64;
template <int P>
__global__ __launch_bounds__(CB) void k_depth_deferred(OpArgs O, StageArgs S)
{
  const int lane = threadIdx.x & (CB - 1);
  const int b0 = blockIdx.x * DEFER_BATCHES;
  const unsigned long long m = b0 + lane < O.n_batches ? O.defer[b0 + lane] : 0ull;
  unsigned long long todo = __ballot(m != 0ull);
  while (todo)
    {
      const int j = __ffsll((long long)todo) - 1;
      todo &= todo - 1;
      const unsigned long long mj = __shfl(m, j);
      depth_deferred_batch<P>(O, S, b0 + j, ((mj >> lane) & 1ull) != 0);
    }
}

// L_zeta, front and depth solve of one batch of cells in one pass: L_zeta stays in
// registers for the front (it is also written out as ki_h[stage+1] for later stages).
// IN32 / S.out32 as for discharge_block (the zeta register's centres included)
template <int P, typename R, bool PAIR = false, bool IN32 = false>
__device__ __forceinline__ void height_stage_block(const OpArgs &O, const StageArgs &S,
                                                   int batch, float *lds, R *xch = nullptr,
                                                   const R *own = nullptr)
{
  constexpr int ND = (P + 1) * (P + 1);
  const DegreeTables &T = basis<double>(P);
  const int lane = threadIdx.x & (CB - 1);
  const bool active = batch * CB + lane < O.n_cells;
  const Phys<double> M = physics<double>(O);
  int k = 0;
  double F[ND], Z[ND], area = 1.0;
  R Racc[1][ND];
  if constexpr (PAIR)
    {
      // faces 0 and 1 for both equations; the momentum integrals go to the discharge wave,
      // the mass integrals of faces 2 and 3 come from it.  The batch's own data are read
      // from the discharge wave's LDS copy (`own`, staged before the first barrier)
      R A[3][ND];
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int i = 0; i < ND; ++i)
          A[c][i] = 0;
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      const double zc =
        !centred<R>() ? 0.0
        : IN32        ? c32_centres(S.src_h, O.n_batches, ND)[(size_t)batch * CB + lane]
                      : S.src_h[vix(batch, 1, 0, ND, 0, lane)];
      __syncthreads();
      if (active)
        {
          area = hx * hy;
          const LdsVec<R> Zo{own + lane}, QX{own + ND * CB + lane}, QY{own + 2 * ND * CB + lane};
          height_cell_term<P, R>(basis<R>(P), R(hx), R(hy), QX, QY, A[0]);
          phase_barrier();
          const AllFaces<R> fn{physics<R>(O)};
          face_terms<P, 3, R, 0x3>(O, basis<R>(P), batch, lane, R(hx), R(hy), Zo, QX, QY, zc,
                                   S.src_h, S.src_q, A, fn, fn, IN32);
        }
      pair_exchange<P, R, 0>(A, xch);
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Racc[0][i] = A[0][i];
    }
  else if (active)
    {
      const double hx = O.hxy[((size_t)batch * 2) * CB + lane];
      const double hy = O.hxy[((size_t)batch * 2 + 1) * CB + lane];
      area = hx * hy;
      R Zo[ND], QX[ND], QY[ND];
      const double zc = load_own<P, R>(S.src_h, S.src_q, batch, lane, Zo, QX, QY);
      height_operator<P, R>(O, basis<R>(P), physics<R>(O), batch, lane, R(hx), R(hy),
                            (const R *)Zo, (const R *)QX, (const R *)QY, zc, S.src_h,
                            S.src_q, Racc);
    }
  if (active)
    {
      {
        // front sum_j a_j k_{j+1} in the reference's order, the earlier k vectors read one
        // whole vector at a time (loads in flight together)
        const size_t base = (size_t)batch * ND * CB + lane;
#pragma unroll
        for (int i = 0; i < ND; ++i)
          reinterpret_cast<R *>(S.dst)[base + (size_t)i * CB] = Racc[0][i];  // ki_h[stage+1]
        if (S.stage == 0)
          {
#pragma unroll
            for (int i = 0; i < ND; ++i)
              F[i] = S.factor[0] * double(Racc[0][i]);
          }
        else
          {
            const R *k0 = reinterpret_cast<const R *>(S.kprev[0]) + base;
#pragma unroll
            for (int i = 0; i < ND; ++i)
              F[i] = double(k0[(size_t)i * CB]);
            issue_barrier();
#pragma unroll
            for (int i = 0; i < ND; ++i)
              F[i] = S.factor[0] * F[i];
            for (int j = 1; j < S.stage; ++j)
              {
                const R *kp = reinterpret_cast<const R *>(S.kprev[j]) + base;
                const double fj = S.factor[j];
                double t[ND];
#pragma unroll
                for (int i = 0; i < ND; ++i)
                  t[i] = double(kp[(size_t)i * CB]);
                issue_barrier();
#pragma unroll
                for (int i = 0; i < ND; ++i)
                  F[i] += fj * t[i];
              }
            const double fs = S.factor[S.stage];
#pragma unroll
            for (int i = 0; i < ND; ++i)
              F[i] += fs * double(Racc[0][i]);
          }
      }
      phase_barrier();
#pragma unroll
      for (int i = 0; i < ND; ++i)
        Z[i] = S.wsol[vix(batch, 1, 0, ND, i, lane)];
      issue_barrier();
    }
  float jr[ND];
  bool have_jr = false;
  if constexpr (centred<R>())
    have_jr = depth_increment_mfma<P>(O, active, F, lds, jr);
  if (active)
    {
      if (centred<R>())
        k = mixed_depth_solve<P, defer_depth_solve(P)>(O, T, M, batch, lane, area, F, Z,
                                                       have_jr ? jr : nullptr);
      else
        {
          unsigned long long bits = 0;
          double hdummy;
          depth_mass<P, false>(O, T, M, batch, lane, Z, F, 1.0, area, bits, hdummy);
          if (S.scratch)
#pragma unroll
            for (int i = 0; i < ND; ++i)
              S.scratch[vix(batch, 1, 0, ND, i, lane)] = F[i];
          k = jacobi_solve<P>(O, T, M, batch, lane, area, F, Z);
        }
      if (centred<R>() && S.out32)
        {
          // the next stage input in its centred FP32 form: zeta minus its DoF-0 value, and
          // that value among the centres
          float *r32 = reinterpret_cast<float *>(S.rk_dst);
#pragma unroll
          for (int i = 0; i < ND; ++i)
            r32[vix(batch, 1, 0, ND, i, lane)] = float(Z[i] - Z[0]);
          S.rk_dst[(size_t)O.n_batches * ND * CB / 2 + (size_t)batch * CB + lane] = Z[0];
        }
      else
#pragma unroll
        for (int i = 0; i < ND; ++i)
          S.rk_dst[vix(batch, 1, 0, ND, i, lane)] = Z[i];
    }
  store_sweeps(S.sweeps_blk, batch, k);
}

// Waves per SIMD the register allocator must leave room for (a SIMD holds 512 registers
// per lane for all its waves): the occupancy each inviscid stage kernel reaches, pinned so
// that a few more live values do not silently drop it by a wave.
template <int P, bool VISC, typename R>
constexpr int stage_min_waves()
{
  return VISC ? 1 : std::is_same<R, float>::value ? (P == 1 ? 5 : P == 2 ? 4 : P == 3 ? 3 : 2)
                                                  : (P == 1 ? 3 : P == 2 ? 2 : P == 3 ? 2 : 1);
}

// Kind of work item w (0 = height stage, 1 = discharge stage) of batch w / 2: the parity of
// the number of set bits of w (the Thue-Morse sequence).  Within an XCD the hardware deals
// consecutive workgroups round-robin over shader engines, then CUs, then SIMDs, all with
// power-of-two periods; with kind = w & 1 every CU received one kind only (measured: 128
// CUs ran only height items and idled while the other 128 ran the heavier discharge
// items).  Bit parity is balanced along every power-of-two stride, and the two items of a
// batch (w = 2b, 2b + 1) still have opposite kinds and sit next to each other.
__device__ __forceinline__ int item_kind(int w) { return __popc((unsigned)w) & 1; }

// the whole stage in one launch: work items w = (batch, kind), WPB waves (items) per
// workgroup, each wave on its own LDS slice; workgroups are dealt so that each XCD holds a
// contiguous range of batches
// LDS words (of R) a height item of the unpaired stage uses: the FP32 depth increment's
// matrix-core operands at p >= 3, rounded up to 64 words; nothing otherwise
template <int P, typename R>
__host__ __device__ constexpr int height_lds_words()
{
  return std::is_same<R, float>::value && P >= 3
           ? (matvec_lds_words<(P + 1) * (P + 1), 1>() + 63) / 64 * 64
           : 0;
}

// With WPB >= 2 the items of a workgroup are the two items of WPB / 2 batches (w = 2b and
// 2b + 1 always have opposite kinds), so the workgroup's LDS holds, per batch, a height
// slice followed by a discharge slice, each of its own size (mixed p = 4, WPB = 2: 29 KiB per
// workgroup, five per CU, against 38 KiB and four with two discharge-sized slices).  With
// WPB = 1 each workgroup holds one discharge-sized slice, whatever kind its item is.
template <int P, bool VISC, typename R, int WPB>
__global__ __launch_bounds__(CB * WPB, (stage_min_waves<P, VISC, R>())) void k_stage(
  OpArgs O, StageArgs H, StageArgs Q)
{
  constexpr int HW = height_lds_words<P, R>(), DW = discharge_lds_words<P, VISC>();
  __shared__ __align__(16) R lds[WPB == 1 ? DW : (WPB / 2) * (HW + DW)];
  static_assert(P < 3 || sizeof(R) * DW >= sizeof(float) * matvec_lds_words<(P + 1) * (P + 1), 1>(),
                "LDS of the stage too small for the matrix-core depth increment");
  static_assert(WPB == 1 || WPB % 2 == 0, "work items per workgroup: 1 or an even number");
  const int n_items = 2 * O.n_batches;
  const int wave = WPB == 1 ? 0 : wave_in_block();
  const int w = logical_batch(blockIdx.x, (n_items + WPB - 1) / WPB, O.xcd_swizzle) * WPB + wave;
  if (w >= n_items)
    return;
  const int kind = item_kind(w);
  R *slice = WPB == 1 ? lds : lds + (wave / 2) * (HW + DW) + (kind ? HW : 0);
  if (kind)
    discharge_block<P, VISC, 0, R>(O, Q, w >> 1, slice);
  else
    height_stage_block<P, R>(O, H, w >> 1, reinterpret_cast<float *>(slice));
}

// LDS offset (words of R) of the paired items' exchange area inside the height wave's slice:
// past the words the FP32 depth increment's matrix-core product uses
template <int P, typename R>
__host__ __device__ constexpr int pair_xch_offset()
{
  return std::is_same<R, float>::value && P >= 3 ? matvec_lds_words<(P + 1) * (P + 1), 1>() : 0;
}

// LDS of a paired workgroup (words of R): the height wave's slice (the FP32 depth
// increment's matrix-core product, then the exchange area), rounded up to 64 words,
// followed by the discharge wave's slice (its own data, later its matrix-core product).
// The height slice is only as large as its contents: at FP64 p = 3 a workgroup takes
// 36 KiB, so four workgroups (two waves per SIMD) fit a CU's 160 KiB of LDS, against
// three with two equal 24 KiB slices.
template <int P, typename R>
__host__ __device__ constexpr int pair_height_words()
{
  return (pair_xch_offset<P, R>() + pair_xch_words<P>() + 63) / 64 * 64;
}

template <int P, typename R>
__host__ __device__ constexpr int pair_lds_words()
{
  return pair_height_words<P, R>() + discharge_lds_words<P, false>();
}

// Which wave of a paired workgroup runs the height item: wave item_kind(blockIdx.x), the
// other the discharge item.  With the height item always on wave 0, the hardware's placement
// of a workgroup's waves on a CU's SIMDs could put every height wave on some SIMDs and every
// discharge wave (the longer item) on the others; alternating the roles along the Thue-Morse
// sequence of the workgroup index mixes the two kinds on every SIMD.
__device__ __forceinline__ bool pair_height_wave()
{
  return wave_in_block() == item_kind((int)blockIdx.x);
}

// the whole stage in one launch with paired work items: workgroup = one batch, one wave its
// height item and the other its discharge item (pair_height_wave), the faces split between
// them (pair_exchange).
// Two workgroup barriers: after the discharge wave has staged the batch's own data in its
// LDS slice (which the height wave then reads too), and in pair_exchange.  Workgroups are
// dealt so that each XCD holds a contiguous range of batches.
// IN32: the stage input comes in its centred FP32 form (c32_values, c32_centres; the device
// time loop's intermediate stages); a compile-time variant, so that neither form's kernel
// carries the other's loads
template <int P, typename R, bool IN32 = false>
__global__ __launch_bounds__(2 * CB, (stage_min_waves<P, false, R>())) void k_stage_pair(
  OpArgs O, StageArgs H, StageArgs Q)
{
  __shared__ __align__(16) R lds[pair_lds_words<P, R>()];
  const int batch = logical_batch(blockIdx.x, O.n_batches, O.xcd_swizzle);
  R *xch = lds + pair_xch_offset<P, R>();
  R *dlds = lds + pair_height_words<P, R>();
  if (pair_height_wave())
    height_stage_block<P, R, true, IN32>(O, H, batch, reinterpret_cast<float *>(lds), xch,
                                         dlds);
  else
    discharge_block<P, false, 0, R, true, IN32>(O, Q, batch, dlds, xch);
}

}  // namespace oceano
