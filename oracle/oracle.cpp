// oracle.cpp -- CPU restatement of oceano's per-stage DG right-hand side.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (oceano_amd/) links, loads or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg do,
// and only as the checker.  It restates, in plain sequential FP64 C++, the algorithm of
//   lib/include/space_discretization/OceanDG.h   (OceanoOperator, lines cited per function)
//   lib/include/model/shallow_water/ShallowWater.h
//   lib/include/numerical_flux/LaxFriedrichsModified.h, NumericalFluxBase.h
//   lib/include/deal.II_oceano/operators.h       (CellwiseInverseMassMatrixLumped)
//   lib/include/model/shallow_water/physics/{BottomFriction,WindStress,Coriolis,ViscosityCoefficient}.h
// and the deal.II 9.5 pieces those files call (MatrixFree::loop / cell_loop, FEEvaluation,
// FEFaceEvaluation, CellwiseInverseMassMatrix, QGauss, QGaussLobatto, FE_DGQ on
// Gauss-Lobatto nodes) from their documented algorithms.  deal.II is not part of the
// reference tree and is not installed here, so no line of it is reproduced.
//
// Parity pins (tests/test_oracle_*.py): quadrature vs numpy.leggauss and exactness,
// Lagrange identities, sum factorisation vs dense brute force, inverse mass vs dense LU,
// Jacobi invariants (converged input is a fixed point, spectral radius, dry lift), flux
// closed forms and (anti)symmetry, free-stream / lake-at-rest / mass-conservation on
// conforming and hanging meshes, vortex convergence order.
//
// Geometry: 2-D axis-aligned rectangular cells (every configured mesh is Cartesian);
// x = x0 + hx*xi, y = y0 + hy*eta on the unit cell, as deal.II's MappingQ(1) gives there.
// DoF layout: per cell, component-major, lexicographic Gauss-Lobatto nodes (x fastest),
// i.e. FESystem(FE_DGQ(p), c) numbering with cells in the caller's order.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

// --------------------------------------------------------------------------------------
// 1-D quadrature rules on [0,1], ascending (deal.II QGauss<1>(n), QGaussLobatto<1>(n)).
// --------------------------------------------------------------------------------------
void legendre(int n, double x, double &p, double &dp, double &pm1)
{
  // P_n(x), P_n'(x) and P_{n-1}(x) on [-1,1] by the three-term recurrence.
  double p0 = 1.0, p1 = x;
  if (n == 0)
    {
      p = 1.0;
      dp = 0.0;
      pm1 = 0.0;
      return;
    }
  for (int k = 2; k <= n; ++k)
    {
      const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = pk;
    }
  p = p1;
  pm1 = p0;
  // derivative from n (x P_n - P_{n-1}) / (x^2 - 1), valid away from +-1
  dp = n * (x * p1 - p0) / (x * x - 1.0);
}

void gauss_rule(int n, std::vector<double> &x01, std::vector<double> &w01)
{
  x01.assign(n, 0.0);
  w01.assign(n, 0.0);
  for (int i = 0; i < n; ++i)
    {
      double x = std::cos(M_PI * (i + 0.75) / (n + 0.5));
      for (int it = 0; it < 100; ++it)
        {
          double p, dp, pm1;
          legendre(n, x, p, dp, pm1);
          const double dx = p / dp;
          x -= dx;
          if (std::fabs(dx) < 1e-17)
            break;
        }
      double p, dp, pm1;
      legendre(n, x, p, dp, pm1);
      const double w = 2.0 / ((1.0 - x * x) * dp * dp);
      // x runs from +1 downwards; map to ascending [0,1]
      x01[i] = 0.5 * (1.0 - x);
      w01[i] = 0.5 * w;
    }
}

void gauss_lobatto_rule(int n, std::vector<double> &x01, std::vector<double> &w01)
{
  // end points plus the roots of P'_{n-1}; weights 2 / (n (n-1) P_{n-1}(x)^2)
  x01.assign(n, 0.0);
  w01.assign(n, 0.0);
  const int m = n - 1;
  std::vector<double> xs(n);
  xs[0] = -1.0;
  xs[n - 1] = 1.0;
  for (int i = 1; i < n - 1; ++i)
    {
      double x = -std::cos(M_PI * i / m);
      for (int it = 0; it < 100; ++it)
        {
          double p, dp, pm1;
          legendre(m, x, p, dp, pm1);
          // Legendre ODE: (1-x^2) P'' = 2x P' - m(m+1) P
          const double d2p = (2.0 * x * dp - m * (m + 1.0) * p) / (1.0 - x * x);
          const double dx = dp / d2p;
          x -= dx;
          if (std::fabs(dx) < 1e-17)
            break;
        }
      xs[i] = x;
    }
  for (int i = 0; i < n; ++i)
    {
      double p, dp, pm1;
      if (i == 0 || i == n - 1)
        {
          // P_m(+-1) = (+-1)^m
          p = (i == 0 && (m % 2)) ? -1.0 : 1.0;
        }
      else
        legendre(m, xs[i], p, dp, pm1);
      const double w = 2.0 / (n * (n - 1.0) * p * p);
      x01[i] = 0.5 * (1.0 + xs[i]);
      w01[i] = 0.5 * w;
    }
}

// Lagrange polynomial k on nodes z, and its derivative, at x.
double lagrange(const std::vector<double> &z, int k, double x)
{
  double v = 1.0;
  for (size_t m = 0; m < z.size(); ++m)
    if ((int)m != k)
      v *= (x - z[m]) / (z[k] - z[m]);
  return v;
}

double lagrange_deriv(const std::vector<double> &z, int k, double x)
{
  double s = 0.0;
  for (size_t m = 0; m < z.size(); ++m)
    {
      if ((int)m == k)
        continue;
      double v = 1.0 / (z[k] - z[m]);
      for (size_t n = 0; n < z.size(); ++n)
        if ((int)n != k && n != m)
          v *= (x - z[n]) / (z[k] - z[n]);
      s += v;
    }
  return s;
}

struct Mat
{
  int r = 0, c = 0;
  std::vector<double> a;
  void resize(int rr, int cc)
  {
    r = rr;
    c = cc;
    a.assign((size_t)rr * cc, 0.0);
  }
  double &operator()(int i, int j) { return a[(size_t)i * c + j]; }
  double operator()(int i, int j) const { return a[(size_t)i * c + j]; }
};

// dense inverse by Gauss-Jordan with partial pivoting (only for the tiny n1 x n1 matrix)
Mat invert(const Mat &m)
{
  const int n = m.r;
  Mat a = m, inv;
  inv.resize(n, n);
  for (int i = 0; i < n; ++i)
    inv(i, i) = 1.0;
  for (int col = 0; col < n; ++col)
    {
      int piv = col;
      for (int i = col + 1; i < n; ++i)
        if (std::fabs(a(i, col)) > std::fabs(a(piv, col)))
          piv = i;
      for (int j = 0; j < n; ++j)
        {
          std::swap(a(col, j), a(piv, j));
          std::swap(inv(col, j), inv(piv, j));
        }
      const double d = 1.0 / a(col, col);
      for (int j = 0; j < n; ++j)
        {
          a(col, j) *= d;
          inv(col, j) *= d;
        }
      for (int i = 0; i < n; ++i)
        if (i != col)
          {
            const double f = a(i, col);
            for (int j = 0; j < n; ++j)
              {
                a(i, j) -= f * a(col, j);
                inv(i, j) -= f * inv(col, j);
              }
          }
    }
  return inv;
}

// boundary condition kinds (IcbcBase.h:65-77)
enum BcKind
{
  BC_NONE = 0,
  BC_WALL = 1,
  BC_SUPERCRITICAL_INFLOW = 2,
  BC_SUPERCRITICAL_OUTFLOW = 3,
  BC_HEIGHT_INFLOW = 4,
  BC_DISCHARGE_INFLOW = 5,
  BC_ABSORBING_OUTFLOW = 6
};

struct InnerFace
{
  int cm, fm, sm, cp, fp, sp;  // cell, local face, subface (-1 = whole face) per side
};
struct BdryFace
{
  int c, f, bid;
};

struct V2
{
  double x, y;
};

// ---------------------------------------------------------------------------------------
// Pointwise model: ShallowWater.h:315-516 and the selected physics variants.
// ---------------------------------------------------------------------------------------
struct Model
{
  double g = 9.81;
  double cu4 = 1e-4;
  double nu = 0.0;

  // ShallowWater.h:315-320
  double depth(double z, double zb) const { return std::max(z + zb, 1e-12); }

  // ShallowWater.h:323-336
  V2 velocity(double z, V2 q, double zb) const
  {
    const double h = depth(z, zb);
    const double h4 = h * h * h * h;
    const double inverse_depth = std::sqrt(2.) * h / std::sqrt(h4 + std::max(h4, cu4));
    return {q.x * inverse_depth, q.y * inverse_depth};
  }

  double square_wavespeed(double z, double zb) const { return g * depth(z, zb); }

  // ShallowWater.h:381-403 with ViscosityCoefficientConstant (ViscosityCoefficient.h:100-105):
  // flux[e][d] = q_e v_d - nu (G[e][d] + G[d][e]); flux[e][e] += nu div G
  void advective_diffusive_flux(double z, V2 q, const double G[2][2], double zb,
                                double flux[2][2]) const
  {
    const V2 v = velocity(z, q, zb);
    const double vv[2] = {v.x, v.y};
    const double qq[2] = {q.x, q.y};
    const double div = G[0][0] + G[1][1];
    for (int d = 0; d < 2; ++d)
      for (int e = 0; e < 2; ++e)
        flux[e][d] = qq[e] * vv[d] - nu * (G[e][d] + G[d][e]);
    for (int e = 0; e < 2; ++e)
      flux[e][e] += nu * div;
  }

  // ShallowWater.h:408-432 with BottomFrictionManning (BottomFriction.h:186-199),
  // WindStressGeneral (WindStress.h:101-108), CoriolisBeta (Coriolis.h:87-95).
  // parameters = {z_b, manning n, tau_x, tau_y, f}
  V2 source(double z, V2 q, V2 dz, const double *par) const
  {
    const V2 v = velocity(z, q, par[0]);
    const double h = depth(z, par[0]);
    // Manning: inv_depth = 1/exp(log(h)*0.33333333333)
    const double inv_depth = 1. / (std::exp(std::log(h) * 0.33333333333));
    const double vnorm = std::sqrt(v.x * v.x + v.y * v.y);
    const double hcf = g * par[1] * par[1] * inv_depth * vnorm;
    const V2 fric = {hcf * v.x, hcf * v.y};
    const V2 wind = {par[2], par[3]};
    const V2 cor = {par[4] * q.y, -par[4] * q.x};
    const double mgh = -g * h;
    V2 s;
    s.x = mgh * dz.x - fric.x + wind.x + cor.x;
    s.y = mgh * dz.y - fric.y + wind.y + cor.y;
    return s;
  }

  // ShallowWater.h:437-458: the non-stiff part of the source (pressure gradient, wind,
  // Coriolis); par = {z_b, manning n, tau_x, tau_y, f}
  V2 source_nonstiff(double z, V2 q, V2 dz, const double *par) const
  {
    const double h = depth(z, par[0]);
    const double mgh = -g * h;
    V2 s;
    s.x = mgh * dz.x + par[2] + par[4] * q.y;
    s.y = mgh * dz.y + par[3] + -par[4] * q.x;
    return s;
  }

  // ShallowWater.h:461-472 with BottomFrictionManning::source (BottomFriction.h:186-199)
  V2 source_stiff(double z, V2 q, double zb, double manning) const
  {
    const V2 v = velocity(z, q, zb);
    const double h = depth(z, zb);
    const double inv_depth = 1. / (std::exp(std::log(h) * 0.33333333333));
    const double vnorm = std::sqrt(v.x * v.x + v.y * v.y);
    const double hcf = g * manning * manning * inv_depth * vnorm;
    return {-(hcf * v.x), -(hcf * v.y)};
  }

  // BottomFrictionManning::jacobian (BottomFriction.h:201-212)
  double friction_jacobian(V2 v, double manning, double h) const
  {
    const double inv_depth = 1. / (std::exp(std::log(h) * 1.33333333333));
    const double vnorm = std::sqrt(v.x * v.x + v.y * v.y);
    return g * manning * manning * inv_depth * vnorm;
  }

  double riemann_invariant_p(double z, V2 q, V2 n, double zb) const
  {
    const V2 v = velocity(z, q, zb);
    const double c = std::sqrt(square_wavespeed(z, zb));
    const double u = v.x * n.x + v.y * n.y;
    return u + 2. * c;
  }
  double riemann_invariant_m(double z, V2 q, V2 n, double zb) const
  {
    const V2 v = velocity(z, q, zb);
    const double c = std::sqrt(square_wavespeed(z, zb));
    const double u = v.x * n.x + v.y * n.y;
    return u - 2. * c;
  }

  // LaxFriedrichsModified.h:161-192
  double numerical_massflux_weak(double z_m, double z_p, V2 q_m, V2 q_p, V2 n, double zb_m,
                                 double zb_p) const
  {
    const V2 v_m = velocity(z_m, q_m, zb_m);
    const V2 v_p = velocity(z_p, q_p, zb_p);
    const double lambda_m =
      std::fabs(v_m.x * n.x + v_m.y * n.y) + std::sqrt(square_wavespeed(z_m, zb_m));
    const double lambda_p =
      std::fabs(v_p.x * n.x + v_p.y * n.y) + std::sqrt(square_wavespeed(z_p, zb_p));
    const double lambda = std::max(lambda_p, lambda_m);
    const double zb_star = std::min(zb_m, zb_p);
    const double hm = depth(z_m, zb_star), hp = depth(z_p, zb_star);
    const V2 flux_m = {hm * v_m.x, hm * v_m.y};
    const V2 flux_p = {hp * v_p.x, hp * v_p.y};
    const double zcorr_m = std::max(z_m, -zb_star);
    const double zcorr_p = std::max(z_p, -zb_star);
    return 0.5 * ((flux_m.x * n.x + flux_m.y * n.y) + (flux_p.x * n.x + flux_p.y * n.y)) +
           0.5 * lambda * (zcorr_m - zcorr_p);
  }

  // LaxFriedrichsModified.h:200-230
  V2 numerical_advflux_weak(double z_m, double z_p, V2 q_m, V2 q_p, V2 n, double zb_m,
                            double zb_p) const
  {
    const V2 v_m = velocity(z_m, q_m, zb_m);
    const V2 v_p = velocity(z_p, q_p, zb_p);
    const double lambda_m =
      std::fabs(v_m.x * n.x + v_m.y * n.y) + std::sqrt(square_wavespeed(z_m, zb_m));
    const double lambda_p =
      std::fabs(v_p.x * n.x + v_p.y * n.y) + std::sqrt(square_wavespeed(z_p, zb_p));
    const double lambda = std::max(lambda_p, lambda_m);
    const double zb_star = std::min(zb_m, zb_p);
    const double hm = depth(z_m, zb_star), hp = depth(z_p, zb_star);
    const V2 hu_m = {hm * v_m.x, hm * v_m.y};
    const V2 hu_p = {hp * v_p.x, hp * v_p.y};
    // advective_flux(z, hu, zb_star): flux[e][d] = hu_e * velocity(z, hu, zb_star)_d
    const V2 w_m = velocity(z_m, hu_m, zb_star);
    const V2 w_p = velocity(z_p, hu_p, zb_star);
    const double fm[2][2] = {{hu_m.x * w_m.x, hu_m.x * w_m.y}, {hu_m.y * w_m.x, hu_m.y * w_m.y}};
    const double fp[2][2] = {{hu_p.x * w_p.x, hu_p.x * w_p.y}, {hu_p.y * w_p.x, hu_p.y * w_p.y}};
    const double hl = 0.5 * lambda;
    V2 r;
    r.x = 0.5 * ((fm[0][0] * n.x + fm[0][1] * n.y) + (fp[0][0] * n.x + fp[0][1] * n.y)) +
          hl * (hu_m.x - hu_p.x);
    r.y = 0.5 * ((fm[1][0] * n.x + fm[1][1] * n.y) + (fp[1][0] * n.x + fp[1][1] * n.y)) +
          hl * (hu_m.y - hu_p.y);
    return r;
  }

  // NumericalFluxBase.h:103-126, CENTER variant
  V2 numerical_presflux_strong(double z_m, double z_p, V2 n, double zb_m, double zb_p) const
  {
    const double zb_star = std::min(zb_m, zb_p);
    const double h_p = depth(z_p, zb_star);
    const double zcorr_m = std::max(z_m, -zb_star);
    const double zcorr_p = std::max(z_p, -zb_star);
    const double h_m = depth(z_m, zb_star);
    const double s = g * 0.25 * (h_p + h_m) * (zcorr_p - zcorr_m);
    return {s * n.x, s * n.y};
  }
};

V2 outward_normal(int f)
{
  switch (f)
    {
      case 0:
        return {-1.0, 0.0};
      case 1:
        return {1.0, 0.0};
      case 2:
        return {0.0, -1.0};
      default:
        return {0.0, 1.0};
    }
}

// ---------------------------------------------------------------------------------------
// The operator
// ---------------------------------------------------------------------------------------
struct Oracle
{
  int p, n1, nq, ng, nd;  // degree, nodes/dir, Gauss cell pts/dir, GL pts/dir, dofs/cell/comp
  std::vector<double> nodes, gq_x, gq_w, gl_x, gl_w, gn_x, gn_w;
  Mat Sq, Dq, Sg, Sgsub[2], Sn, Sninv;
  Model model;
  int N = 0;
  std::vector<double> geom;  // [N][4] x0 y0 hx hy
  std::vector<InnerFace> inner;
  std::vector<BdryFace> bdry;
  std::vector<int> bc_kind;  // by boundary id
  // static stores (OceanDG.h:498-569), public order
  std::vector<double> cell0;  // [N][nq*nq][5]
  std::vector<double> cell1;  // [N][ng*ng]
  std::vector<double> cell2;  // [N][n1*n1][2]
  std::vector<double> facez;  // [n_inner][ng][2] (zb_m at x-1e-12n, zb_p at x+1e-12n)
  std::vector<double> bdryz;  // [n_bdry][nq]
  std::vector<double> bvals;  // [n_bdry][nq][3] boundary function values (set per stage)
  int lane_width = 1;
  int max_iteration_height = 15;
  // threads > 1: the loops run on OpenMP threads (CPU-baseline timing only; the face
  // scatter then accumulates atomically, so sums are not in a fixed order)
  int threads = 1;
  long last_sweeps = 0;

  Oracle(int degree) : p(degree)
  {
    n1 = p + 1;
    nq = (int)std::floor(1.5 * p) + 1;  // main.cc:216
    ng = p + 2;                         // OceanDG.h:448-450
    nd = n1 * n1;
    std::vector<double> w;
    gauss_lobatto_rule(n1, nodes, w);
    gauss_rule(nq, gq_x, gq_w);
    gauss_lobatto_rule(ng, gl_x, gl_w);
    gauss_rule(n1, gn_x, gn_w);
    Sq.resize(nq, n1);
    Dq.resize(nq, n1);
    for (int a = 0; a < nq; ++a)
      for (int i = 0; i < n1; ++i)
        {
          Sq(a, i) = lagrange(nodes, i, gq_x[a]);
          Dq(a, i) = lagrange_deriv(nodes, i, gq_x[a]);
        }
    Sg.resize(ng, n1);
    for (int s = 0; s < 2; ++s)
      Sgsub[s].resize(ng, n1);
    for (int a = 0; a < ng; ++a)
      for (int i = 0; i < n1; ++i)
        {
          Sg(a, i) = lagrange(nodes, i, gl_x[a]);
          for (int s = 0; s < 2; ++s)
            Sgsub[s](a, i) = lagrange(nodes, i, 0.5 * (gl_x[a] + s));
        }
    Sn.resize(n1, n1);
    for (int a = 0; a < n1; ++a)
      for (int i = 0; i < n1; ++i)
        Sn(a, i) = lagrange(nodes, i, gn_x[a]);
    Sninv = invert(Sn);
    bc_kind.assign(256, BC_NONE);
  }

  // -------------------- tensor-product kernels (deal.II FEEvaluation semantics) ----------
  // values: out[a + m*b] = sum_ij S(a,i) S(b,j) U[i + n1*j]
  void eval_values(const double *U, const Mat &S, double *out) const
  {
    const int m = S.r;
    std::vector<double> T((size_t)m * n1);
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < m; ++a)
        {
          double s = 0.0;
          for (int i = 0; i < n1; ++i)
            s += S(a, i) * U[i + n1 * j];
          T[a + m * j] = s;
        }
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a)
        {
          double s = 0.0;
          for (int j = 0; j < n1; ++j)
            s += S(b, j) * T[a + m * j];
          out[a + m * b] = s;
        }
  }
  // reference-coordinate gradient at Gauss(nq) points: gx = (D x S) U, gy = (S x D) U
  void eval_ref_gradients(const double *U, double *gx, double *gy) const
  {
    const int m = nq;
    std::vector<double> TS((size_t)m * n1), TD((size_t)m * n1);
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < m; ++a)
        {
          double s = 0.0, d = 0.0;
          for (int i = 0; i < n1; ++i)
            {
              s += Sq(a, i) * U[i + n1 * j];
              d += Dq(a, i) * U[i + n1 * j];
            }
          TS[a + m * j] = s;
          TD[a + m * j] = d;
        }
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a)
        {
          double sx = 0.0, sy = 0.0;
          for (int j = 0; j < n1; ++j)
            {
              sx += Sq(b, j) * TD[a + m * j];
              sy += Dq(b, j) * TS[a + m * j];
            }
          gx[a + m * b] = sx;
          gy[a + m * b] = sy;
        }
  }
  // R_ij += sum_ab S(a,i) S(b,j) v[a + m b]
  void integrate_values(const double *v, const Mat &S, double *R) const
  {
    const int m = S.r;
    std::vector<double> T((size_t)m * n1);
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < m; ++a)
        {
          double s = 0.0;
          for (int b = 0; b < m; ++b)
            s += S(b, j) * v[a + m * b];
          T[a + m * j] = s;
        }
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i)
        {
          double s = 0.0;
          for (int a = 0; a < m; ++a)
            s += S(a, i) * T[a + m * j];
          R[i + n1 * j] += s;
        }
  }
  // R_ij += sum_ab [D(a,i) S(b,j) gx + S(a,i) D(b,j) gy]   (reference-coordinate test grads)
  void integrate_ref_gradients(const double *gx, const double *gy, double *R) const
  {
    const int m = nq;
    std::vector<double> TS((size_t)m * n1), TD((size_t)m * n1);
    for (int j = 0; j < n1; ++j)
      for (int a = 0; a < m; ++a)
        {
          double s = 0.0, d = 0.0;
          for (int b = 0; b < m; ++b)
            {
              s += Sq(b, j) * gx[a + m * b];
              d += Dq(b, j) * gy[a + m * b];
            }
          TS[a + m * j] = s;  // to be tested with D in x
          TD[a + m * j] = d;  // to be tested with S in x
        }
    for (int j = 0; j < n1; ++j)
      for (int i = 0; i < n1; ++i)
        {
          double s = 0.0;
          for (int a = 0; a < m; ++a)
            s += Dq(a, i) * TS[a + m * j] + Sq(a, i) * TD[a + m * j];
          R[i + n1 * j] += s;
        }
  }

  // dof index of the k-th node along local face f
  int face_dof(int f, int k) const
  {
    switch (f)
      {
        case 0:
          return n1 * k;
        case 1:
          return (n1 - 1) + n1 * k;
        case 2:
          return k;
        default:
          return k + n1 * (n1 - 1);
      }
  }
  double face_length(int c, int f) const
  {
    return (f < 2) ? geom[4 * c + 3] : geom[4 * c + 2];
  }

  // trace of component array U (nd values) on face f at points of rule M (rows)
  void face_eval(const double *U, int f, const Mat &M, double *out) const
  {
    for (int a = 0; a < M.r; ++a)
      {
        double s = 0.0;
        for (int k = 0; k < n1; ++k)
          s += M(a, k) * U[face_dof(f, k)];
        out[a] = s;
      }
  }
  void face_integrate(const double *v, int f, const Mat &M, double *R) const
  {
    for (int k = 0; k < n1; ++k)
      {
        double s = 0.0;
        for (int a = 0; a < M.r; ++a)
          s += M(a, k) * v[a];
        if (threads > 1)
          {
#pragma omp atomic
            R[face_dof(f, k)] += s;
          }
        else
          R[face_dof(f, k)] += s;
      }
  }

  // ---------------------------- local_apply_cell_height (OceanDG.h:670-693) -----------
  void cell_height(const double *src_q, double *dst) const
  {
    const int m = nq, nqp = nq * nq;
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> qx(nqp), qy(nqp), fx(nqp), fy(nqp);
#pragma omp for schedule(static)
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        const double *Q = src_q + (size_t)c * 2 * nd;
        eval_values(Q, Sq, qx.data());
        eval_values(Q + nd, Sq, qy.data());
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              const int q = a + m * b;
              const double JxW = gq_w[a] * gq_w[b] * hx * hy;
              // submit_gradient(mass_flux(q)) = q; J^{-T} applied for the test gradient
              fx[q] = qx[q] / hx * JxW;
              fy[q] = qy[q] / hy * JxW;
            }
        integrate_ref_gradients(fx.data(), fy.data(), dst + (size_t)c * nd);
      }
    }
  }

  // ---------------------------- local_apply_cell_discharge (OceanDG.h:696-743) --------
  // nonstiff = true: local_apply_cell_nonstiff_discharge (OceanDG.h:746-791), friction out
  void cell_discharge(const double *src_h, const double *src_q, const double *src_v,
                      double *dst, bool nonstiff = false) const
  {
    const int m = nq, nqp = nq * nq;
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> z(nqp), zx(nqp), zy(nqp), qx(nqp), qy(nqp), uxx(nqp), uxy(nqp),
      uyx(nqp), uyy(nqp);
    std::vector<double> vx(nqp), vy(nqp), g0x(nqp), g0y(nqp), g1x(nqp), g1y(nqp);
#pragma omp for schedule(static)
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        const double *H = src_h + (size_t)c * nd;
        const double *Q = src_q + (size_t)c * 2 * nd;
        const double *U = src_v + (size_t)c * 2 * nd;
        eval_values(H, Sq, z.data());
        eval_ref_gradients(H, zx.data(), zy.data());
        eval_values(Q, Sq, qx.data());
        eval_values(Q + nd, Sq, qy.data());
        eval_ref_gradients(U, uxx.data(), uxy.data());
        eval_ref_gradients(U + nd, uyx.data(), uyy.data());
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              const int q = a + m * b;
              const double JxW = gq_w[a] * gq_w[b] * hx * hy;
              const double *par = &cell0[((size_t)c * nqp + q) * 5];
              const double zq = z[q];
              const V2 dz = {zx[q] / hx, zy[q] / hy};
              const V2 qq = {qx[q], qy[q]};
              const double h = model.depth(zq, par[0]);
              // du[e][d] = d u_e / d x_d; G = h * du
              const double G[2][2] = {{h * (uxx[q] / hx), h * (uxy[q] / hy)},
                                      {h * (uyx[q] / hx), h * (uyy[q] / hy)}};
              double F[2][2];
              model.advective_diffusive_flux(zq, qq, G, par[0], F);
              const V2 S = nonstiff ? model.source_nonstiff(zq, qq, dz, par)
                                    : model.source(zq, qq, dz, par);
              g0x[q] = F[0][0] / hx * JxW;
              g0y[q] = F[0][1] / hy * JxW;
              g1x[q] = F[1][0] / hx * JxW;
              g1y[q] = F[1][1] / hy * JxW;
              vx[q] = S.x * JxW;
              vy[q] = S.y * JxW;
            }
        double *R = dst + (size_t)c * 2 * nd;
        integrate_values(vx.data(), Sq, R);
        integrate_ref_gradients(g0x.data(), g0y.data(), R);
        integrate_values(vy.data(), Sq, R + nd);
        integrate_ref_gradients(g1x.data(), g1y.data(), R + nd);
      }
    }
  }

  // geometry of an inner face: JxW factor (face length of the finer side)
  double inner_face_length(const InnerFace &F) const
  {
    if (F.sm >= 0)
      return face_length(F.cp, F.fp);
    if (F.sp >= 0)
      return face_length(F.cm, F.fm);
    return face_length(F.cm, F.fm);
  }

  // traces of (zeta, qx, qy) of both sides at the GL(ng) points of inner face `fi`
  void inner_traces(const InnerFace &F, const double *src_h, const double *src_q, double *zm,
                    double *qxm, double *qym, double *zp, double *qxp, double *qyp) const
  {
    const Mat &Mm = F.sm >= 0 ? Sgsub[F.sm] : Sg;
    const Mat &Mp = F.sp >= 0 ? Sgsub[F.sp] : Sg;
    face_eval(src_h + (size_t)F.cm * nd, F.fm, Mm, zm);
    face_eval(src_q + (size_t)F.cm * 2 * nd, F.fm, Mm, qxm);
    face_eval(src_q + (size_t)F.cm * 2 * nd + nd, F.fm, Mm, qym);
    face_eval(src_h + (size_t)F.cp * nd, F.fp, Mp, zp);
    face_eval(src_q + (size_t)F.cp * 2 * nd, F.fp, Mp, qxp);
    face_eval(src_q + (size_t)F.cp * 2 * nd + nd, F.fp, Mp, qyp);
  }

  // ---------------------------- local_apply_face_height (OceanDG.h:937-982) -----------
  void face_height(const double *src_h, const double *src_q, double *dst) const
  {
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> zm(ng), qxm(ng), qym(ng), zp(ng), qxp(ng), qyp(ng), vm(ng), vp(ng);
#pragma omp for schedule(static)
    for (size_t fi = 0; fi < inner.size(); ++fi)
      {
        const InnerFace &F = inner[fi];
        inner_traces(F, src_h, src_q, zm.data(), qxm.data(), qym.data(), zp.data(),
                     qxp.data(), qyp.data());
        const double len = inner_face_length(F);
        const V2 n = outward_normal(F.fm);
        for (int a = 0; a < ng; ++a)
          {
            const double JxW = gl_w[a] * len;
            const double zbm = facez[(fi * ng + a) * 2 + 0];
            const double zbp = facez[(fi * ng + a) * 2 + 1];
            const double flux = model.numerical_massflux_weak(zm[a], zp[a], {qxm[a], qym[a]},
                                                              {qxp[a], qyp[a]}, n, zbm, zbp);
            vm[a] = -flux * JxW;
            vp[a] = flux * JxW;
          }
        face_integrate(vp.data(), F.fp, F.sp >= 0 ? Sgsub[F.sp] : Sg, dst + (size_t)F.cp * nd);
        face_integrate(vm.data(), F.fm, F.sm >= 0 ? Sgsub[F.sm] : Sg, dst + (size_t)F.cm * nd);
      }
    }
  }

  // ---------------------------- local_apply_face_discharge (OceanDG.h:985-1050) -------
  void face_discharge(const double *src_h, const double *src_q, double *dst) const
  {
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> zm(ng), qxm(ng), qym(ng), zp(ng), qxp(ng), qyp(ng);
    std::vector<double> vmx(ng), vmy(ng), vpx(ng), vpy(ng);
#pragma omp for schedule(static)
    for (size_t fi = 0; fi < inner.size(); ++fi)
      {
        const InnerFace &F = inner[fi];
        inner_traces(F, src_h, src_q, zm.data(), qxm.data(), qym.data(), zp.data(),
                     qxp.data(), qyp.data());
        const double len = inner_face_length(F);
        const V2 n = outward_normal(F.fm);
        for (int a = 0; a < ng; ++a)
          {
            const double JxW = gl_w[a] * len;
            const double zbm = facez[(fi * ng + a) * 2 + 0];
            const double zbp = facez[(fi * ng + a) * 2 + 1];
            const V2 fp_ = model.numerical_advflux_weak(zm[a], zp[a], {qxm[a], qym[a]},
                                                        {qxp[a], qyp[a]}, n, zbm, zbp);
            V2 fm_ = {-fp_.x, -fp_.y};
            const V2 pm = model.numerical_presflux_strong(zm[a], zp[a], n, zbm, zbp);
            const V2 pp = model.numerical_presflux_strong(zp[a], zm[a], n, zbp, zbm);
            fm_.x -= pm.x;
            fm_.y -= pm.y;
            const V2 fpp = {fp_.x + pp.x, fp_.y + pp.y};
            vmx[a] = fm_.x * JxW;
            vmy[a] = fm_.y * JxW;
            vpx[a] = fpp.x * JxW;
            vpy[a] = fpp.y * JxW;
          }
        const Mat &Mm = F.sm >= 0 ? Sgsub[F.sm] : Sg;
        const Mat &Mp = F.sp >= 0 ? Sgsub[F.sp] : Sg;
        face_integrate(vpx.data(), F.fp, Mp, dst + (size_t)F.cp * 2 * nd);
        face_integrate(vpy.data(), F.fp, Mp, dst + (size_t)F.cp * 2 * nd + nd);
        face_integrate(vmx.data(), F.fm, Mm, dst + (size_t)F.cm * 2 * nd);
        face_integrate(vmy.data(), F.fm, Mm, dst + (size_t)F.cm * 2 * nd + nd);
      }
    }
  }

  // exterior state on a boundary face (OceanDG.h:1130-1203 / 1246-1317)
  bool boundary_state(int kind, double z_m, V2 q_m, V2 n, double zb_m, const double *bv,
                      double &z_p, V2 &q_p) const
  {
    const double rho_u_dot_n = q_m.x * n.x + q_m.y * n.y;
    bool at_outflow = false;
    switch (kind)
      {
        case BC_WALL:
          z_p = z_m;
          q_p = {q_m.x - 2. * rho_u_dot_n * n.x, q_m.y - 2. * rho_u_dot_n * n.y};
          break;
        case BC_SUPERCRITICAL_INFLOW:
          {
            z_p = bv[0];
            const double h = model.depth(z_p, zb_m);
            q_p = {bv[1] * h, bv[2] * h};
            break;
          }
        case BC_SUPERCRITICAL_OUTFLOW:
          z_p = z_m;
          q_p = q_m;
          at_outflow = true;
          break;
        case BC_HEIGHT_INFLOW:
          z_p = bv[0];
          q_p = q_m;
          break;
        case BC_DISCHARGE_INFLOW:
          z_p = z_m;
          q_p = {bv[1] * -n.x, bv[1] * -n.y};
          break;
        case BC_ABSORBING_OUTFLOW:
          {
            z_p = bv[0];
            const double h = model.depth(z_p, zb_m);
            q_p = {bv[1] * h, bv[2] * h};
            const double r_p = model.riemann_invariant_p(z_m, q_m, n, zb_m);
            const double r_m = model.riemann_invariant_m(z_p, q_p, n, zb_m);
            const double c_b = 0.25 * (r_p - r_m);
            const double h_b = c_b * c_b / model.g;
            const double u_b = 0.5 * (r_p + r_m);
            z_p = h_b - zb_m;
            const double hh = model.depth(z_p, zb_m);
            q_p = {u_b * hh * n.x, u_b * hh * n.y};
            break;
          }
        default:
          throw std::string("Unknown boundary id, did you set a boundary condition for "
                            "this part of the domain boundary?");
      }
    return at_outflow;
  }

  // ------------------ local_apply_boundary_face_height / _discharge ------------------
  void bdry_face(const double *src_h, const double *src_q, double *dst_h, double *dst_q) const
  {
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> zm(nq), qxm(nq), qym(nq), vh(nq), vx(nq), vy(nq);
#pragma omp for schedule(static)
    for (size_t bi = 0; bi < bdry.size(); ++bi)
      {
        const BdryFace &B = bdry[bi];
        const int c = B.c, f = B.f;
        face_eval(src_h + (size_t)c * nd, f, Sq, zm.data());
        face_eval(src_q + (size_t)c * 2 * nd, f, Sq, qxm.data());
        face_eval(src_q + (size_t)c * 2 * nd + nd, f, Sq, qym.data());
        const double len = face_length(c, f);
        const V2 n = outward_normal(f);
        const int kind = bc_kind[B.bid];
        for (int a = 0; a < nq; ++a)
          {
            const double JxW = gq_w[a] * len;
            const double zb = bdryz[bi * nq + a];
            const double *bv = &bvals[(bi * nq + a) * 3];
            const V2 qm = {qxm[a], qym[a]};
            double z_p;
            V2 q_p;
            const bool at_outflow = boundary_state(kind, zm[a], qm, n, zb, bv, z_p, q_p);
            if (dst_h)
              {
                const double flux = model.numerical_massflux_weak(zm[a], z_p, qm, q_p, n, zb, zb);
                vh[a] = -flux * JxW;
              }
            if (dst_q)
              {
                V2 flux = model.numerical_advflux_weak(zm[a], z_p, qm, q_p, n, zb, zb);
                const V2 pr = model.numerical_presflux_strong(zm[a], z_p, n, zb, zb);
                flux.x += pr.x;
                flux.y += pr.y;
                if (at_outflow && (qm.x * n.x + qm.y * n.y) < -1e-12)
                  flux = {0., 0.};
                vx[a] = -flux.x * JxW;
                vy[a] = -flux.y * JxW;
              }
          }
        if (dst_h)
          face_integrate(vh.data(), f, Sq, dst_h + (size_t)c * nd);
        if (dst_q)
          {
            face_integrate(vx.data(), f, Sq, dst_q + (size_t)c * 2 * nd);
            face_integrate(vy.data(), f, Sq, dst_q + (size_t)c * 2 * nd + nd);
          }
      }
    }
  }

  // L_zeta and L_q (the two MatrixFree::loop calls of OceanDG.h:1805-1823, dst zeroed)
  void rhs(const double *src_h, const double *src_q, const double *src_v, double *Lh,
           double *Lq) const
  {
    std::fill(Lh, Lh + (size_t)N * nd, 0.0);
    cell_height(src_q, Lh);
    face_height(src_h, src_q, Lh);
    bdry_face(src_h, src_q, Lh, nullptr);
    std::fill(Lq, Lq + (size_t)N * 2 * nd, 0.0);
    cell_discharge(src_h, src_q, src_v, Lq);
    face_discharge(src_h, src_q, Lq);
    bdry_face(src_h, src_q, nullptr, Lq);
  }

  // ---------------------------- local_apply_cell_mass_height (OceanDG.h:829-855) ------
  void cell_mass_height(const double *src, double *dst) const
  {
    const int m = ng, nqp = ng * ng;
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> z(nqp), v(nqp);
#pragma omp for schedule(static)
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        eval_values(src + (size_t)c * nd, Sg, z.data());
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              const int q = a + m * b;
              const double JxW = gl_w[a] * gl_w[b] * hx * hy;
              v[q] = model.depth(z[q], cell1[(size_t)c * nqp + q]) * JxW;
            }
        integrate_values(v.data(), Sg, dst + (size_t)c * nd);
      }
    }
  }

  // ---------- local_apply_inverse_mass_matrix_height (OceanDG.h:1374-1507) ------------
  // dst holds the initial guess on entry (zeta^n) and the solution on exit; src = front.
  // Cells are batched in groups of `lane_width` consecutive cells, the reference's
  // VectorizedArray lanes: the dry lift is per lane, the l1 stop test sums the batch.
  void inverse_mass_height(double *dst, const double *src)
  {
    const int m = ng, nqp = ng * ng;
    const int W = lane_width;
    const int n_batches = (N + W - 1) / W;
    long sweeps = 0;
#pragma omp parallel num_threads(threads) if (threads > 1) reduction(+ : sweeps)
    {
    std::vector<double> z(nqp), v(nqp), rhs(nd), resid(nd), D(nd), dz(nd);
#pragma omp for schedule(static)
    for (int bb = 0; bb < n_batches; ++bb)
      {
        const int c0 = bb * W;
        const int nl = std::min(W, N - c0);
        // dry lift per lane (OceanDG.h:1399-1422)
        for (int l = 0; l < nl; ++l)
          {
            const int c = c0 + l;
            double *Z = dst + (size_t)c * nd;
            eval_values(Z, Sg, z.data());
            double zb_max_dry = -std::numeric_limits<double>::max();
            int n_dry = 0;
            for (int q = 0; q < nqp; ++q)
              {
                const double zb = cell1[(size_t)c * nqp + q];
                if (z[q] <= -zb)
                  {
                    zb_max_dry = std::max(zb, zb_max_dry);
                    ++n_dry;
                  }
              }
            if (n_dry == nqp)
              for (int i = 0; i < nd; ++i)
                Z[i] = -zb_max_dry + 1e-09;
          }
        double norm = 1.;
        for (int k = 0; (k < max_iteration_height) && (norm > 1e-12); ++k)
          {
            ++sweeps;
            norm = 0.;
            for (int l = 0; l < nl; ++l)
              {
                const int c = c0 + l;
                const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
                double *Z = dst + (size_t)c * nd;
                for (int i = 0; i < nd; ++i)
                  rhs[i] = src[(size_t)c * nd + i];
                eval_values(Z, Sg, z.data());
                for (int b = 0; b < m; ++b)
                  for (int a = 0; a < m; ++a)
                    {
                      const int q = a + m * b;
                      const double JxW = gl_w[a] * gl_w[b] * hx * hy;
                      v[q] = -model.depth(z[q], cell1[(size_t)c * nqp + q]) * JxW;
                    }
                integrate_values(v.data(), Sg, rhs.data());  // rhs = front - M h(z^k)
                for (int i = 0; i < nd; ++i)
                  resid[i] = Z[i];
                int n_dry = 0;
                for (int b = 0; b < m; ++b)
                  for (int a = 0; a < m; ++a)
                    {
                      const int q = a + m * b;
                      const double JxW = gl_w[a] * gl_w[b] * hx * hy;
                      const double mask = (z[q] <= -cell1[(size_t)c * nqp + q]) ? 0. : 1.;
                      n_dry += (mask == 0.);
                      v[q] = mask * JxW;
                    }
                std::fill(D.begin(), D.end(), 0.0);
                integrate_values(v.data(), Sg, D.data());
                // CellwiseInverseMassMatrixLumped::apply (operators.h:276-297)
                for (int i = 0; i < nd; ++i)
                  dz[i] = (n_dry < nqp) ? 1. / D[i] * rhs[i] : 0.;
                for (int i = 0; i < nd; ++i)
                  Z[i] += dz[i];  // distribute_local_to_global adds
                for (int i = 0; i < nd; ++i)
                  norm += std::fabs(resid[i] - Z[i]);
              }
          }
      }
    }
    last_sweeps += sweeps;
  }

  // ---- local_apply_inverse_mass_matrix_discharge (OceanDG.h:1533-1571): deal.II -------
  // CellwiseInverseMassMatrix: M^{-1} = (S^{-1} x S^{-1}) W^{-1} (S^{-T} x S^{-T}) with S the
  // GL-nodal basis at the Gauss(p+1) points and W = JxW there (exact on affine cells).
  void inverse_mass_tensor(const double *in, double *out, int n_comp) const
  {
#pragma omp parallel num_threads(threads) if (threads > 1)
    {
    std::vector<double> T(nd), X(nd);
#pragma omp for schedule(static)
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        for (int d = 0; d < n_comp; ++d)
          {
            const double *I = in + ((size_t)c * n_comp + d) * nd;
            double *O = out + ((size_t)c * n_comp + d) * nd;
            // X[a,b] = sum_ij Sinv(i,a) Sinv(j,b) I[i,j]
            for (int j = 0; j < n1; ++j)
              for (int a = 0; a < n1; ++a)
                {
                  double s = 0.0;
                  for (int i = 0; i < n1; ++i)
                    s += Sninv(i, a) * I[i + n1 * j];
                  T[a + n1 * j] = s;
                }
            for (int b = 0; b < n1; ++b)
              for (int a = 0; a < n1; ++a)
                {
                  double s = 0.0;
                  for (int j = 0; j < n1; ++j)
                    s += Sninv(j, b) * T[a + n1 * j];
                  X[a + n1 * b] = s / (gn_w[a] * gn_w[b] * hx * hy);
                }
            // O[i,j] = sum_ab Sinv(i,a) Sinv(j,b) X[a,b]
            for (int b = 0; b < n1; ++b)
              for (int i = 0; i < n1; ++i)
                {
                  double s = 0.0;
                  for (int a = 0; a < n1; ++a)
                    s += Sninv(i, a) * X[a + n1 * b];
                  T[i + n1 * b] = s;
                }
            for (int j = 0; j < n1; ++j)
              for (int i = 0; i < n1; ++i)
                {
                  double s = 0.0;
                  for (int b = 0; b < n1; ++b)
                    s += Sninv(j, b) * T[i + n1 * b];
                  O[i + n1 * j] = s;
                }
          }
      }
    }
  }

  // ------------------ local_apply_cell_stiff_discharge (OceanDG.h:794-826) ------------
  void cell_stiff_discharge(const double *src_h, const double *src_q, double *dst) const
  {
    const int m = nq, nqp = nq * nq;
    std::vector<double> z(nqp), qx(nqp), qy(nqp), vx(nqp), vy(nqp);
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        eval_values(src_h + (size_t)c * nd, Sq, z.data());
        eval_values(src_q + (size_t)c * 2 * nd, Sq, qx.data());
        eval_values(src_q + (size_t)c * 2 * nd + nd, Sq, qy.data());
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              const int q = a + m * b;
              const double JxW = gq_w[a] * gq_w[b] * hx * hy;
              const double *par = &cell0[((size_t)c * nqp + q) * 5];
              const V2 S = model.source_stiff(z[q], V2{qx[q], qy[q]}, par[0], par[1]);
              vx[q] = S.x * JxW;
              vy[q] = S.y * JxW;
            }
        integrate_values(vx.data(), Sq, dst + (size_t)c * 2 * nd);
        integrate_values(vy.data(), Sq, dst + (size_t)c * 2 * nd + nd);
      }
  }

  // ------------------ local_apply_cell_mass_discharge (OceanDG.h:858-879) -------------
  // dst += M src at Gauss(p+1) (exact on the affine cells)
  void cell_mass_discharge(const double *src, double *dst) const
  {
    const int m = n1, nqp = n1 * n1;
    std::vector<double> v(nqp);
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        for (int d = 0; d < 2; ++d)
          {
            eval_values(src + ((size_t)c * 2 + d) * nd, Sn, v.data());
            for (int b = 0; b < m; ++b)
              for (int a = 0; a < m; ++a)
                v[a + m * b] *= gn_w[a] * gn_w[b] * hx * hy;
            integrate_values(v.data(), Sn, dst + ((size_t)c * 2 + d) * nd);
          }
      }
  }

  // ------- local_apply_inverse_modified_mass_matrix_discharge (OceanDG.h:1574-1645) ---
  // out = (S^T W' S)^-1 in with W' = JxW (1 + fm J(v, n, h)) at Gauss(p+1), the state
  // (z, q) from ri_h, ri_q and (z_b, n) from cell_2 (FE degree >= 1 branch)
  void inverse_modified_mass(const double *in, const double *ri_h, const double *ri_q,
                             double fm, double *out) const
  {
    const int nqp = n1 * n1;
    std::vector<double> z(nqp), qx(nqp), qy(nqp), ijxw(nqp), T(nd), X(nd);
    for (int c = 0; c < N; ++c)
      {
        const double hx = geom[4 * c + 2], hy = geom[4 * c + 3];
        eval_values(ri_h + (size_t)c * nd, Sn, z.data());
        eval_values(ri_q + (size_t)c * 2 * nd, Sn, qx.data());
        eval_values(ri_q + (size_t)c * 2 * nd + nd, Sn, qy.data());
        for (int b = 0; b < n1; ++b)
          for (int a = 0; a < n1; ++a)
            {
              const int q = a + n1 * b;
              const double zb = cell2[((size_t)c * nqp + q) * 2];
              const double cf = cell2[((size_t)c * nqp + q) * 2 + 1];
              const V2 v = model.velocity(z[q], V2{qx[q], qy[q]}, zb);
              ijxw[q] = 1. / (gn_w[a] * gn_w[b] * hx * hy);
              ijxw[q] *= 1. / (1. + fm * model.friction_jacobian(v, cf, model.depth(z[q], zb)));
            }
        for (int d = 0; d < 2; ++d)
          {
            const double *I = in + ((size_t)c * 2 + d) * nd;
            double *O = out + ((size_t)c * 2 + d) * nd;
            for (int j = 0; j < n1; ++j)
              for (int a = 0; a < n1; ++a)
                {
                  double s = 0.0;
                  for (int i = 0; i < n1; ++i)
                    s += Sninv(i, a) * I[i + n1 * j];
                  T[a + n1 * j] = s;
                }
            for (int b = 0; b < n1; ++b)
              for (int a = 0; a < n1; ++a)
                {
                  double s = 0.0;
                  for (int j = 0; j < n1; ++j)
                    s += Sninv(j, b) * T[a + n1 * j];
                  X[a + n1 * b] = s * ijxw[a + n1 * b];
                }
            for (int b = 0; b < n1; ++b)
              for (int i = 0; i < n1; ++i)
                {
                  double s = 0.0;
                  for (int a = 0; a < n1; ++a)
                    s += Sninv(i, a) * X[a + n1 * b];
                  T[i + n1 * b] = s;
                }
            for (int j = 0; j < n1; ++j)
              for (int i = 0; i < n1; ++i)
                {
                  double s = 0.0;
                  for (int b = 0; b < n1; ++b)
                    s += Sninv(j, b) * T[i + n1 * b];
                  O[i + n1 * j] = s;
                }
          }
      }
  }

  // ------------- perform_stage_hydro, additive (IMEX) RK (OceanDG.h:1936-2119) --------
  // ki_h: S+1 vectors; ki_q: 2S vectors ([0] front, [2j+1] explicit L_q of stage j,
  // [2j+2] stiff friction term of stage j < S-1); fm = d_tilde[stage] dt.
  void perform_stage_hydro_ark(int stage, int n_stages, const double *factor,
                               const double *factor_tilde, double fm, const double *ri_h,
                               const double *ri_q, const double *ri_v, double *const *ki_h,
                               double *const *ki_q, double *w_h, double *w_q, double *r_h,
                               double *r_q)
  {
    const size_t NH = (size_t)N * nd, NQ = (size_t)N * 2 * nd;
    const bool last = stage == n_stages - 1;
    last_sweeps = 0;
    double *Lh = ki_h[stage + 1];
    std::fill(Lh, Lh + NH, 0.0);
    cell_height(ri_q, Lh);
    face_height(ri_h, ri_q, Lh);
    bdry_face(ri_h, ri_q, Lh, nullptr);
    double *kex = ki_q[2 * stage + 1];
    std::fill(kex, kex + NQ, 0.0);
    cell_discharge(ri_h, ri_q, ri_v, kex, !last);
    face_discharge(ri_h, ri_q, kex);
    bdry_face(ri_h, ri_q, nullptr, kex);
    if (!last)
      {
        double *kim = ki_q[2 * stage + 2];
        std::fill(kim, kim + NQ, 0.0);
        cell_stiff_discharge(ri_h, ri_q, kim);
      }
    for (size_t i = 0; i < NH; ++i)
      {
        double k_i = ki_h[1][i];
        ki_h[0][i] = factor[0] * k_i;
        for (int j = 1; j < stage + 1; ++j)
          {
            k_i = ki_h[j + 1][i];
            ki_h[0][i] += factor[j] * k_i;
          }
      }
    cell_mass_height(w_h, ki_h[0]);
    if (last)
      inverse_mass_height(w_h, ki_h[0]);
    else
      {
        std::memcpy(r_h, w_h, NH * sizeof(double));
        inverse_mass_height(r_h, ki_h[0]);
      }
    if (last)
      {
        for (size_t i = 0; i < NQ; ++i)
          {
            double kex_i = ki_q[1][i];
            ki_q[0][i] = factor[0] * kex_i;
            for (int j = 1; j < stage + 1; ++j)
              {
                kex_i = ki_q[2 * j + 1][i];
                const double kim_i = ki_q[2 * j][i];
                ki_q[0][i] += factor[j] * kex_i + factor[j - 1] * kim_i;
              }
          }
        inverse_mass_tensor(ki_q[0], r_q, 2);
        for (size_t i = 0; i < NQ; ++i)
          {
            const double sol_i = r_q[i];
            w_q[i] += sol_i;
          }
      }
    else
      {
        for (size_t i = 0; i < NQ; ++i)
          {
            double kex_i = ki_q[1][i];
            double kim_i = ki_q[2][i];
            ki_q[0][i] = factor[0] * kex_i + factor_tilde[0] * kim_i;
            for (int j = 1; j < stage + 1; ++j)
              {
                kex_i = ki_q[2 * j + 1][i];
                kim_i = ki_q[2 * j + 2][i];
                ki_q[0][i] += factor[j] * kex_i + factor_tilde[j] * kim_i;
              }
          }
        cell_mass_discharge(w_q, ki_q[0]);
        inverse_modified_mass(ki_q[0], ri_h, ri_q, fm, r_q);
      }
  }

  // ------------------ perform_stage_hydro, explicit RK (OceanDG.h:1782-1912) ----------
  void perform_stage_hydro(int stage, int n_ki, const double *factor, const double *ri_h,
                           const double *ri_q, const double *ri_v, double *const *ki_h,
                           double *const *ki_q, double *w_h, double *w_q, double *r_h,
                           double *r_q)
  {
    const int n_stages = n_ki - 1;
    const size_t NH = (size_t)N * nd, NQ = (size_t)N * 2 * nd;
    last_sweeps = 0;
    rhs(ri_h, ri_q, ri_v, ki_h[stage + 1], ki_q[0]);
    // pre-op: front = sum_j factor[j] * ki_h[j+1]  (OceanDG.h:1835-1847)
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (size_t i = 0; i < NH; ++i)
      {
        double k_i = ki_h[1][i];
        ki_h[0][i] = factor[0] * k_i;
        for (int j = 1; j < stage + 1; ++j)
          {
            k_i = ki_h[j + 1][i];
            ki_h[0][i] += factor[j] * k_i;
          }
      }
    cell_mass_height(w_h, ki_h[0]);
    if (stage == n_stages - 1)
      inverse_mass_height(w_h, ki_h[0]);
    else
      {
        std::memcpy(r_h, w_h, NH * sizeof(double));
        inverse_mass_height(r_h, ki_h[0]);
      }
    inverse_mass_tensor(ki_q[0], ki_q[stage + 1], 2);
    // post-op (OceanDG.h:1878-1909)
    double *dst = (stage == n_stages - 1) ? w_q : r_q;
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
    for (size_t i = 0; i < NQ; ++i)
      {
        double k_i = ki_q[1][i];
        const double sol_i = w_q[i];
        dst[i] = sol_i + factor[0] * k_i;
        for (int j = 1; j < stage + 1; ++j)
          {
            k_i = ki_q[j + 1][i];
            dst[i] += factor[j] * k_i;
          }
      }
  }
};

thread_local std::string g_err;

}  // namespace

extern "C" {

const char *ocr_last_error() { return g_err.c_str(); }

// 1-D rules: kind 0 = Gauss, 1 = Gauss-Lobatto; on [0,1], ascending
void ocr_rule(int kind, int n, double *x, double *w)
{
  std::vector<double> xx, ww;
  if (kind == 0)
    gauss_rule(n, xx, ww);
  else
    gauss_lobatto_rule(n, xx, ww);
  for (int i = 0; i < n; ++i)
    {
      x[i] = xx[i];
      w[i] = ww[i];
    }
}

// Lagrange basis on `nn` nodes evaluated at `npts` points: val/der [npts][nn]
void ocr_lagrange(int nn, const double *nodes, int npts, const double *pts, double *val,
                  double *der)
{
  std::vector<double> z(nodes, nodes + nn);
  for (int a = 0; a < npts; ++a)
    for (int k = 0; k < nn; ++k)
      {
        val[a * nn + k] = lagrange(z, k, pts[a]);
        der[a * nn + k] = lagrange_deriv(z, k, pts[a]);
      }
}

// inner: [n_inner][6] = cm, fm, sm, cp, fp, sp ; bdry: [n_bdry][3] = c, f, bid
// bc_kind: [256] kind per boundary id; params: g, threshold_for_wetdry, viscosity
void *ocr_create(int degree, int n_cells, const double *geom, int n_inner, const int *inner,
                 int n_bdry, const int *bdry, const int *bc_kind, const double *params)
{
  try
    {
      Oracle *o = new Oracle(degree);
      o->N = n_cells;
      o->geom.assign(geom, geom + 4 * (size_t)n_cells);
      o->inner.resize(n_inner);
      for (int i = 0; i < n_inner; ++i)
        o->inner[i] = {inner[6 * i], inner[6 * i + 1], inner[6 * i + 2],
                       inner[6 * i + 3], inner[6 * i + 4], inner[6 * i + 5]};
      o->bdry.resize(n_bdry);
      for (int i = 0; i < n_bdry; ++i)
        o->bdry[i] = {bdry[3 * i], bdry[3 * i + 1], bdry[3 * i + 2]};
      for (int i = 0; i < 256; ++i)
        o->bc_kind[i] = bc_kind[i];
      o->model.g = params[0];
      const double cu = params[1];
      o->model.cu4 = cu * cu * cu * cu;  // ShallowWater.h:295-296
      o->model.nu = params[2];
      const int nq = o->nq, ng = o->ng, n1 = o->n1;
      o->cell0.assign((size_t)n_cells * nq * nq * 5, 0.0);
      o->cell1.assign((size_t)n_cells * ng * ng, 0.0);
      o->cell2.assign((size_t)n_cells * n1 * n1 * 2, 0.0);
      o->facez.assign((size_t)n_inner * ng * 2, 0.0);
      o->bdryz.assign((size_t)n_bdry * nq, 0.0);
      o->bvals.assign((size_t)n_bdry * nq * 3, 0.0);
      return o;
    }
  catch (...)
    {
      g_err = "ocr_create failed";
      return nullptr;
    }
}

void ocr_destroy(void *h) { delete static_cast<Oracle *>(h); }

// store: 0 cell0 [N][nq^2][5], 1 cell1 [N][ng^2], 2 cell2 [N][n1^2][2],
//        3 face [n_inner][ng][2], 4 boundary [n_bdry][nq], 5 boundary values [n_bdry][nq][3]
int ocr_set_static(void *h, int store, const double *values)
{
  Oracle *o = static_cast<Oracle *>(h);
  std::vector<double> *v = nullptr;
  switch (store)
    {
      case 0:
        v = &o->cell0;
        break;
      case 1:
        v = &o->cell1;
        break;
      case 2:
        v = &o->cell2;
        break;
      case 3:
        v = &o->facez;
        break;
      case 4:
        v = &o->bdryz;
        break;
      case 5:
        v = &o->bvals;
        break;
      default:
        g_err = "unknown store";
        return 1;
    }
  std::copy(values, values + v->size(), v->begin());
  return 0;
}

void ocr_set_lane_width(void *h, int w) { static_cast<Oracle *>(h)->lane_width = w; }
// loops on n OpenMP threads (n > 1: face sums accumulate atomically; CPU-baseline timing)
void ocr_set_threads(void *h, int n) { static_cast<Oracle *>(h)->threads = n < 1 ? 1 : n; }
void ocr_set_max_iterations(void *h, int k)
{
  static_cast<Oracle *>(h)->max_iteration_height = k;
}
long ocr_last_sweeps(void *h) { return static_cast<Oracle *>(h)->last_sweeps; }

int ocr_rhs(void *h, const double *src_h, const double *src_q, const double *src_v, double *Lh,
            double *Lq)
{
  try
    {
      static_cast<Oracle *>(h)->rhs(src_h, src_q, src_v, Lh, Lq);
      return 0;
    }
  catch (const std::string &s)
    {
      g_err = s;
      return 2;
    }
}

int ocr_cell_mass_height(void *h, const double *src, double *dst)
{
  static_cast<Oracle *>(h)->cell_mass_height(src, dst);
  return 0;
}

int ocr_inverse_mass_height(void *h, double *dst, const double *src)
{
  Oracle *o = static_cast<Oracle *>(h);
  o->last_sweeps = 0;
  o->inverse_mass_height(dst, src);
  return 0;
}

int ocr_inverse_mass_tensor(void *h, const double *in, double *out, int n_comp)
{
  static_cast<Oracle *>(h)->inverse_mass_tensor(in, out, n_comp);
  return 0;
}

int ocr_perform_stage_hydro(void *h, int stage, int n_ki, const double *factor,
                            const double *ri_h, const double *ri_q, const double *ri_v,
                            double *const *ki_h, double *const *ki_q, double *w_h, double *w_q,
                            double *r_h, double *r_q)
{
  try
    {
      static_cast<Oracle *>(h)->perform_stage_hydro(stage, n_ki, factor, ri_h, ri_q, ri_v, ki_h,
                                                    ki_q, w_h, w_q, r_h, r_q);
      return 0;
    }
  catch (const std::string &s)
    {
      g_err = s;
      return 2;
    }
}

// quadrature points of the static stores, in the same order as the stores
// (OceanDG.h:498-569; FEEvaluation::quadrature_point on the affine cell)
int ocr_quadrature_points(void *h, int store, double *xy)
{
  Oracle *o = static_cast<Oracle *>(h);
  auto cell_rule = [&](const std::vector<double> &x) {
    const int m = (int)x.size();
    size_t k = 0;
    for (int c = 0; c < o->N; ++c)
      {
        const double *g = &o->geom[4 * c];
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              xy[k++] = g[0] + g[2] * x[a];
              xy[k++] = g[1] + g[3] * x[b];
            }
      }
  };
  auto face_pt = [&](int c, int f, double t, double &x, double &y) {
    const double *g = &o->geom[4 * c];
    x = (f == 0) ? g[0] : (f == 1) ? g[0] + g[2] : g[0] + t * g[2];
    y = (f == 2) ? g[1] : (f == 3) ? g[1] + g[3] : g[1] + t * g[3];
  };
  switch (store)
    {
      case 0:
        cell_rule(o->gq_x);
        return 0;
      case 1:
        cell_rule(o->gl_x);
        return 0;
      case 2:
        cell_rule(o->gn_x);
        return 0;
      case 3:
        {
          size_t k = 0;
          for (const InnerFace &F : o->inner)
            {
              const V2 n = outward_normal(F.fm);
              for (int a = 0; a < o->ng; ++a)
                {
                  double x, y;
                  face_pt(F.cm, F.fm, F.sm >= 0 ? 0.5 * (o->gl_x[a] + F.sm) : o->gl_x[a], x, y);
                  xy[k++] = x - 1e-12 * n.x;
                  xy[k++] = y - 1e-12 * n.y;
                  face_pt(F.cp, F.fp, F.sp >= 0 ? 0.5 * (o->gl_x[a] + F.sp) : o->gl_x[a], x, y);
                  xy[k++] = x + 1e-12 * n.x;
                  xy[k++] = y + 1e-12 * n.y;
                }
            }
          return 0;
        }
      case 4:
        {
          size_t k = 0;
          for (const BdryFace &B : o->bdry)
            for (int a = 0; a < o->nq; ++a)
              {
                double x, y;
                face_pt(B.c, B.f, o->gq_x[a], x, y);
                xy[k++] = x;
                xy[k++] = y;
              }
          return 0;
        }
    }
  g_err = "unknown store";
  return 1;
}

// transform_from_q_points_to_basis at the Gauss(p+1) collocation points:
// values [N][n1^2][n_comp] -> coefficients [N][n_comp][n1^2]
int ocr_project_collocation(void *h, int n_comp, const double *values, double *coeffs)
{
  Oracle *o = static_cast<Oracle *>(h);
  const int n1 = o->n1, nd = o->nd;
  std::vector<double> T(nd);
  for (int c = 0; c < o->N; ++c)
    for (int d = 0; d < n_comp; ++d)
      {
        const double *u = values + (size_t)c * nd * n_comp;
        double *U = coeffs + ((size_t)c * n_comp + d) * nd;
        for (int b = 0; b < n1; ++b)
          for (int i = 0; i < n1; ++i)
            {
              double s = 0;
              for (int a = 0; a < n1; ++a)
                s += o->Sninv(i, a) * u[(a + n1 * b) * n_comp + d];
              T[i + n1 * b] = s;
            }
        for (int j = 0; j < n1; ++j)
          for (int i = 0; i < n1; ++i)
            {
              double s = 0;
              for (int b = 0; b < n1; ++b)
                s += o->Sninv(j, b) * T[i + n1 * b];
              U[i + n1 * j] = s;
            }
      }
  return 0;
}

// compute_errors_hydro (OceanDG.h:2382-2459) with per-field results:
// exact [N][nq^2][3] (free surface, u, v at the cell_0 points); err[0] = L2 of the depth
// error, err[1] = L2 of the discharge error (hu, hv together, as the reference),
// err[2], err[3] = L2 of hu and hv separately, err[4..6] = Linf of depth, hu, hv.
int ocr_errors(void *h, const double *zeta, const double *q, const double *exact, double *err)
{
  Oracle *o = static_cast<Oracle *>(h);
  const int m = o->nq, np = m * m, nd = o->nd;
  std::vector<double> z(np), qx(np), qy(np);
  double e2[3] = {0, 0, 0}, einf[3] = {0, 0, 0};
  for (int c = 0; c < o->N; ++c)
    {
      const double hx = o->geom[4 * c + 2], hy = o->geom[4 * c + 3];
      o->eval_values(zeta + (size_t)c * nd, o->Sq, z.data());
      o->eval_values(q + (size_t)c * 2 * nd, o->Sq, qx.data());
      o->eval_values(q + (size_t)c * 2 * nd + nd, o->Sq, qy.data());
      for (int b = 0; b < m; ++b)
        for (int a = 0; a < m; ++a)
          {
            const int qp = a + m * b;
            const double JxW = o->gq_w[a] * o->gq_w[b] * hx * hy;
            const double zb = o->cell0[((size_t)c * np + qp) * 5];
            const double *ex = exact + ((size_t)c * np + qp) * 3;
            const double hex = o->model.depth(ex[0], zb);
            const double d0 = hex - o->model.depth(z[qp], zb);
            const double d1 = ex[1] * hex - qx[qp];
            const double d2 = ex[2] * hex - qy[qp];
            e2[0] += d0 * d0 * JxW;
            e2[1] += d1 * d1 * JxW;
            e2[2] += d2 * d2 * JxW;
            einf[0] = std::max(einf[0], std::fabs(d0));
            einf[1] = std::max(einf[1], std::fabs(d1));
            einf[2] = std::max(einf[2], std::fabs(d2));
          }
    }
  err[0] = std::sqrt(e2[0]);
  err[1] = std::sqrt(e2[1] + e2[2]);
  err[2] = std::sqrt(e2[1]);
  err[3] = std::sqrt(e2[2]);
  err[4] = einf[0];
  err[5] = einf[1];
  err[6] = einf[2];
  return 0;
}

int ocr_perform_stage_hydro_ark(void *h, int stage, int n_stages, const double *factor,
                                const double *factor_tilde, double fm, const double *ri_h,
                                const double *ri_q, const double *ri_v, double *const *ki_h,
                                double *const *ki_q, double *w_h, double *w_q, double *r_h,
                                double *r_q)
{
  static_cast<Oracle *>(h)->perform_stage_hydro_ark(stage, n_stages, factor, factor_tilde, fm,
                                                    ri_h, ri_q, ri_v, ki_h, ki_q, w_h, w_q,
                                                    r_h, r_q);
  return 0;
}

// evaluate_velocity_field (OceanDG.h:2600-2662), FE_DGQ degree >= 1 branch: velocity at the
// Gauss(p+1) points with z_b from cell_2, then transform_from_q_points_to_basis (the inverse
// of the Gauss(p+1) collocation).  zeta [N][nd], q and vel [N][2][nd], cell-major.
int ocr_velocity_field(void *h, const double *zeta, const double *q, double *vel)
{
  Oracle *o = static_cast<Oracle *>(h);
  const int n1 = o->n1, nd = o->nd;
  std::vector<double> z(nd), qx(nd), qy(nd), vq(2 * nd);
  for (int c = 0; c < o->N; ++c)
    {
      o->eval_values(zeta + (size_t)c * nd, o->Sn, z.data());
      o->eval_values(q + (size_t)c * 2 * nd, o->Sn, qx.data());
      o->eval_values(q + (size_t)c * 2 * nd + nd, o->Sn, qy.data());
      for (int k = 0; k < nd; ++k)
        {
          const double zb = o->cell2[((size_t)c * nd + k) * 2];
          const V2 v = o->model.velocity(z[k], V2{qx[k], qy[k]}, zb);
          vq[2 * k] = v.x;
          vq[2 * k + 1] = v.y;
        }
      std::vector<double> T(nd);
      for (int d = 0; d < 2; ++d)
        {
          double *U = vel + ((size_t)c * 2 + d) * nd;
          for (int b = 0; b < n1; ++b)
            for (int i = 0; i < n1; ++i)
              {
                double s = 0;
                for (int a = 0; a < n1; ++a)
                  s += o->Sninv(i, a) * vq[2 * (a + n1 * b) + d];
                T[i + n1 * b] = s;
              }
          for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i)
              {
                double s = 0;
                for (int b = 0; b < n1; ++b)
                  s += o->Sninv(j, b) * T[i + n1 * b];
                U[i + n1 * j] = s;
              }
        }
    }
  return 0;
}

// compute_cell_transport_speed (OceanDG.h:2506-2580) at the GL(p+2) points with z_b from
// cell_1, for axis-aligned cells (inverse Jacobian diag(1/hx, 1/hy)); the five power
// iterations of the largest eigenvalue of J^-T J^-1 are kept as written.
double ocr_transport_speed(void *h, const double *zeta, const double *q)
{
  Oracle *o = static_cast<Oracle *>(h);
  const int m = o->ng, np = m * m, nd = o->nd;
  std::vector<double> z(np), qx(np), qy(np);
  double max_transport = 0.0;
  for (int c = 0; c < o->N; ++c)
    {
      const double ij[2] = {1.0 / o->geom[4 * c + 2], 1.0 / o->geom[4 * c + 3]};
      o->eval_values(zeta + (size_t)c * nd, o->Sg, z.data());
      o->eval_values(q + (size_t)c * 2 * nd, o->Sg, qx.data());
      o->eval_values(q + (size_t)c * 2 * nd + nd, o->Sg, qy.data());
      double local_max = 0.0;
      for (int k = 0; k < np; ++k)
        {
          const double zb = o->cell1[(size_t)c * np + k];
          const V2 v = o->model.velocity(z[k], V2{qx[k], qy[k]}, zb);
          const double cs[2] = {ij[0] * v.x, ij[1] * v.y};
          double conv = 0.0;
          for (int d = 0; d < 2; ++d)
            conv = std::max(conv, std::fabs(cs[d]));
          const double sos = std::sqrt(o->model.square_wavespeed(z[k], zb));
          double ev[2] = {1.0, 1.0};
          for (int i = 0; i < 5; ++i)
            {
              for (int d = 0; d < 2; ++d)
                ev[d] = ij[d] * (ij[d] * ev[d]);
              double nrm = 0.0;
              for (int d = 0; d < 2; ++d)
                nrm = std::max(nrm, std::fabs(ev[d]));
              for (int d = 0; d < 2; ++d)
                ev[d] /= nrm;
            }
          const double je[2] = {ij[0] * ev[0], ij[1] * ev[1]};
          const double lam = std::sqrt((je[0] * je[0] + je[1] * je[1]) /
                                       (ev[0] * ev[0] + ev[1] * ev[1]));
          local_max = std::max(local_max, lam * sos + conv);
        }
      max_transport = std::max(max_transport, local_max);
    }
  return max_transport;
}

// check_mass (OceanDG.h:2138-2184), cell part: sum over the DoFs of
// local_apply_cell_mass_height(zeta) = the integral of the depth at GL(p+2)^2
double ocr_volume(void *h, const double *zeta)
{
  Oracle *o = static_cast<Oracle *>(h);
  std::vector<double> I((size_t)o->N * o->nd, 0.0);
  o->cell_mass_height(zeta, I.data());
  double s = 0.0;
  for (double v : I)
    s += v;
  return s;
}

// The exterior state of one boundary quadrature point (OceanDG.h:1130-1203): kind as in
// bc_kind, interior state (z_m, q_m), outward normal n, bed z_b, boundary data bv[3].
// out = {z_p, qx_p, qy_p}; returns 1 at a supercritical outflow (where the momentum flux
// is cut under backflow, OceanDG.h:1326-1332), 0 otherwise, -1 for an unknown kind.
int ocr_boundary_state(void *h, int kind, double z_m, double qx_m, double qy_m, double nx,
                       double ny, double zb, const double *bv, double *out)
{
  Oracle *o = static_cast<Oracle *>(h);
  try
    {
      double z_p;
      V2 q_p;
      const bool at_outflow = o->boundary_state(kind, z_m, {qx_m, qy_m}, {nx, ny}, zb, bv, z_p,
                                                q_p);
      out[0] = z_p;
      out[1] = q_p.x;
      out[2] = q_p.y;
      return at_outflow ? 1 : 0;
    }
  catch (const std::string &e)
    {
      g_err = e;
      return -1;
    }
}

}  // extern "C"
