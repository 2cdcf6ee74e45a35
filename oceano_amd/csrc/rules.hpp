// rules.hpp -- 1-D quadrature rules and the Gauss-Lobatto nodal basis used by the operator.
//
// FE_DGQ(p) in deal.II is the Lagrange basis on the p+1 Gauss-Lobatto points of [0,1];
// the operator integrates with QGauss(floor(1.5p)+1) on cells and boundary faces,
// QGaussLobatto(p+2) on inner faces and the depth mass, and QGauss(p+1) for the
// tensor-product inverse mass (OceanDG.h:438-470, main.cc:216).
#pragma once

#include <cmath>
#include <vector>

namespace oceano {

struct Rule1D
{
  std::vector<double> x;  // ascending on [0,1]
  std::vector<double> w;
};

// P_n, P_n' at t in (-1,1); recurrence (k+1) P_{k+1} = (2k+1) t P_k - k P_{k-1}
inline void legendre_value_slope(int n, double t, double &pn, double &dpn, double &pnm1)
{
  double a = 1.0, b = t;
  if (n == 0)
    {
      pn = 1.0;
      dpn = 0.0;
      pnm1 = 0.0;
      return;
    }
  for (int k = 1; k < n; ++k)
    {
      const double c = ((2 * k + 1) * t * b - k * a) / (k + 1);
      a = b;
      b = c;
    }
  pn = b;
  pnm1 = a;
  dpn = n * (a - t * b) / (1.0 - t * t);
}

inline Rule1D gauss(int n)
{
  Rule1D r;
  r.x.resize(n);
  r.w.resize(n);
  for (int i = 0; i < n; ++i)
    {
      // i-th root from below; initial guess -cos(pi (i + 3/4) / (n + 1/2))
      double t = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
      for (int it = 0; it < 64; ++it)
        {
          double p, dp, pm;
          legendre_value_slope(n, t, p, dp, pm);
          const double step = p / dp;
          t -= step;
          if (std::abs(step) <= 1e-17)
            break;
        }
      double p, dp, pm;
      legendre_value_slope(n, t, p, dp, pm);
      r.x[i] = 0.5 * (t + 1.0);
      r.w[i] = 1.0 / ((1.0 - t * t) * dp * dp);  // (2/((1-t^2) P'^2)) / 2
    }
  return r;
}

inline Rule1D gauss_lobatto(int n)
{
  Rule1D r;
  r.x.resize(n);
  r.w.resize(n);
  const int m = n - 1;  // interior nodes are the roots of P_m'
  for (int i = 0; i < n; ++i)
    {
      double t;
      if (i == 0)
        t = -1.0;
      else if (i == n - 1)
        t = 1.0;
      else
        {
          t = -std::cos(M_PI * i / m);
          for (int it = 0; it < 64; ++it)
            {
              double p, dp, pm;
              legendre_value_slope(m, t, p, dp, pm);
              const double d2p = (2.0 * t * dp - m * (m + 1.0) * p) / (1.0 - t * t);
              const double step = dp / d2p;
              t -= step;
              if (std::abs(step) <= 1e-17)
                break;
            }
        }
      double pmv;
      if (i == 0)
        pmv = (m % 2) ? -1.0 : 1.0;
      else if (i == n - 1)
        pmv = 1.0;
      else
        {
          double dp, pm;
          legendre_value_slope(m, t, pmv, dp, pm);
        }
      r.x[i] = 0.5 * (t + 1.0);
      r.w[i] = 1.0 / (n * (n - 1.0) * pmv * pmv);
    }
  return r;
}

// l_k(x) on nodes z and its derivative
inline double lagrange_value(const std::vector<double> &z, int k, double x)
{
  double v = 1.0;
  for (int m = 0; m < (int)z.size(); ++m)
    if (m != k)
      v *= (x - z[m]) / (z[k] - z[m]);
  return v;
}

inline double lagrange_slope(const std::vector<double> &z, int k, double x)
{
  double s = 0.0;
  for (int m = 0; m < (int)z.size(); ++m)
    {
      if (m == k)
        continue;
      double v = 1.0 / (z[k] - z[m]);
      for (int j = 0; j < (int)z.size(); ++j)
        if (j != k && j != m)
          v *= (x - z[j]) / (z[k] - z[j]);
      s += v;
    }
  return s;
}

// inverse of a small dense matrix (row-major, n x n), Gauss-Jordan with pivoting
inline std::vector<double> small_inverse(std::vector<double> a, int n)
{
  std::vector<double> inv(n * n, 0.0);
  for (int i = 0; i < n; ++i)
    inv[i * n + i] = 1.0;
  for (int c = 0; c < n; ++c)
    {
      int piv = c;
      for (int i = c + 1; i < n; ++i)
        if (std::abs(a[i * n + c]) > std::abs(a[piv * n + c]))
          piv = i;
      for (int j = 0; j < n; ++j)
        {
          std::swap(a[c * n + j], a[piv * n + j]);
          std::swap(inv[c * n + j], inv[piv * n + j]);
        }
      const double d = 1.0 / a[c * n + c];
      for (int j = 0; j < n; ++j)
        {
          a[c * n + j] *= d;
          inv[c * n + j] *= d;
        }
      for (int i = 0; i < n; ++i)
        if (i != c)
          {
            const double f = a[i * n + c];
            for (int j = 0; j < n; ++j)
              {
                a[i * n + j] -= f * a[c * n + j];
                inv[i * n + j] -= f * inv[c * n + j];
              }
          }
    }
  return inv;
}

}  // namespace oceano
