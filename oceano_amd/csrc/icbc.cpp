// icbc.cpp -- test-case functions evaluated by the host at quadrature points.
//
// Travelling vortex: the exact solution of Icbc_ShallowWaterVortex.h:139-226 (the
// compactly supported cos^4 vortex, regularity branch p = 2), with centre, radius,
// amplitude and background velocity as parameters instead of the file's constants
// (h0 = 1, uoo = 6, hmin = 0.9, radius0 = 0.25, centre (0.5, 0.5)), and optional periodic
// images.  Channel bed: the subcritical SWASHES bump plus seeded sub-cell bumps, in the
// reference's bathymetry convention z_b = depth of the bed below the datum (h = zeta + z_b,
// Icbc_ChannelFlow.h:126-144).

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "oceano_dg.h"

namespace {

// H(x) of the p = 2 vortex, times corr = 4 (Icbc_ShallowWaterVortex.h:185-197)
double vortex_H(double X)
{
  const double cosx = std::cos(X), sinx = std::sin(X);
  const double cos2x = std::cos(2. * X), sin2x = std::sin(2. * X);
  const double cosx2 = cosx * cosx, cos2x2 = cos2x * cos2x, x2 = X * X;
  const double H = 0.091145833 * cos2x + 0.182291667 * X * sin2x +
                   std::pow(cosx, 6) * (0.015625 * cosx2 + 0.024305556) + 0.011393229 * cos2x2 +
                   0.13671875 * x2 + 0.045572917 * X * cos2x * sin2x +
                   0.125 * X * std::pow(cosx, 5) * sinx * (cosx2 + 1.166666667);
  return 4. * H;
}

double nearest_image(double d, double period)
{
  if (period > 0)
    d -= period * std::floor(d / period + 0.5);
  return d;
}

}  // namespace

extern "C" {

int oc_icbc_vortex(const double prm[10], double t, int64_t n, const double *xy, double *out)
{
  if (!prm || (n > 0 && (!xy || !out)))
    return OC_ERR_ARGUMENT;
  const double h0 = prm[0], uinf = prm[1], vinf = prm[2], amp = prm[3], r0 = prm[4];
  const double xc = prm[5] + uinf * t, yc = prm[6] + vinf * t, g = prm[7];
  const double px = prm[8], py = prm[9];
  const double hmin = h0 - amp;
  const double corr = 4., twoPowp = 4.;
  const double Gamma = M_PI / (twoPowp * r0) *
                       std::sqrt(1. / corr * g * (h0 - hmin) / (0.034179687 * M_PI * M_PI - 0.222222222));
  const double H_pi_half = vortex_H(0.5 * M_PI);
  const double inv_gpi = 1. / (g * M_PI * M_PI);
  const double num = twoPowp * Gamma * r0;
  for (int64_t i = 0; i < n; ++i)
    {
      const double dx = nearest_image(xy[2 * i] - xc, px);
      const double dy = nearest_image(xy[2 * i + 1] - yc, py);
      const double radius = std::sqrt(dx * dx + dy * dy);
      double depth = h0, u = uinf, v = vinf;
      if (radius < r0)
        {
          const double rho_half = 0.5 * M_PI * radius / r0;
          const double c = std::cos(rho_half);
          const double omega = twoPowp * Gamma * (c * c) * (c * c);
          depth = h0 - inv_gpi * num * num * (H_pi_half - vortex_H(rho_half));
          u = uinf - dy * omega;
          v = vinf + dx * omega;
        }
      out[3 * i + 0] = depth;
      out[3 * i + 1] = u;
      out[3 * i + 2] = v;
    }
  return OC_OK;
}

int oc_icbc_channel(const double prm[8], uint64_t seed, int64_t n, const double *xy,
                    double *out)
{
  if (!prm || (n > 0 && (!xy || !out)))
    return OC_ERR_ARGUMENT;
  std::mt19937_64 engine(seed);
  auto uniform01 = [&]() { return double(engine() >> 11) * (1.0 / 9007199254740992.0); };
  int K = (int)prm[1];
  if (K < 0)
    K = 5 + (int)(engine() % 11);
  std::vector<double> A(K), W(K), X(K);
  for (int k = 0; k < K; ++k)
    {
      A[k] = prm[2] + (prm[3] - prm[2]) * uniform01();
      W[k] = prm[4] + (prm[5] - prm[4]) * uniform01();
      X[k] = prm[6] + (prm[7] - prm[6]) * uniform01();
    }
  for (int64_t i = 0; i < n; ++i)
    {
      const double x = xy[2 * i];
      double b = 0.0;
      if (x > 8.0 && x < 12.0)
        b = 0.2 - 0.05 * (x - 10.0) * (x - 10.0);
      for (int k = 0; k < K; ++k)
        {
          const double s = (x - X[k]) / W[k];
          b += A[k] * std::exp(-0.5 * s * s);
        }
      out[5 * i + 0] = -b;
      out[5 * i + 1] = prm[0];
      out[5 * i + 2] = 0.0;
      out[5 * i + 3] = 0.0;
      out[5 * i + 4] = 0.0;
    }
  return OC_OK;
}

}  // extern "C"
