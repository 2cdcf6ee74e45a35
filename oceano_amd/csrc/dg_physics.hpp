// dg_physics.hpp -- pointwise shallow-water model and numerical fluxes (device code).
//
// Restates, for a scalar number type T, the inline functions the reference evaluates at
// every quadrature point:
//   ShallowWater.h:315-336  depth (floor 1e-12) and desingularised velocity
//   ShallowWater.h:381-432  advective/diffusive flux and source (Manning friction,
//                           general wind stress, Coriolis; BottomFriction.h:186-199,
//                           WindStress.h:101-108, Coriolis.h:87-95)
//   ShallowWater.h:476-516  Riemann invariants (absorbing outflow)
//   LaxFriedrichsModified.h:161-230  LLF mass / advective flux, hydrostatic reconstruction
//   NumericalFluxBase.h:103-126      centred strong-form pressure jump
// The faces of the configured meshes are axis-aligned, so the normal is one of +-e_x,
// +-e_y; the functions still take a general normal.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace oceano {

template <typename T>
struct Vec2
{
  T x, y;
};

// 1/sqrt(x) for x > 0 (the depth floor keeps every argument positive): the hardware
// reciprocal square root refined by two Newton-Raphson steps y <- y + y (1 - x y^2) / 2,
// within a few ulp of the correctly rounded 1/sqrt(x).  FP64: this replaces a correctly
// rounded sqrt and a division (two multi-instruction sequences) by ~10 instructions.
__device__ __forceinline__ double rsqrt_nr(double x)
{
  double y = __builtin_amdgcn_rsq(x);
  double r = __builtin_fma(-(x * y), y, 1.0);
  y = __builtin_fma(0.5 * y, r, y);
  r = __builtin_fma(-(x * y), y, 1.0);
  return __builtin_fma(0.5 * y, r, y);
}
// FP32 (mixed path): the hardware reciprocal square root (within 1 ulp), unrefined
__device__ __forceinline__ float rsqrt_nr(float x) { return __builtin_amdgcn_rsqf(x); }

// a . n for the axis-aligned normals of the kernels: with one component of n zero the
// other product is the whole dot product (a.x * 1 + a.y * 0 == a.x for finite a); the
// tests fold at compile time because every normal is a constant of an unrolled face loop
template <typename T>
__host__ __device__ __forceinline__ T dot_n(Vec2<T> a, Vec2<T> n)
{
  if (n.y == T(0))
    return a.x * n.x;
  if (n.x == T(0))
    return a.y * n.y;
  return a.x * n.x + a.y * n.y;
}

template <typename T>
struct Phys
{
  T g;
  T cu4;  // Threshold_for_wetdry^4 (ShallowWater.h:295-296)
  T nu;   // constant horizontal viscosity

  __host__ __device__ __forceinline__ T depth(T z, T zb) const
  {
    const T h = z + zb;
    return h > T(1e-12) ? h : T(1e-12);
  }

  // inverse_depth = sqrt(2) h / sqrt(h^4 + max(h^4, cu^4))
  __device__ __forceinline__ T inv_depth(T h) const
  {
    const T h4 = h * h * h * h;
    const T mx = h4 > cu4 ? h4 : cu4;
    return (T(1.4142135623730951) * h) * rsqrt_nr(h4 + mx);
  }

  __device__ __forceinline__ Vec2<T> velocity(T z, Vec2<T> q, T zb) const
  {
    const T s = inv_depth(depth(z, zb));
    return {q.x * s, q.y * s};
  }

  // sqrt(g h) = (g h) / sqrt(g h); g h >= g 1e-12 > 0
  __device__ __forceinline__ T wave_speed(T z, T zb) const
  {
    const T gh = g * depth(z, zb);
    return gh * rsqrt_nr(gh);
  }

  // LaxFriedrichsModified::numerical_massflux_weak
  __device__ __forceinline__ T mass_flux(T z_m, T z_p, Vec2<T> q_m, Vec2<T> q_p, Vec2<T> n,
                                        T zb_m, T zb_p) const
  {
    const Vec2<T> v_m = velocity(z_m, q_m, zb_m);
    const Vec2<T> v_p = velocity(z_p, q_p, zb_p);
    const T lm = std::fabs(dot_n(v_m, n)) + wave_speed(z_m, zb_m);
    const T lp = std::fabs(dot_n(v_p, n)) + wave_speed(z_p, zb_p);
    const T lambda = lp > lm ? lp : lm;
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hm = depth(z_m, zbs), hp = depth(z_p, zbs);
    const T zcm = z_m > -zbs ? z_m : -zbs;
    const T zcp = z_p > -zbs ? z_p : -zbs;
    return T(0.5) * (dot_n(Vec2<T>{hm * v_m.x, hm * v_m.y}, n) +
                     dot_n(Vec2<T>{hp * v_p.x, hp * v_p.y}, n)) +
           T(0.5) * lambda * (zcm - zcp);
  }

  // LaxFriedrichsModified::numerical_advflux_weak
  __device__ __forceinline__ Vec2<T> adv_flux(T z_m, T z_p, Vec2<T> q_m, Vec2<T> q_p,
                                             Vec2<T> n, T zb_m, T zb_p) const
  {
    const T s_m = inv_depth(depth(z_m, zb_m)), s_p = inv_depth(depth(z_p, zb_p));
    const Vec2<T> v_m = {q_m.x * s_m, q_m.y * s_m};
    const Vec2<T> v_p = {q_p.x * s_p, q_p.y * s_p};
    const T lm = std::fabs(dot_n(v_m, n)) + wave_speed(z_m, zb_m);
    const T lp = std::fabs(dot_n(v_p, n)) + wave_speed(z_p, zb_p);
    const T lambda = lp > lm ? lp : lm;
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hm = depth(z_m, zbs), hp = depth(z_p, zbs);
    const Vec2<T> hu_m = {hm * v_m.x, hm * v_m.y};
    const Vec2<T> hu_p = {hp * v_p.x, hp * v_p.y};
    // advective_flux(z, hu, zb*): F[e][d] = hu_e * velocity(z, hu, zb*)_d; with a
    // continuous bathymetry zb* = zb_m = zb_p and the inverse depths are the ones above
    T sm = s_m, sp = s_p;
    if (zb_m != zb_p)
      {
        sm = inv_depth(hm);
        sp = inv_depth(hp);
      }
    const Vec2<T> w_m = {hu_m.x * sm, hu_m.y * sm};
    const Vec2<T> w_p = {hu_p.x * sp, hu_p.y * sp};
    const T wn_m = dot_n(w_m, n);
    const T wn_p = dot_n(w_p, n);
    const T hl = T(0.5) * lambda;
    return {T(0.5) * (hu_m.x * wn_m + hu_p.x * wn_p) + hl * (hu_m.x - hu_p.x),
            T(0.5) * (hu_m.y * wn_m + hu_p.y * wn_p) + hl * (hu_m.y - hu_p.y)};
  }

  // NumericalFluxBase::numerical_presflux_strong, CENTER variant
  __device__ __forceinline__ Vec2<T> pres_flux(T z_m, T z_p, Vec2<T> n, T zb_m, T zb_p) const
  {
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hp = depth(z_p, zbs), hm = depth(z_m, zbs);
    const T zcm = z_m > -zbs ? z_m : -zbs;
    const T zcp = z_p > -zbs ? z_p : -zbs;
    const T s = g * T(0.25) * (hp + hm) * (zcp - zcm);
    // s * n, with the zero component of an axis-aligned normal left zero
    return {n.x == T(0) ? T(0) : s * n.x, n.y == T(0) ? T(0) : s * n.y};
  }

  __device__ __forceinline__ T riemann_p(T z, Vec2<T> q, Vec2<T> n, T zb) const
  {
    const Vec2<T> v = velocity(z, q, zb);
    return dot_n(v, n) + T(2) * wave_speed(z, zb);
  }
  __device__ __forceinline__ T riemann_m(T z, Vec2<T> q, Vec2<T> n, T zb) const
  {
    const Vec2<T> v = velocity(z, q, zb);
    return dot_n(v, n) - T(2) * wave_speed(z, zb);
  }
};

}  // namespace oceano
