// dg_physics.hpp -- pointwise shallow-water model and numerical fluxes (device + host).
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
  __host__ __device__ __forceinline__ T inv_depth(T h) const
  {
    const T h4 = h * h * h * h;
    const T mx = h4 > cu4 ? h4 : cu4;
    return T(1.4142135623730951) * h / std::sqrt(h4 + mx);
  }

  __host__ __device__ __forceinline__ Vec2<T> velocity(T z, Vec2<T> q, T zb) const
  {
    const T s = inv_depth(depth(z, zb));
    return {q.x * s, q.y * s};
  }

  __host__ __device__ __forceinline__ T wave_speed(T z, T zb) const
  {
    return std::sqrt(g * depth(z, zb));
  }

  // LaxFriedrichsModified::numerical_massflux_weak
  __host__ __device__ __forceinline__ T mass_flux(T z_m, T z_p, Vec2<T> q_m, Vec2<T> q_p,
                                                  Vec2<T> n, T zb_m, T zb_p) const
  {
    const Vec2<T> v_m = velocity(z_m, q_m, zb_m);
    const Vec2<T> v_p = velocity(z_p, q_p, zb_p);
    const T lm = std::fabs(v_m.x * n.x + v_m.y * n.y) + wave_speed(z_m, zb_m);
    const T lp = std::fabs(v_p.x * n.x + v_p.y * n.y) + wave_speed(z_p, zb_p);
    const T lambda = lp > lm ? lp : lm;
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hm = depth(z_m, zbs), hp = depth(z_p, zbs);
    const T zcm = z_m > -zbs ? z_m : -zbs;
    const T zcp = z_p > -zbs ? z_p : -zbs;
    return T(0.5) * ((hm * v_m.x * n.x + hm * v_m.y * n.y) + (hp * v_p.x * n.x + hp * v_p.y * n.y)) +
           T(0.5) * lambda * (zcm - zcp);
  }

  // LaxFriedrichsModified::numerical_advflux_weak
  __host__ __device__ __forceinline__ Vec2<T> adv_flux(T z_m, T z_p, Vec2<T> q_m, Vec2<T> q_p,
                                                       Vec2<T> n, T zb_m, T zb_p) const
  {
    const Vec2<T> v_m = velocity(z_m, q_m, zb_m);
    const Vec2<T> v_p = velocity(z_p, q_p, zb_p);
    const T lm = std::fabs(v_m.x * n.x + v_m.y * n.y) + wave_speed(z_m, zb_m);
    const T lp = std::fabs(v_p.x * n.x + v_p.y * n.y) + wave_speed(z_p, zb_p);
    const T lambda = lp > lm ? lp : lm;
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hm = depth(z_m, zbs), hp = depth(z_p, zbs);
    const Vec2<T> hu_m = {hm * v_m.x, hm * v_m.y};
    const Vec2<T> hu_p = {hp * v_p.x, hp * v_p.y};
    // advective_flux(z, hu, zb*): F[e][d] = hu_e * velocity(z, hu, zb*)_d
    const T sm = inv_depth(hm), sp = inv_depth(hp);
    const Vec2<T> w_m = {hu_m.x * sm, hu_m.y * sm};
    const Vec2<T> w_p = {hu_p.x * sp, hu_p.y * sp};
    const T wn_m = w_m.x * n.x + w_m.y * n.y;
    const T wn_p = w_p.x * n.x + w_p.y * n.y;
    const T hl = T(0.5) * lambda;
    return {T(0.5) * (hu_m.x * wn_m + hu_p.x * wn_p) + hl * (hu_m.x - hu_p.x),
            T(0.5) * (hu_m.y * wn_m + hu_p.y * wn_p) + hl * (hu_m.y - hu_p.y)};
  }

  // NumericalFluxBase::numerical_presflux_strong, CENTER variant
  __host__ __device__ __forceinline__ Vec2<T> pres_flux(T z_m, T z_p, Vec2<T> n, T zb_m,
                                                        T zb_p) const
  {
    const T zbs = zb_m < zb_p ? zb_m : zb_p;
    const T hp = depth(z_p, zbs), hm = depth(z_m, zbs);
    const T zcm = z_m > -zbs ? z_m : -zbs;
    const T zcp = z_p > -zbs ? z_p : -zbs;
    const T s = g * T(0.25) * (hp + hm) * (zcp - zcm);
    return {s * n.x, s * n.y};
  }

  __host__ __device__ __forceinline__ T riemann_p(T z, Vec2<T> q, Vec2<T> n, T zb) const
  {
    const Vec2<T> v = velocity(z, q, zb);
    return (v.x * n.x + v.y * n.y) + T(2) * wave_speed(z, zb);
  }
  __host__ __device__ __forceinline__ T riemann_m(T z, Vec2<T> q, Vec2<T> n, T zb) const
  {
    const Vec2<T> v = velocity(z, q, zb);
    return (v.x * n.x + v.y * n.y) - T(2) * wave_speed(z, zb);
  }
};

}  // namespace oceano
