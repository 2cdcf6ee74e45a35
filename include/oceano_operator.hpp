// oceano_operator.hpp -- C++ adapter with the reference operator's interface.
//
// Header-only, above the C ABI of liboceano_dg.so (oceano_dg.h).  It keeps the member
// signatures of SpaceDiscretization::OceanoOperator (OceanDG.h:157-380) and the explicit
// Runge-Kutta driver of TimeIntegrator::ExplicitRungeKuttaIntegrator
// (ExplicitRungeKuttaIntegrator.h:77-298) so that the time loop of main.cc:1546-1629
// compiles against it unchanged in structure.  Status codes of the C ABI are turned back
// into exceptions, which is the reference's error convention (main.cc:1709-1734).
//
// VectorType must provide `double *get_values()` / `const double *get_values() const`
// returning device memory of `size()` doubles in the operator's storage order, and
// `reinit(size_t n)`; oceano::DeviceVector below does, and so does deal.II's
// LinearAlgebra::distributed::Vector<double, MemorySpace::Default>.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "oceano_dg.h"

namespace oceano {

class Error : public std::runtime_error
{
public:
  Error(int code, const std::string &where)
    : std::runtime_error(where + " failed (status " + std::to_string(code) +
                         "): " + oc_last_error()),
      code(code)
  {}
  int code;
};

inline void check(int code, const char *where)
{
  if (code != OC_OK)
    throw Error(code, where);
}

// device-resident vector; copies are device-to-device
class DeviceVector
{
public:
  DeviceVector() = default;
  explicit DeviceVector(size_t n) { reinit(n); }
  DeviceVector(const DeviceVector &o) { *this = o; }
  DeviceVector(DeviceVector &&o) noexcept { swap(o); }
  DeviceVector &operator=(const DeviceVector &o)
  {
    if (this != &o)
      {
        reinit(o.n_);
        if (n_)
          check(hipMemcpy(p_, o.p_, n_ * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess
                  ? OC_OK
                  : OC_ERR_HIP,
                "DeviceVector copy");
      }
    return *this;
  }
  DeviceVector &operator=(DeviceVector &&o) noexcept
  {
    swap(o);
    return *this;
  }
  ~DeviceVector() { release(); }
  void reinit(size_t n)
  {
    if (n == n_)
      {
        if (n_)
          (void)hipMemset(p_, 0, n_ * sizeof(double));
        return;
      }
    release();
    n_ = n;
    if (n_)
      {
        if (hipMalloc(&p_, n_ * sizeof(double)) != hipSuccess)
          throw Error(OC_ERR_HIP, "DeviceVector hipMalloc");
        (void)hipMemset(p_, 0, n_ * sizeof(double));
      }
  }
  double *get_values() { return p_; }
  const double *get_values() const { return p_; }
  size_t size() const { return n_; }
  void swap(DeviceVector &o) noexcept
  {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
  }

private:
  void release()
  {
    if (p_)
      (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  double *p_ = nullptr;
  size_t n_ = 0;
};

template <typename VectorType>
class OceanoOperator
{
public:
  OceanoOperator(const oc_mesh_desc &mesh, const oc_params &params, hipStream_t stream = nullptr)
    : stream_(stream)
  {
    check(oc_dg_create(&mesh, &params, &op_), "oc_dg_create");
  }
  OceanoOperator(const OceanoOperator &) = delete;
  OceanoOperator &operator=(const OceanoOperator &) = delete;
  ~OceanoOperator()
  {
    if (op_)
      oc_dg_destroy(op_);
  }

  oc_dg *handle() const { return op_; }

  // OceanDG.h:346-348: variable 0 = free surface, 1 = discharge (and velocity)
  void initialize_vector(VectorType &vector, const unsigned int variable) const
  {
    int64_t n = 0;
    check(oc_dg_vector_length(op_, variable == 0 ? 1 : 2, &n), "oc_dg_vector_length");
    vector.reinit((size_t)n);
  }

  // OceanDG.h:350 + 498-569: problem_data(x, y, out[5]) evaluated at the quadrature points
  template <typename Function>
  void initialize_data_at_quadrature(const Function &problem_data)
  {
    for (int store = OC_STORE_CELL_GAUSS; store <= OC_STORE_BOUNDARY; ++store)
      {
        int64_t n = 0;
        check(oc_dg_n_points(op_, store, &n), "oc_dg_n_points");
        std::vector<double> xy(2 * n), v;
        check(oc_dg_quadrature_points(op_, store, xy.data()), "oc_dg_quadrature_points");
        const int nv = store == OC_STORE_CELL_GAUSS ? 5 : store == OC_STORE_CELL_COLLOC ? 2 : 1;
        v.resize((size_t)n * nv);
        double out[5];
        for (int64_t i = 0; i < n; ++i)
          {
            problem_data(xy[2 * i], xy[2 * i + 1], out);
            for (int c = 0; c < nv; ++c)
              v[(size_t)i * nv + c] = out[c];
          }
        check(oc_dg_set_static_data(op_, store, v.data()), "oc_dg_set_static_data");
      }
  }

  // OceanDG.h:182-194
  void perform_stage_hydro(const unsigned int cur_stage, const double cur_time,
                           const double *factor_residual,
                           const std::vector<VectorType> &current_ri,
                           std::vector<VectorType> &vec_ki_height,
                           std::vector<VectorType> &vec_ki_discharge,
                           VectorType &solution_height, VectorType &solution_discharge,
                           VectorType &next_ri_height, VectorType &next_ri_discharge) const
  {
    const size_t n_ki = vec_ki_height.size();
    if (vec_ki_discharge.size() != n_ki || current_ri.size() < 2)
      throw Error(OC_ERR_ARGUMENT, "perform_stage_hydro");
    std::vector<double *> kh(n_ki), kq(n_ki);
    for (size_t j = 0; j < n_ki; ++j)
      {
        kh[j] = vec_ki_height[j].get_values();
        kq[j] = vec_ki_discharge[j].get_values();
      }
    const double *vel = current_ri.size() > 3 && current_ri.back().size()
                          ? current_ri.back().get_values()
                          : nullptr;
    check(oc_dg_perform_stage_hydro(op_, cur_stage, cur_time, factor_residual, cur_stage + 1,
                                    current_ri[0].get_values(), current_ri[1].get_values(), vel,
                                    kh.data(), kq.data(), (uint32_t)n_ki,
                                    solution_height.get_values(),
                                    solution_discharge.get_values(), next_ri_height.get_values(),
                                    next_ri_discharge.get_values(), 0, stream_),
          "oc_dg_perform_stage_hydro");
  }

  // OceanDG.h:196-208: the TIMEINTEGRATOR_ADDITIVERUNGEKUTTA overload; reads the public
  // member factor_matrix that the integrator sets before every stage
  void perform_stage_hydro(const unsigned int cur_stage, const double cur_time,
                           const double *factor_residual, const double *factor_tilde_residual,
                           const std::vector<VectorType> &current_ri,
                           std::vector<VectorType> &vec_ki_height,
                           std::vector<VectorType> &vec_ki_discharge,
                           VectorType &solution_height, VectorType &solution_discharge,
                           VectorType &next_ri_height, VectorType &next_ri_discharge) const
  {
    const size_t n_h = vec_ki_height.size(), n_q = vec_ki_discharge.size();
    if (current_ri.size() < 2)
      throw Error(OC_ERR_ARGUMENT, "perform_stage_hydro");
    std::vector<double *> kh(n_h), kq(n_q);
    for (size_t j = 0; j < n_h; ++j)
      kh[j] = vec_ki_height[j].get_values();
    for (size_t j = 0; j < n_q; ++j)
      kq[j] = vec_ki_discharge[j].get_values();
    const double *vel = current_ri.size() > 3 && current_ri.back().size()
                          ? current_ri.back().get_values()
                          : nullptr;
    check(oc_dg_perform_stage_hydro_ark(
            op_, cur_stage, cur_time, factor_residual, factor_tilde_residual, cur_stage + 1,
            factor_matrix, current_ri[0].get_values(), current_ri[1].get_values(), vel,
            kh.data(), (uint32_t)n_h, kq.data(), (uint32_t)n_q, solution_height.get_values(),
            solution_discharge.get_values(), next_ri_height.get_values(),
            next_ri_discharge.get_values(), stream_),
          "oc_dg_perform_stage_hydro_ark");
  }

  // d~_stage dt of the implicit stage, set by AdditiveRungeKuttaIntegrator (OceanDG.h:163)
  double factor_matrix = 0.0;

  // OceanDG.h:178-180 with dst = M^-1 L(src) for both fields
  void apply(const double current_time, const VectorType &src_height,
             const VectorType &src_discharge, VectorType &dst_height,
             VectorType &dst_discharge, const VectorType *velocity = nullptr) const
  {
    check(oc_dg_apply(op_, current_time, src_height.get_values(), src_discharge.get_values(),
                      velocity ? velocity->get_values() : nullptr, dst_height.get_values(),
                      dst_discharge.get_values(), stream_),
          "oc_dg_apply");
  }

private:
  oc_dg *op_ = nullptr;
  hipStream_t stream_ = nullptr;
};

// TimeIntegrator::ExplicitRungeKuttaIntegrator (ExplicitRungeKuttaIntegrator.h:77-298) with
// the tableaux of time_stepping.h:137-255; vec_ri copies as in the reference (207-210).
class ExplicitRungeKuttaIntegrator
{
public:
  enum Scheme
  {
    stage_1_order_1,
    stage_2_order_2,
    stage_3_order_3,
    stage_4_order_4,
    stage_3_order_2
  };

  explicit ExplicitRungeKuttaIntegrator(Scheme scheme)
  {
    switch (scheme)
      {
        case stage_1_order_1:
          ai = {{1.0}};
          bi = {1.0};
          ci = {0.0};
          break;
        case stage_2_order_2:
          ai = {{1.0}};
          bi = {0.5, 0.5};
          ci = {0.0, 1.0};
          break;
        case stage_3_order_3:
          ai = {{1.0}, {0.25, 0.25}};
          bi = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
          ci = {0.0, 1.0, 0.5};
          break;
        case stage_4_order_4:
          ai = {{0.5}, {0.0, 0.5}, {0.0, 0.0, 1.0}};
          bi = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
          ci = {0.0, 0.5, 0.5, 1.0};
          break;
        case stage_3_order_2:
          {
            const double chi = 2.0 - std::sqrt(2.0), a32 = 0.5;
            ai = {{chi}, {1.0 - a32, a32}};
            bi = {0.5 - chi / 4.0, 0.5 - chi / 4.0, chi / 2.0};
            ci = {0.0, chi, 1.0};
            break;
          }
      }
  }

  unsigned int n_stages() const { return (unsigned int)bi.size(); }

  template <typename VectorType, typename Operator>
  void reinit(const Operator &op, VectorType &vec_ri_height, VectorType &vec_ri_discharge,
              std::vector<VectorType> &vec_ki_height,
              std::vector<VectorType> &vec_ki_discharge) const
  {
    op.initialize_vector(vec_ri_height, 0);
    op.initialize_vector(vec_ri_discharge, 1);
    vec_ki_height.resize(n_stages() + 1);
    vec_ki_discharge.resize(n_stages() + 1);
    for (auto &v : vec_ki_height)
      op.initialize_vector(v, 0);
    for (auto &v : vec_ki_discharge)
      op.initialize_vector(v, 1);
  }

  template <typename VectorType, typename Operator>
  void perform_time_step(const Operator &op, const double current_time, const double time_step,
                         VectorType &solution_height, VectorType &solution_discharge,
                         VectorType &postprocess_velocity, VectorType &vec_ri_height,
                         VectorType &vec_ri_discharge, std::vector<VectorType> &vec_ki_height,
                         std::vector<VectorType> &vec_ki_discharge) const
  {
    std::vector<double> b_i = bi;
    for (auto &v : b_i)
      v *= time_step;
    std::vector<std::vector<double>> a_i = ai;
    for (auto &row : a_i)
      for (auto &v : row)
        v *= time_step;
    const VectorType empty;
    for (unsigned int stage = 0; stage < ci.size(); ++stage)
      {
        std::vector<VectorType> vec_ri =
          stage == 0 ? std::vector<VectorType>{solution_height, solution_discharge, empty,
                                               postprocess_velocity}
                     : std::vector<VectorType>{vec_ri_height, vec_ri_discharge, empty,
                                               postprocess_velocity};
        op.perform_stage_hydro(stage, current_time + ci[stage] * time_step,
                               stage == ci.size() - 1 ? &b_i[0] : &a_i[stage][0], vec_ri,
                               vec_ki_height, vec_ki_discharge, solution_height,
                               solution_discharge, vec_ri_height, vec_ri_discharge);
      }
  }

private:
  std::vector<std::vector<double>> ai;
  std::vector<double> bi, ci;
};

// TimeIntegrator::AdditiveRungeKuttaIntegrator (AdditiveRungeKuttaIntegrator.h:77-339):
// stage_3_order_2 = explicit THREE_STAGE_SECOND_ORDER (time_stepping.h:223-245) with the
// implicit TRAPEZOIDAL_BDF2 (time_stepping.h:340-379); vec_ri copies as in the reference.
// The reference reads d~[S-1] at the last stage, one past the two coefficients of the
// implicit tableau; the last stage has no implicit solve, so 0 is used there.
class AdditiveRungeKuttaIntegrator
{
public:
  enum Scheme
  {
    stage_3_order_2
  };

  explicit AdditiveRungeKuttaIntegrator(Scheme = stage_3_order_2)
  {
    const double chi = 2.0 - std::sqrt(2.0), a32 = 0.5;
    ai = {{chi}, {1.0 - a32, a32}};
    bi = {0.5 - chi / 4.0, 0.5 - chi / 4.0, chi / 2.0};
    ci = {0.0, chi, 1.0};
    atildei = {{chi / 2.0}, {1.0 / (2.0 * std::sqrt(2.0)), 1.0 / (2.0 * std::sqrt(2.0))}};
    dtildei = {chi / 2.0, 1.0 - 1.0 / std::sqrt(2.0)};
  }

  unsigned int n_stages() const { return (unsigned int)bi.size(); }

  // AdditiveRungeKuttaIntegrator.h:158-182: n_stages+1 height and 2 n_stages discharge k
  template <typename VectorType, typename Operator>
  void reinit(const Operator &op, VectorType &vec_ri_height, VectorType &vec_ri_discharge,
              std::vector<VectorType> &vec_ki_height,
              std::vector<VectorType> &vec_ki_discharge) const
  {
    op.initialize_vector(vec_ri_height, 0);
    op.initialize_vector(vec_ri_discharge, 1);
    vec_ki_height.resize(n_stages() + 1);
    vec_ki_discharge.resize(2 * n_stages());
    for (auto &v : vec_ki_height)
      op.initialize_vector(v, 0);
    for (auto &v : vec_ki_discharge)
      op.initialize_vector(v, 1);
  }

  template <typename VectorType, typename Operator>
  void perform_time_step(Operator &op, const double current_time, const double time_step,
                         VectorType &solution_height, VectorType &solution_discharge,
                         VectorType &postprocess_velocity, VectorType &vec_ri_height,
                         VectorType &vec_ri_discharge, std::vector<VectorType> &vec_ki_height,
                         std::vector<VectorType> &vec_ki_discharge) const
  {
    std::vector<double> b_i = bi;
    for (auto &v : b_i)
      v *= time_step;
    std::vector<std::vector<double>> a_i = ai, at_i = atildei;
    for (auto &row : a_i)
      for (auto &v : row)
        v *= time_step;
    for (auto &row : at_i)
      for (auto &v : row)
        v *= time_step;
    const VectorType empty;
    for (unsigned int stage = 0; stage < ci.size(); ++stage)
      {
        const bool last = stage == ci.size() - 1;
        op.factor_matrix = stage < dtildei.size() ? dtildei[stage] * time_step : 0.0;
        std::vector<VectorType> vec_ri =
          stage == 0 ? std::vector<VectorType>{solution_height, solution_discharge, empty,
                                               postprocess_velocity}
                     : std::vector<VectorType>{vec_ri_height, vec_ri_discharge, empty,
                                               postprocess_velocity};
        op.perform_stage_hydro(stage, current_time + ci[stage] * time_step,
                               last ? &b_i[0] : &a_i[stage][0],
                               last ? &b_i[0] : &at_i[stage][0], vec_ri, vec_ki_height,
                               vec_ki_discharge, solution_height, solution_discharge,
                               vec_ri_height, vec_ri_discharge);
      }
  }

private:
  std::vector<std::vector<double>> ai, atildei;
  std::vector<double> bi, ci, dtildei;
};

}  // namespace oceano
