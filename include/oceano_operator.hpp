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
// returning device memory of `size()` doubles, and `reinit(size_t n)`; oceano::DeviceVector
// below does, and so does deal.II's LinearAlgebra::distributed::Vector<double,
// MemorySpace::Default>.
//
// Vector layout (the Layout argument of the constructor):
//   operator_order  the caller's vectors are in the operator's storage order (what
//                   initialize_vector sizes them for, as MatrixFree::initialize_dof_vector
//                   does); every call passes them straight to the kernels.
//   cell_major      the caller's vectors are in deal.II's DoF order for a DG FESystem: cell
//                   by cell in the order of the mesh given at construction, component-major
//                   inside a cell, lexicographic Gauss-Lobatto nodes (FE_DGQ numbering;
//                   main.cc:580-583).  The adapter keeps operator-order mirrors of the time
//                   integrator's registers: stage 0 packs the solution into its mirror, the
//                   stages run on the mirrors, the last stage unpacks the new solution back
//                   into the caller's vectors (2 permutations of the state per step).  The
//                   caller's next_ri / vec_ki vectors are then scratch the operator does not
//                   read back; only the solution carries the state between steps, which is
//                   all the reference time loop reads (output, errors, mass check).
//                   apply() packs its inputs and unpacks its outputs on every call.
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

enum class Layout
{
  operator_order,
  cell_major
};

template <typename VectorType>
class OceanoOperator
{
public:
  OceanoOperator(const oc_mesh_desc &mesh, const oc_params &params, hipStream_t stream = nullptr,
                 Layout layout = Layout::operator_order)
    : stream_(stream), layout_(layout)
  {
    check(oc_dg_create(&mesh, &params, &op_), "oc_dg_create");
    int64_t sizes[7];
    check(oc_dg_sizes(op_, sizes), "oc_dg_sizes");
    n_cells_ = sizes[0];
    dofs_per_cell_ = sizes[2];
  }
  OceanoOperator(const OceanoOperator &) = delete;
  OceanoOperator &operator=(const OceanoOperator &) = delete;
  ~OceanoOperator()
  {
    if (op_)
      oc_dg_destroy(op_);
  }

  oc_dg *handle() const { return op_; }
  Layout layout() const { return layout_; }

  // OceanDG.h:346-348: variable 0 = free surface, 1 = discharge (and velocity)
  void initialize_vector(VectorType &vector, const unsigned int variable) const
  {
    vector.reinit(length(variable == 0 ? 1 : 2, layout_));
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
    if (layout_ == Layout::operator_order)
      {
        check(oc_dg_perform_stage_hydro(op_, cur_stage, cur_time, factor_residual,
                                        cur_stage + 1, current_ri[0].get_values(),
                                        current_ri[1].get_values(), vel, kh.data(), kq.data(),
                                        (uint32_t)n_ki, solution_height.get_values(),
                                        solution_discharge.get_values(),
                                        next_ri_height.get_values(),
                                        next_ri_discharge.get_values(), 0, stream_),
              "oc_dg_perform_stage_hydro");
        return;
      }
    // cell_major: the stages run on operator-order mirrors of the integrator's registers
    Mirrors &m = mirrors(n_ki, n_ki);
    if (cur_stage == 0)
      {
        to_operator(1, solution_height.get_values(), m.w_h.get_values());
        to_operator(2, solution_discharge.get_values(), m.w_q.get_values());
      }
    if (vel)
      {
        to_operator(2, vel, m.vel.get_values());
        vel = m.vel.get_values();
      }
    const bool last = cur_stage + 2 == n_ki;
    const double *ri_h = cur_stage == 0 ? m.w_h.get_values() : m.r_h[cur_stage % 2].get_values();
    const double *ri_q = cur_stage == 0 ? m.w_q.get_values() : m.r_q[cur_stage % 2].get_values();
    check(oc_dg_perform_stage_hydro(op_, cur_stage, cur_time, factor_residual, cur_stage + 1,
                                    ri_h, ri_q, vel, m.kh_ptr.data(), m.kq_ptr.data(),
                                    (uint32_t)n_ki, m.w_h.get_values(), m.w_q.get_values(),
                                    m.r_h[(cur_stage + 1) % 2].get_values(),
                                    m.r_q[(cur_stage + 1) % 2].get_values(), 0, stream_),
          "oc_dg_perform_stage_hydro");
    if (last)
      {
        to_cell_major(1, m.w_h.get_values(), solution_height.get_values());
        to_cell_major(2, m.w_q.get_values(), solution_discharge.get_values());
      }
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
    if (layout_ == Layout::operator_order)
      {
        check(oc_dg_perform_stage_hydro_ark(
                op_, cur_stage, cur_time, factor_residual, factor_tilde_residual,
                cur_stage + 1, factor_matrix, current_ri[0].get_values(),
                current_ri[1].get_values(), vel, kh.data(), (uint32_t)n_h, kq.data(),
                (uint32_t)n_q, solution_height.get_values(), solution_discharge.get_values(),
                next_ri_height.get_values(), next_ri_discharge.get_values(), stream_),
              "oc_dg_perform_stage_hydro_ark");
        return;
      }
    Mirrors &m = mirrors(n_h, n_q);
    if (cur_stage == 0)
      {
        to_operator(1, solution_height.get_values(), m.w_h.get_values());
        to_operator(2, solution_discharge.get_values(), m.w_q.get_values());
      }
    if (vel)
      {
        to_operator(2, vel, m.vel.get_values());
        vel = m.vel.get_values();
      }
    const bool last = cur_stage + 2 == n_h;
    const double *ri_h = cur_stage == 0 ? m.w_h.get_values() : m.r_h[cur_stage % 2].get_values();
    const double *ri_q = cur_stage == 0 ? m.w_q.get_values() : m.r_q[cur_stage % 2].get_values();
    check(oc_dg_perform_stage_hydro_ark(
            op_, cur_stage, cur_time, factor_residual, factor_tilde_residual, cur_stage + 1,
            factor_matrix, ri_h, ri_q, vel, m.kh_ptr.data(), (uint32_t)n_h, m.kq_ptr.data(),
            (uint32_t)n_q, m.w_h.get_values(), m.w_q.get_values(),
            m.r_h[(cur_stage + 1) % 2].get_values(), m.r_q[(cur_stage + 1) % 2].get_values(),
            stream_),
          "oc_dg_perform_stage_hydro_ark");
    if (last)
      {
        to_cell_major(1, m.w_h.get_values(), solution_height.get_values());
        to_cell_major(2, m.w_q.get_values(), solution_discharge.get_values());
      }
  }

  // d~_stage dt of the implicit stage, set by AdditiveRungeKuttaIntegrator (OceanDG.h:163)
  double factor_matrix = 0.0;

  // OceanDG.h:178-180, apply(current_time, src, dst): dst = M^-1 L(src) for both fields
  // (step-67 semantics; the reference's body is dead code), with the boundary data at
  // current_time.  The fields travel as the stage's current_ri does: src = {free surface,
  // discharge[, tracer[, velocity]]} (the velocity is read when the viscosity is nonzero),
  // dst = {free surface, discharge}.
  void apply(const double current_time, const std::vector<VectorType> &src,
             std::vector<VectorType> &dst) const
  {
    if (src.size() < 2 || dst.size() < 2)
      throw Error(OC_ERR_ARGUMENT, "apply: src needs {zeta, q, ...}, dst {zeta, q}");
    const VectorType *vel = src.size() > 3 && src.back().size() ? &src.back() : nullptr;
    apply(current_time, src[0], src[1], dst[0], dst[1], vel);
  }

  // the same with the fields as separate arguments
  void apply(const double current_time, const VectorType &src_height,
             const VectorType &src_discharge, VectorType &dst_height,
             VectorType &dst_discharge, const VectorType *velocity = nullptr) const
  {
    if (layout_ == Layout::operator_order)
      {
        check(oc_dg_apply(op_, current_time, src_height.get_values(),
                          src_discharge.get_values(),
                          velocity ? velocity->get_values() : nullptr, dst_height.get_values(),
                          dst_discharge.get_values(), stream_),
              "oc_dg_apply");
        return;
      }
    Mirrors &m = mirrors(2, 2);
    to_operator(1, src_height.get_values(), m.a_h.get_values());
    to_operator(2, src_discharge.get_values(), m.a_q.get_values());
    if (velocity)
      to_operator(2, velocity->get_values(), m.vel.get_values());
    check(oc_dg_apply(op_, current_time, m.a_h.get_values(), m.a_q.get_values(),
                      velocity ? m.vel.get_values() : nullptr, m.b_h.get_values(),
                      m.b_q.get_values(), stream_),
          "oc_dg_apply");
    to_cell_major(1, m.b_h.get_values(), dst_height.get_values());
    to_cell_major(2, m.b_q.get_values(), dst_discharge.get_values());
  }

private:
  // operator-order registers of the cell_major layout
  struct Mirrors
  {
    DeviceVector w_h, w_q, vel, r_h[2], r_q[2], a_h, a_q, b_h, b_q;
    std::vector<DeviceVector> kh, kq;
    std::vector<double *> kh_ptr, kq_ptr;
  };

  size_t length(int n_components, Layout layout) const
  {
    if (layout == Layout::cell_major)
      return (size_t)n_cells_ * n_components * dofs_per_cell_;
    int64_t n = 0;
    check(oc_dg_vector_length(op_, n_components, &n), "oc_dg_vector_length");
    return (size_t)n;
  }

  Mirrors &mirrors(size_t n_kh, size_t n_kq) const
  {
    Mirrors &m = mirrors_;
    if (m.w_h.size() == 0)
      {
        const size_t n1 = length(1, Layout::operator_order);
        const size_t n2 = length(2, Layout::operator_order);
        m.w_h.reinit(n1);
        m.w_q.reinit(n2);
        m.vel.reinit(n2);
        m.a_h.reinit(n1);
        m.a_q.reinit(n2);
        m.b_h.reinit(n1);
        m.b_q.reinit(n2);
        for (int i = 0; i < 2; ++i)
          {
            m.r_h[i].reinit(n1);
            m.r_q[i].reinit(n2);
          }
      }
    while (m.kh.size() < n_kh)
      m.kh.emplace_back(length(1, Layout::operator_order));
    while (m.kq.size() < n_kq)
      m.kq.emplace_back(length(2, Layout::operator_order));
    m.kh_ptr.resize(m.kh.size());
    m.kq_ptr.resize(m.kq.size());
    for (size_t j = 0; j < m.kh.size(); ++j)
      m.kh_ptr[j] = m.kh[j].get_values();
    for (size_t j = 0; j < m.kq.size(); ++j)
      m.kq_ptr[j] = m.kq[j].get_values();
    return m;
  }

  void to_operator(int n_components, const double *cell_major, double *op_order) const
  {
    check(oc_dg_pack(op_, n_components, cell_major, op_order, stream_), "oc_dg_pack");
  }
  void to_cell_major(int n_components, const double *op_order, double *cell_major) const
  {
    check(oc_dg_unpack(op_, n_components, op_order, cell_major, stream_), "oc_dg_unpack");
  }

  oc_dg *op_ = nullptr;
  hipStream_t stream_ = nullptr;
  Layout layout_ = Layout::operator_order;
  int64_t n_cells_ = 0, dofs_per_cell_ = 0;
  mutable Mirrors mirrors_;
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
