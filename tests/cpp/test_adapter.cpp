// Drop-in check of include/oceano_operator.hpp: the reference-shaped time loops
// (ExplicitRungeKuttaIntegrator and AdditiveRungeKuttaIntegrator with vec_ri copies,
// OceanoOperator::perform_stage_hydro and its additive overload with factor_matrix) give
// bit-identical states to direct C-ABI stage calls, and C-ABI failures surface as
// exceptions with the status code.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "oceano_operator.hpp"

using oceano::DeviceVector;

static std::vector<double> host(const DeviceVector &v)
{
  std::vector<double> h(v.size());
  if (hipMemcpy(h.data(), v.get_values(), h.size() * sizeof(double), hipMemcpyDeviceToHost) !=
      hipSuccess)
    throw std::runtime_error("copy");
  return h;
}

int main()
{
  const int32_t ids[4] = {1, 2, 0, 0};
  oc_mesh *mesh = nullptr;
  oceano::check(oc_mesh_create_rectangle(0, 2, 0, 1, 20, 10, 1, 1, ids, &mesh), "mesh");
  oceano::check(oc_mesh_refine_global(mesh, 1), "refine");
  oceano::check(oc_mesh_refine_fraction(mesh, 0.3, 4), "refine fraction");
  int64_t info[4];
  oc_mesh_info(mesh, info);
  std::vector<double> geom(4 * info[0]);
  std::vector<int32_t> inner(6 * info[1]), bdry(3 * info[2] + 1);
  oc_mesh_export(mesh, geom.data(), nullptr, inner.data(), bdry.data());
  oc_mesh_desc desc{(int32_t)info[0], geom.data(), (int32_t)info[1], inner.data(),
                    (int32_t)info[2], bdry.data()};
  oc_params prm{};
  prm.degree = 3;
  prm.g = 9.81;
  prm.threshold_wetdry = 0.1;
  prm.viscosity = 0.0;
  prm.max_iteration_height = 15;
  prm.precision = OC_PRECISION_FP64;

  oceano::OceanoOperator<DeviceVector> op(desc, prm);
  op.initialize_data_at_quadrature([](double, double, double *out) {
    for (int c = 0; c < 5; ++c)
      out[c] = 0.0;
  });

  // initial state: the travelling vortex projected at the Gauss(p+1) points
  int64_t np = 0;
  oc_dg_n_points(op.handle(), OC_STORE_CELL_COLLOC, &np);
  std::vector<double> xy(2 * np), w(3 * np);
  oc_dg_quadrature_points(op.handle(), OC_STORE_CELL_COLLOC, xy.data());
  const double vprm[10] = {1.0, 1.0, 0.0, 0.1, 0.35, 0.5, 0.5, 9.81, 2.0, 1.0};
  oc_icbc_vortex(vprm, 0.0, np, xy.data(), w.data());
  std::vector<double> zv(np), qv(2 * np);
  for (int64_t i = 0; i < np; ++i)
    {
      zv[i] = w[3 * i];
      qv[2 * i] = w[3 * i + 1] * w[3 * i];
      qv[2 * i + 1] = w[3 * i + 2] * w[3 * i];
    }
  std::vector<double> zc(np), qc(2 * np);
  oceano::check(oc_dg_project_collocation(op.handle(), 1, zv.data(), zc.data()), "project");
  oceano::check(oc_dg_project_collocation(op.handle(), 2, qv.data(), qc.data()), "project");
  DeviceVector tz(zc.size()), tq(qc.size());
  if (hipMemcpy(tz.get_values(), zc.data(), zc.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(tq.get_values(), qc.data(), qc.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
    throw std::runtime_error("copy");
  DeviceVector h0, q0;
  op.initialize_vector(h0, 0);
  op.initialize_vector(q0, 1);
  oceano::check(oc_dg_pack(op.handle(), 1, tz.get_values(), h0.get_values(), nullptr), "pack");
  oceano::check(oc_dg_pack(op.handle(), 2, tq.get_values(), q0.get_values(), nullptr), "pack");

  // (a) reference-shaped loop through the adapter
  oceano::ExplicitRungeKuttaIntegrator integrator(oceano::ExplicitRungeKuttaIntegrator::stage_4_order_4);
  DeviceVector wh = h0, wq = q0, vel, rh, rq;
  std::vector<DeviceVector> kh, kq;
  integrator.reinit(op, rh, rq, kh, kq);
  const double dt = 2.8e-4;
  for (int n = 0; n < 10; ++n)
    integrator.perform_time_step(op, n * dt, dt, wh, wq, vel, rh, rq, kh, kq);

  // (b) direct C-ABI calls with two alternating r registers
  DeviceVector bh = h0, bq = q0, r2h[2], r2q[2];
  std::vector<DeviceVector> k2h(5), k2q(5);
  for (int i = 0; i < 2; ++i)
    {
      op.initialize_vector(r2h[i], 0);
      op.initialize_vector(r2q[i], 1);
    }
  for (int j = 0; j < 5; ++j)
    {
      op.initialize_vector(k2h[j], 0);
      op.initialize_vector(k2q[j], 1);
    }
  // the integrator scales its tableau by the step (ExplicitRungeKuttaIntegrator.h:202-205)
  const double a[3][3] = {{0.5 * dt}, {0.0 * dt, 0.5 * dt}, {0.0 * dt, 0.0 * dt, 1.0 * dt}};
  const double b[4] = {(1.0 / 6.0) * dt, (1.0 / 3.0) * dt, (1.0 / 3.0) * dt, (1.0 / 6.0) * dt};
  for (int n = 0; n < 10; ++n)
    for (unsigned s = 0; s < 4; ++s)
      {
        double *khp[5], *kqp[5];
        for (int j = 0; j < 5; ++j)
          {
            khp[j] = k2h[j].get_values();
            kqp[j] = k2q[j].get_values();
          }
        const double *rih = s == 0 ? bh.get_values() : r2h[s % 2].get_values();
        const double *riq = s == 0 ? bq.get_values() : r2q[s % 2].get_values();
        oceano::check(oc_dg_perform_stage_hydro(op.handle(), s, 0.0, s == 3 ? b : a[s], s + 1,
                                                rih, riq, nullptr, khp, kqp, 5, bh.get_values(),
                                                bq.get_values(), r2h[(s + 1) % 2].get_values(),
                                                r2q[(s + 1) % 2].get_values(), 0, nullptr),
                      "stage");
      }
  const auto A = host(wh), B = host(bh), Aq = host(wq), Bq = host(bq);
  const bool same = A.size() == B.size() && std::memcmp(A.data(), B.data(), A.size() * 8) == 0 &&
                    std::memcmp(Aq.data(), Bq.data(), Aq.size() * 8) == 0;
  double dmax = 0;
  for (size_t i = 0; i < A.size(); ++i)
    dmax = std::fmax(dmax, std::fabs(A[i] - B[i]));
  std::printf("adapter vs C ABI after 10 RK4 steps: bitwise %s, max|diff| %.3e, cells %lld\n",
              same ? "identical" : "DIFFERENT", dmax, (long long)info[0]);

  // (c) the additive (IMEX) integrator through the adapter vs direct C-ABI calls, with a
  // Manning friction field so that the stiff terms are active
  bool same_ark = false;
  {
    oceano::OceanoOperator<DeviceVector> opf(desc, prm);
    opf.initialize_data_at_quadrature([](double x, double, double *out) {
      for (int c = 0; c < 5; ++c)
        out[c] = 0.0;
      out[1] = 0.05 * (1.0 + 0.5 * std::cos(3.0 * x));
    });
    oceano::AdditiveRungeKuttaIntegrator ark;
    DeviceVector ah = h0, aq = q0, avel, arh, arq;
    std::vector<DeviceVector> akh, akq;
    ark.reinit(opf, arh, arq, akh, akq);
    for (int n = 0; n < 5; ++n)
      ark.perform_time_step(opf, n * dt, dt, ah, aq, avel, arh, arq, akh, akq);
    // direct: two alternating r registers
    const double chi = 2.0 - std::sqrt(2.0);
    const double ea[2][2] = {{chi * dt}, {0.5 * dt, 0.5 * dt}};
    const double ta[2][2] = {{chi / 2.0 * dt},
                             {1.0 / (2.0 * std::sqrt(2.0)) * dt, 1.0 / (2.0 * std::sqrt(2.0)) * dt}};
    const double eb[3] = {(0.5 - chi / 4.0) * dt, (0.5 - chi / 4.0) * dt, chi / 2.0 * dt};
    const double fm[3] = {chi / 2.0 * dt, (1.0 - 1.0 / std::sqrt(2.0)) * dt, 0.0};
    DeviceVector ch = h0, cq = q0, r3h[2], r3q[2];
    std::vector<DeviceVector> k3h(4), k3q(6);
    for (int i = 0; i < 2; ++i)
      {
        opf.initialize_vector(r3h[i], 0);
        opf.initialize_vector(r3q[i], 1);
      }
    for (auto &v : k3h)
      opf.initialize_vector(v, 0);
    for (auto &v : k3q)
      opf.initialize_vector(v, 1);
    for (int n = 0; n < 5; ++n)
      for (unsigned s = 0; s < 3; ++s)
        {
          double *khp[4], *kqp[6];
          for (int j = 0; j < 4; ++j)
            khp[j] = k3h[j].get_values();
          for (int j = 0; j < 6; ++j)
            kqp[j] = k3q[j].get_values();
          const double *rih = s == 0 ? ch.get_values() : r3h[s % 2].get_values();
          const double *riq = s == 0 ? cq.get_values() : r3q[s % 2].get_values();
          oceano::check(oc_dg_perform_stage_hydro_ark(
                          opf.handle(), s, 0.0, s == 2 ? eb : ea[s], s == 2 ? eb : ta[s], s + 1,
                          fm[s], rih, riq, nullptr, khp, 4, kqp, 6, ch.get_values(),
                          cq.get_values(), r3h[(s + 1) % 2].get_values(),
                          r3q[(s + 1) % 2].get_values(), nullptr),
                        "ark stage");
        }
    const auto X = host(ah), Y = host(ch), Xq = host(aq), Yq = host(cq);
    same_ark = std::memcmp(X.data(), Y.data(), X.size() * 8) == 0 &&
               std::memcmp(Xq.data(), Yq.data(), Xq.size() * 8) == 0;
    double d = 0;
    for (size_t i = 0; i < Xq.size(); ++i)
      d = std::fmax(d, std::fabs(Xq[i] - Yq[i]));
    std::printf("ARK adapter vs C ABI after 5 IMEX steps: bitwise %s, max|diff q| %.3e\n",
                same_ark ? "identical" : "DIFFERENT", d);
  }

  // (d) deal.II-ordered (cell-major) vectors: the same loop through an adapter in the
  // cell_major layout gives the unpacked state of (a) bitwise; apply(t, src, dst) likewise
  bool same_cm = false;
  {
    oceano::OceanoOperator<DeviceVector> opc(desc, prm, nullptr, oceano::Layout::cell_major);
    opc.initialize_data_at_quadrature([](double, double, double *out) {
      for (int c = 0; c < 5; ++c)
        out[c] = 0.0;
    });
    DeviceVector ch, cq, cvel, crh, crq;
    opc.initialize_vector(ch, 0);
    opc.initialize_vector(cq, 1);
    if (ch.size() != zc.size() || cq.size() != qc.size())
      throw std::runtime_error("cell-major vector length");
    ch = tz;
    cq = tq;
    std::vector<DeviceVector> ckh, ckq;
    integrator.reinit(opc, crh, crq, ckh, ckq);
    for (int n = 0; n < 10; ++n)
      integrator.perform_time_step(opc, n * dt, dt, ch, cq, cvel, crh, crq, ckh, ckq);
    DeviceVector uh(zc.size()), uq(qc.size());
    oceano::check(oc_dg_unpack(op.handle(), 1, wh.get_values(), uh.get_values(), nullptr),
                  "unpack");
    oceano::check(oc_dg_unpack(op.handle(), 2, wq.get_values(), uq.get_values(), nullptr),
                  "unpack");
    const auto U = host(uh), Uq = host(uq), Cm = host(ch), Cq = host(cq);
    same_cm = std::memcmp(U.data(), Cm.data(), U.size() * 8) == 0 &&
              std::memcmp(Uq.data(), Cq.data(), Uq.size() * 8) == 0;
    // apply(t, src, dst) in both layouts
    std::vector<DeviceVector> src_o{h0, q0}, dst_o(2), src_c{tz, tq}, dst_c(2);
    op.initialize_vector(dst_o[0], 0);
    op.initialize_vector(dst_o[1], 1);
    opc.initialize_vector(dst_c[0], 0);
    opc.initialize_vector(dst_c[1], 1);
    op.apply(0.0, src_o, dst_o);
    opc.apply(0.0, src_c, dst_c);
    DeviceVector ah(zc.size()), aq(qc.size());
    oceano::check(oc_dg_unpack(op.handle(), 1, dst_o[0].get_values(), ah.get_values(), nullptr),
                  "unpack");
    oceano::check(oc_dg_unpack(op.handle(), 2, dst_o[1].get_values(), aq.get_values(), nullptr),
                  "unpack");
    const auto Ah = host(ah), Aq2 = host(aq), Ch = host(dst_c[0]), Cq2 = host(dst_c[1]);
    double amax = 0;
    for (double v : Aq2)
      amax = std::fmax(amax, std::fabs(v));
    same_cm = same_cm && amax > 0 && std::memcmp(Ah.data(), Ch.data(), Ah.size() * 8) == 0 &&
              std::memcmp(Aq2.data(), Cq2.data(), Aq2.size() * 8) == 0;
    std::printf("cell-major layout vs operator order (10 RK4 steps, apply): bitwise %s\n",
                same_cm ? "identical" : "DIFFERENT");
  }

  // errors become exceptions carrying the status code
  int code = 0;
  oc_mesh *m2 = nullptr;
  try
    {
      const int32_t bid[1] = {7};
      const int32_t kind[1] = {OC_BC_WALL};
      const int32_t ids2[4] = {1, 2, 0, 0};
      oceano::check(oc_mesh_create_rectangle(0, 1, 0, 1, 2, 2, 0, 0, ids2, &m2), "mesh");
      int64_t i2[4];
      oc_mesh_info(m2, i2);
      std::vector<double> g2(4 * i2[0]);
      std::vector<int32_t> in2(6 * i2[1] + 1), bd2(3 * i2[2]);
      oc_mesh_export(m2, g2.data(), nullptr, in2.data(), bd2.data());
      oc_mesh_desc d2{(int32_t)i2[0], g2.data(), (int32_t)i2[1], in2.data(), (int32_t)i2[2],
                      bd2.data()};
      oc_params p2 = prm;
      p2.n_boundary_kinds = 1;
      p2.boundary_ids = bid;
      p2.boundary_kinds = kind;
      oceano::OceanoOperator<DeviceVector> bad(d2, p2);
    }
  catch (const oceano::Error &e)
    {
      code = e.code;
      std::printf("expected failure: %s\n", e.what());
    }
  oc_mesh_destroy(m2);
  oc_mesh_destroy(mesh);
  const bool ok = same && same_ark && same_cm && code == OC_ERR_BOUNDARY;
  std::printf("%s\n", ok ? "OK" : "FAILED");
  return ok ? 0 : 1;
}
