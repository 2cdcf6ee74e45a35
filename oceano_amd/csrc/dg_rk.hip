// dg_rk.hip -- device-resident explicit Runge-Kutta time loop over oc_dg_perform_stage_hydro.
//
// Restates the stage sequence of TimeIntegrator::ExplicitRungeKuttaIntegrator
// (ExplicitRungeKuttaIntegrator.h:179-298) and the per-step velocity update of the time
// loop (main.cc:1556-1577).  The reference copies the r register into vec_ri before every
// stage (207-210, 255-258); here two r registers alternate, so stage s reads r[s % 2] and
// writes r[(s + 1) % 2] (the results are the same: the stage never reads what it writes).
// A whole step can be captured into a HIP graph once and replayed, which removes the host
// launch overhead from the time loop.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dg_internal.hpp"
#include "oceano_dg.h"

namespace oceano {
void set_error(const std::string &msg);
}

struct oc_rk
{
  oc_dg *op = nullptr;
  int S = 0;
  std::vector<std::vector<double>> rows;  // unscaled: row s feeds stage s's update
  std::vector<double *> kh, kq;           // S + 1 each
  double *rh[2] = {nullptr, nullptr}, *rq[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

using oceano::set_error;

namespace {

void release(oc_rk *rk)
{
  if (rk->exec)
    (void)hipGraphExecDestroy(rk->exec);
  if (rk->graph)
    (void)hipGraphDestroy(rk->graph);
  rk->exec = nullptr;
  rk->graph = nullptr;
  for (double *p : rk->kh)
    (void)hipFree(p);
  for (double *p : rk->kq)
    (void)hipFree(p);
  for (int i = 0; i < 2; ++i)
    {
      (void)hipFree(rk->rh[i]);
      (void)hipFree(rk->rq[i]);
    }
}

int zeros(double **p, size_t n)
{
  if (hipMalloc(p, n * sizeof(double)) != hipSuccess ||
      hipMemset(*p, 0, n * sizeof(double)) != hipSuccess)
    {
      set_error("oc_rk_create: device allocation failed");
      return OC_ERR_HIP;
    }
  return OC_OK;
}

int step(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity, hipStream_t s)
{
  const int S = rk->S;
  for (int st = 0; st < S; ++st)
    {
      std::vector<double> f(rk->rows[st]);
      for (double &v : f)
        v *= dt;
      const double *ri_h = st == 0 ? w_h : rk->rh[st % 2];
      const double *ri_q = st == 0 ? w_q : rk->rq[st % 2];
      const int rc = oc_dg_perform_stage_hydro(
        rk->op, (uint32_t)st, 0.0, f.data(), (uint32_t)st + 1, ri_h, ri_q, velocity,
        rk->kh.data(), rk->kq.data(), (uint32_t)S + 1, w_h, w_q, rk->rh[(st + 1) % 2],
        rk->rq[(st + 1) % 2], 0, s);
      if (rc)
        return rc;
    }
  if (velocity)
    return oc_dg_velocity(rk->op, w_h, w_q, velocity, s);
  return OC_OK;
}

}  // namespace

extern "C" {

int oc_rk_create(oc_dg *op, int32_t n_stages, const double *rows, oc_rk **out)
{
  if (!op || !rows || !out || n_stages < 1 || n_stages > oceano::MAXKI - 1)
    {
      set_error("oc_rk_create: null argument or stage count outside 1..8");
      return OC_ERR_ARGUMENT;
    }
  oc_rk *rk = new oc_rk;
  rk->op = op;
  rk->S = n_stages;
  size_t o = 0;
  for (int s = 0; s < n_stages; ++s)
    {
      rk->rows.emplace_back(rows + o, rows + o + s + 1);
      o += (size_t)s + 1;
    }
  const size_t n1 = (size_t)op->NB * oceano::CB * op->nd, n2 = 2 * n1;
  rk->kh.assign(n_stages + 1, nullptr);
  rk->kq.assign(n_stages + 1, nullptr);
  int rc = OC_OK;
  for (int j = 0; j <= n_stages && !rc; ++j)
    rc = zeros(&rk->kh[j], n1) ? OC_ERR_HIP : zeros(&rk->kq[j], n2);
  for (int i = 0; i < 2 && !rc; ++i)
    rc = zeros(&rk->rh[i], n1) ? OC_ERR_HIP : zeros(&rk->rq[i], n2);
  if (rc)
    {
      release(rk);
      delete rk;
      return rc;
    }
  *out = rk;
  return OC_OK;
}

void oc_rk_destroy(oc_rk *rk)
{
  if (!rk)
    return;
  release(rk);
  delete rk;
}

int oc_rk_step(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity, void *stream)
{
  if (!rk || !w_h || !w_q)
    {
      set_error("oc_rk_step: null argument");
      return OC_ERR_ARGUMENT;
    }
  return step(rk, dt, w_h, w_q, velocity, (hipStream_t)stream);
}

int oc_rk_capture(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity,
                  void *stream)
{
  if (!rk || !w_h || !w_q || !stream)
    {
      set_error("oc_rk_capture: null argument (capture needs a non-null stream)");
      return OC_ERR_ARGUMENT;
    }
  hipStream_t s = (hipStream_t)stream;
  if (rk->exec)
    (void)hipGraphExecDestroy(rk->exec);
  if (rk->graph)
    (void)hipGraphDestroy(rk->graph);
  rk->exec = nullptr;
  rk->graph = nullptr;
  // per-kernel event timing is not recorded into the graph
  const int timing = rk->op->timing;
  rk->op->timing = 0;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    {
      rk->op->timing = timing;
      set_error("oc_rk_capture: hipStreamBeginCapture failed");
      return OC_ERR_HIP;
    }
  const int rc = step(rk, dt, w_h, w_q, velocity, s);
  const hipError_t e = hipStreamEndCapture(s, &rk->graph);
  rk->op->timing = timing;
  if (rc)
    return rc;
  if (e != hipSuccess ||
      hipGraphInstantiate(&rk->exec, rk->graph, nullptr, nullptr, 0) != hipSuccess)
    {
      set_error("oc_rk_capture: graph capture or instantiation failed");
      return OC_ERR_HIP;
    }
  return OC_OK;
}

int oc_rk_replay(oc_rk *rk, int32_t n, void *stream)
{
  if (!rk || !rk->exec || n < 0)
    {
      set_error("oc_rk_replay: no captured step");
      return OC_ERR_ARGUMENT;
    }
  for (int i = 0; i < n; ++i)
    if (hipGraphLaunch(rk->exec, (hipStream_t)stream) != hipSuccess)
      {
        set_error("oc_rk_replay: hipGraphLaunch failed");
        return OC_ERR_HIP;
      }
  return OC_OK;
}

int oc_rk_registers(oc_rk *rk, double **ki_h, double **ki_q, double **r_h, double **r_q)
{
  if (!rk)
    return OC_ERR_ARGUMENT;
  for (int j = 0; j <= rk->S; ++j)
    {
      if (ki_h)
        ki_h[j] = rk->kh[j];
      if (ki_q)
        ki_q[j] = rk->kq[j];
    }
  for (int i = 0; i < 2; ++i)
    {
      if (r_h)
        r_h[i] = rk->rh[i];
      if (r_q)
        r_q[i] = rk->rq[i];
    }
  return OC_OK;
}

}  // extern "C"
