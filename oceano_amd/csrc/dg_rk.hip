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

#include <algorithm>
#include <cmath>
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
  std::vector<double> c;                  // stage s runs at t + c[s] dt (row sums)
  std::vector<double *> kh, kq;           // S + 1 each
  double *rh[2] = {nullptr, nullptr}, *rq[2] = {nullptr, nullptr};
  double t = 0.0;                         // time of the next step
  int *d_tick = nullptr;                  // boundary-table entry of the replayed step
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // a loop over several operators (oc_rk_create_multi): their registers, one launch per
  // stage for all of them (oc_dg_perform_stage_hydro_multi)
  struct Regs
  {
    std::vector<double *> kh, kq;
    double *rh[2] = {nullptr, nullptr}, *rq[2] = {nullptr, nullptr};
  };
  std::vector<oc_dg *> ops;
  std::vector<Regs> regs;
  std::vector<double> ts;                 // multi: the time of each operator's next step
  std::vector<double> graph_dts;          // multi: the steps recorded per operator
  std::vector<uint64_t> graph_generations;
  double graph_dt = 0.0;
  int graph_steps = 1;                    // RK steps recorded in the graph
  bool graph_ticks = false;               // the graph indexes the table by d_tick
  uint64_t graph_generation = 0;
};

using oceano::set_error;

namespace {

__global__ void k_set_tick(int *tick, int v) { *tick = v; }
__global__ void k_add_tick(int *tick, int v) { *tick += v; }

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
  for (auto &r : rk->regs)
    {
      for (double *p : r.kh)
        (void)hipFree(p);
      for (double *p : r.kq)
        (void)hipFree(p);
      for (int i = 0; i < 2; ++i)
        {
          (void)hipFree(r.rh[i]);
          (void)hipFree(r.rq[i]);
        }
    }
  (void)hipFree(rk->d_tick);
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

// the S stages of one step at time rk->t (ExplicitRungeKuttaIntegrator.h:212-271: stage s
// at current_time + c_s time_step)
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
      // the intermediate stage inputs live only in the integrator's r registers: the FP32
      // paired stage may pass them in their centred FP32 form (oc_dg::rk_in32 / rk_out32)
      rk->op->rk_in32 = st > 0;
      rk->op->rk_out32 = st + 1 < S;
      const int rc = oc_dg_perform_stage_hydro(
        rk->op, (uint32_t)st, rk->t + rk->c[st] * dt, f.data(), (uint32_t)st + 1, ri_h, ri_q,
        velocity, rk->kh.data(), rk->kq.data(), (uint32_t)S + 1, w_h, w_q,
        rk->rh[(st + 1) % 2], rk->rq[(st + 1) % 2], 0, s);
      rk->op->rk_in32 = rk->op->rk_out32 = 0;
      if (rc)
        return rc;
    }
  if (velocity)
    return oc_dg_velocity(rk->op, w_h, w_q, velocity, s);
  return OC_OK;
}

// the S stages of one step of every operator of a multi-operator loop, one launch per stage
int step_multi(oc_rk *rk, const double *dt, double *const *w_h, double *const *w_q,
               hipStream_t s)
{
  const int S = rk->S, n = (int)rk->ops.size();
  std::vector<const double *> ri_h(n), ri_q(n), fp(n);
  std::vector<double *const *> kh(n), kq(n);
  std::vector<double *> rh(n), rq(n);
  std::vector<std::vector<double>> f(n);
  std::vector<double> t(n);
  for (int st = 0; st < S; ++st)
    {
      for (int g = 0; g < n; ++g)
        {
          f[g] = rk->rows[st];
          for (double &v : f[g])
            v *= dt[g];
          fp[g] = f[g].data();
          t[g] = rk->ts[g] + rk->c[st] * dt[g];
        }
      for (int g = 0; g < n; ++g)
        {
          oc_rk::Regs &r = rk->regs[g];
          ri_h[g] = st == 0 ? w_h[g] : r.rh[st % 2];
          ri_q[g] = st == 0 ? w_q[g] : r.rq[st % 2];
          kh[g] = r.kh.data();
          kq[g] = r.kq.data();
          rh[g] = r.rh[(st + 1) % 2];
          rq[g] = r.rq[(st + 1) % 2];
        }
      const int rc = oc_dg_perform_stage_hydro_multi(
        n, rk->ops.data(), (uint32_t)st, t.data(), fp.data(), (uint32_t)st + 1, ri_h.data(),
        ri_q.data(), kh.data(), kq.data(), (uint32_t)S + 1, w_h, w_q, rh.data(), rq.data(), s);
      if (rc)
        return rc;
    }
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
  rk->c.assign(1, 0.0);
  for (int s = 0; s < n_stages; ++s)
    {
      rk->rows.emplace_back(rows + o, rows + o + s + 1);
      if (s + 1 < n_stages)
        {
          double cs = 0.0;
          for (int j = 0; j <= s; ++j)
            cs += rows[o + j];
          rk->c.push_back(cs);
        }
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
  if (!rc && hipMalloc((void **)&rk->d_tick, sizeof(int)) != hipSuccess)
    {
      set_error("oc_rk_create: device allocation failed");
      rc = OC_ERR_HIP;
    }
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
  const int rc = step(rk, dt, w_h, w_q, velocity, (hipStream_t)stream);
  if (!rc)
    rk->t += dt;
  return rc;
}

int oc_rk_set_time(oc_rk *rk, double t)
{
  if (!rk || !std::isfinite(t))
    {
      set_error("oc_rk_set_time: null handle or non-finite time");
      return OC_ERR_ARGUMENT;
    }
  rk->t = t;
  for (double &v : rk->ts)
    v = t;
  return OC_OK;
}

int oc_rk_time(const oc_rk *rk, double *t)
{
  if (!rk || !t)
    {
      set_error("oc_rk_time: null argument");
      return OC_ERR_ARGUMENT;
    }
  *t = rk->t;
  return OC_OK;
}

int oc_rk_capture(oc_rk *rk, double dt, double *w_h, double *w_q, double *velocity,
                  void *stream)
{
  return oc_rk_capture_steps(rk, 1, dt, w_h, w_q, velocity, stream);
}

int oc_rk_capture_steps(oc_rk *rk, int32_t n_steps, double dt, double *w_h, double *w_q,
                        double *velocity, void *stream)
{
  if (!rk || !w_h || !w_q || !stream || n_steps < 1)
    {
      set_error("oc_rk_capture: null argument (capture needs a non-null stream) or no step");
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
  oc_dg *op = rk->op;
  const int timing = op->timing;
  op->timing = 0;
  // time-dependent boundary data: the recorded stages index the table by the device tick,
  // which oc_rk_replay sets to the entry of the first replayed stage and the graph advances
  // by S at the end of every step
  const bool ticks = op->needs_bdata && !op->bt_any;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    {
      op->timing = timing;
      set_error("oc_rk_capture: hipStreamBeginCapture failed");
      return OC_ERR_HIP;
    }
  if (ticks)
    op->capture_tick = rk->d_tick;
  int rc = OC_OK;
  for (int i = 0; i < n_steps && !rc; ++i)
    {
      rc = step(rk, dt, w_h, w_q, velocity, s);
      if (!rc && ticks)
        {
          hipLaunchKernelGGL(k_add_tick, dim3(1), dim3(1), 0, s, rk->d_tick, rk->S);
          if (hipGetLastError() != hipSuccess)
            {
              set_error("oc_rk_capture: tick kernel launch failed");
              rc = OC_ERR_HIP;
            }
        }
    }
  op->capture_tick = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &rk->graph);
  op->timing = timing;
  if (rc)
    return rc;
  if (e != hipSuccess ||
      hipGraphInstantiate(&rk->exec, rk->graph, nullptr, nullptr, 0) != hipSuccess)
    {
      set_error("oc_rk_capture: graph capture or instantiation failed");
      return OC_ERR_HIP;
    }
  rk->graph_dt = dt;
  rk->graph_steps = n_steps;
  rk->graph_ticks = ticks;
  rk->graph_generation = op->generation;
  return OC_OK;
}

int oc_rk_replay(oc_rk *rk, int32_t n, void *stream)
{
  if (!rk || !rk->exec || n < 0)
    {
      set_error("oc_rk_replay: no captured step");
      return OC_ERR_ARGUMENT;
    }
  oc_dg *op = rk->op;
  bool stale = op->generation != rk->graph_generation;
  for (size_t g = 0; g < rk->graph_generations.size(); ++g)
    stale = stale || rk->ops[g]->generation != rk->graph_generations[g];
  if (stale)
    {
      set_error("oc_rk_replay: the operator's static data, boundary data or closure changed "
                "since oc_rk_capture; capture again");
      return OC_ERR_ARGUMENT;
    }
  hipStream_t s = (hipStream_t)stream;
  const long steps = (long)n * rk->graph_steps;
  if (rk->graph_ticks && n > 0)
    {
      // the steps read consecutive table entries from the first entry at which the stage
      // times of a whole step start (a step's last stage time usually repeats as the next
      // step's first, so the entry of the time alone can be the previous step's last)
      const double dt = rk->graph_dt;
      const size_t nt = op->bt_times.size(), S = (size_t)rk->S;
      auto matches = [&](size_t e, long r, size_t st) {
        const double t = rk->t + r * dt + rk->c[st] * dt;
        return std::fabs(op->bt_times[e] - t) <= 1e-12 * std::max(1.0, std::fabs(t));
      };
      long e0 = -1;
      for (size_t k = 0; k + S <= nt && e0 < 0; ++k)
        {
          bool all = true;
          for (size_t st = 0; st < S && all; ++st)
            all = matches(k + st, 0, st);
          if (all)
            e0 = (long)k;
        }
      bool ok = e0 >= 0 && (size_t)e0 + (size_t)steps * S <= nt;
      for (long r = 0; ok && r < steps; ++r)
        for (size_t st = 0; ok && st < S; ++st)
          ok = matches((size_t)e0 + (size_t)r * S + st, r, st);
      if (!ok)
        {
          set_error("oc_rk_replay: the boundary table does not list the stage times of the "
                    "replayed steps consecutively from the current time (entries "
                    "t + r dt + c_s dt, r = 0..n-1, s = 0..S-1)");
          return OC_ERR_ARGUMENT;
        }
      hipLaunchKernelGGL(k_set_tick, dim3(1), dim3(1), 0, s, rk->d_tick, (int)e0);
      if (hipGetLastError() != hipSuccess)
        {
          set_error("oc_rk_replay: tick kernel launch failed");
          return OC_ERR_HIP;
        }
    }
  for (int i = 0; i < n; ++i)
    if (hipGraphLaunch(rk->exec, s) != hipSuccess)
      {
        set_error("oc_rk_replay: hipGraphLaunch failed");
        return OC_ERR_HIP;
      }
  rk->t += steps * rk->graph_dt;
  for (size_t g = 0; g < rk->ts.size(); ++g)
    rk->ts[g] += steps * rk->graph_dts[g];
  return OC_OK;
}

int oc_rk_create_multi(int32_t n_ops, oc_dg *const *ops, int32_t n_stages, const double *rows,
                       oc_rk **out)
{
  if (!ops || !out || n_ops < 1 || n_ops > OC_MULTI_MAX_OPERATORS)
    {
      set_error("oc_rk_create_multi: null argument or operator count outside 1.." +
                std::to_string(OC_MULTI_MAX_OPERATORS));
      return OC_ERR_ARGUMENT;
    }
  for (int g = 0; g < n_ops; ++g)
    if (!ops[g])
      {
        set_error("oc_rk_create_multi: null operator");
        return OC_ERR_ARGUMENT;
      }
  oc_rk *rk = nullptr;
  // the first operator's registers are the single-operator ones (unused by multi steps)
  if (const int rc = oc_rk_create(ops[0], n_stages, rows, &rk))
    return rc;
  rk->ops.assign(ops, ops + n_ops);
  rk->regs.resize(n_ops);
  rk->ts.assign(n_ops, 0.0);
  int rc = OC_OK;
  for (int g = 0; g < n_ops && !rc; ++g)
    {
      oc_rk::Regs &r = rk->regs[g];
      const size_t n1 = (size_t)ops[g]->NB * oceano::CB * ops[g]->nd, n2 = 2 * n1;
      r.kh.assign(n_stages + 1, nullptr);
      r.kq.assign(n_stages + 1, nullptr);
      for (int j = 0; j <= n_stages && !rc; ++j)
        rc = zeros(&r.kh[j], n1) ? OC_ERR_HIP : zeros(&r.kq[j], n2);
      for (int i = 0; i < 2 && !rc; ++i)
        rc = zeros(&r.rh[i], n1) ? OC_ERR_HIP : zeros(&r.rq[i], n2);
    }
  if (rc)
    {
      oc_rk_destroy(rk);
      return rc;
    }
  *out = rk;
  return OC_OK;
}

int oc_rk_step_multi(oc_rk *rk, const double *dt, double *const *w_h, double *const *w_q,
                     void *stream)
{
  if (!rk || rk->ops.empty() || !dt || !w_h || !w_q)
    {
      set_error("oc_rk_step_multi: null argument or not a multi-operator loop");
      return OC_ERR_ARGUMENT;
    }
  const int rc = step_multi(rk, dt, w_h, w_q, (hipStream_t)stream);
  if (!rc)
    for (size_t g = 0; g < rk->ts.size(); ++g)
      rk->ts[g] += dt[g];
  return rc;
}

int oc_rk_capture_multi(oc_rk *rk, int32_t n_steps, const double *dt, double *const *w_h,
                        double *const *w_q, void *stream)
{
  if (!rk || rk->ops.empty() || !dt || !w_h || !w_q || !stream || n_steps < 1)
    {
      set_error("oc_rk_capture_multi: null argument, null stream, no step or not a "
                "multi-operator loop");
      return OC_ERR_ARGUMENT;
    }
  for (oc_dg *op : rk->ops)
    if (op->needs_bdata && !op->bt_any)
      {
        set_error("oc_rk_capture_multi: time-dependent boundary data are not supported in a "
                  "multi-operator graph");
        return OC_ERR_UNSUPPORTED;
      }
  hipStream_t s = (hipStream_t)stream;
  if (rk->exec)
    (void)hipGraphExecDestroy(rk->exec);
  if (rk->graph)
    (void)hipGraphDestroy(rk->graph);
  rk->exec = nullptr;
  rk->graph = nullptr;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess)
    {
      set_error("oc_rk_capture_multi: hipStreamBeginCapture failed");
      return OC_ERR_HIP;
    }
  int rc = OC_OK;
  for (int i = 0; i < n_steps && !rc; ++i)
    rc = step_multi(rk, dt, w_h, w_q, s);
  const hipError_t e = hipStreamEndCapture(s, &rk->graph);
  if (rc)
    return rc;
  if (e != hipSuccess ||
      hipGraphInstantiate(&rk->exec, rk->graph, nullptr, nullptr, 0) != hipSuccess)
    {
      set_error("oc_rk_capture_multi: graph capture or instantiation failed");
      return OC_ERR_HIP;
    }
  rk->graph_dt = dt[0];
  rk->graph_dts.assign(dt, dt + rk->ops.size());
  rk->graph_steps = n_steps;
  rk->graph_ticks = false;
  rk->graph_generation = rk->op->generation;
  rk->graph_generations.clear();
  for (oc_dg *op : rk->ops)
    rk->graph_generations.push_back(op->generation);
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
