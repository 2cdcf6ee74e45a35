"""Throughput of the per-stage DG right-hand side on one MI355X.

Workload (the configuration the headline metric is quoted on): travelling vortex at DG
degree 3 on the seeded non-conforming mesh -- [0,2]x[0,1], 20x10 coarse cells refined
globally 4 times, then 30% of the cells (std::mt19937_64 seed 20261202) refined once more:
97,280 cells with 85,944 hanging subfaces, 4.67 M scalar DoF (zeta, hu, hv), periodic.
A "step" is one explicit RK stage (oc_dg_perform_stage_hydro) of classic RK4 over the
whole mesh; inputs are resident in HBM when timing starts.  The metric counts
DoF x stages / second.  The timed region runs the device-resident time loop: one RK4 step
captured into a HIP graph and replayed (plus eager stages for a remainder), exactly
--steps stages between two synchronisations.

Extra fields: `roofline` for the kernel that takes most of the stage time (algorithmic
HBM bytes per launch / mean launch time from HIP events recorded on the launch stream
around each kernel in a second pass over the same number of stages launched eagerly,
against 8 TB/s; `eager_value` is that pass's rate); `cpu_baseline` times the CPU
restatement of the same stage (single core) on a bounded sample of the same workload.
"""
import argparse
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "vortex_p3_hanging_97k"
# HBM bytes per launch of the stage kernel on this workload, measured with rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE scaled by a 512 MiB copy (profiles/k_stage_vortex_p3_hanging_97k.md)
MEASURED_TRAFFIC = {"fp64": 259.9e6, "mixed": 260.3e6}


def build_problem():
    from oceano_amd import problems
    return problems.vortex(degree=3, refine=4, refine_fraction=0.3, seed=20261202)


def algorithmic_bytes(op, stage, static):
    """Compulsory HBM bytes of one stage (FP64 vectors, int32 topology), per kernel.

    stage (one pass): reads zeta_r, q_r, k_h[1..stage], k_q[1..stage], w_h, w_q; writes
    k_h[stage+1], k_q[stage+1], r_h, r_q (w_h, w_q at the last stage); plus per cell the
    face tables (4 kinds + 8 neighbours + 8 face ids, int32), the cell sizes (2 doubles)
    and the stored static fields (none for the vortex: every problem-data field is
    uniform and elided).  Neighbour traces are re-reads of zeta_r, q_r, not compulsory."""
    n, nd = op.n_cells, op.dofs_per_cell
    v1, v2 = 8 * nd * n, 16 * nd * n
    topo = (4 + 8 + 8) * 4 * n
    geom = 16 * n
    stat = sum(static.values())
    one_pass = (v1 + v2) * (1 + stage + 1 + 2) + topo + geom + stat
    rhs = v1 + v2 + v1 + topo + geom
    dis = v1 + v2 + v2 + stage * v2 + 2 * v2 + topo + geom
    jac = (stage + 1) * v1 + 2 * v1 + geom
    return {"stage": one_pass, "height_rhs": rhs, "discharge": dis,
            "rk_update": (stage + 3) * v2, "height_jacobi": jac}


def cpu_baseline(problem, n_stages):
    """Single-core CPU restatement (oracle/) on the first n_stages RK4 stages."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from oceano_amd.integrator import butcher_rows
    o = pyoracle.Oracle.from_mesh(problem.mesh, problem.degree, bcs=problem.bcs)
    d = o.initialize_data_at_quadrature(problem.problem_data)
    h, q = o.project_hydro(problem.initial, d["cell_2"][:, 0])
    rows, _ = butcher_rows("stage_4_order_4", 1e-5)
    S = len(rows)
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(S + 1)]
    w_h, w_q, r_h, r_q = h.copy(), q.copy(), o.zeros(1), o.zeros(2)
    t0 = time.perf_counter()
    for k in range(n_stages):
        s = k % S
        ri_h, ri_q = (w_h.copy(), w_q.copy()) if s == 0 else (r_h.copy(), r_q.copy())
        o.perform_stage_hydro(s, rows[s][:s + 1], ri_h, ri_q, None, ki_h, ki_q, w_h, w_q, r_h,
                              r_q)
    dt = time.perf_counter() - t0
    dofs = problem.dofs()
    return {"value": dofs * n_stages / dt, "unit": "DoF-RK-stages/s", "cores": 1,
            "kind": "port",
            "sample": "%d RK4 stages of the same workload (%d cells, %d DoF) from the "
                      "projected initial state, %.1f s" % (n_stages, problem.mesh.n_cells,
                                                          dofs, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed RK stages")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dt", type=float, default=2e-5,
                    help="RK4 step (CFL ~0.03 on the finest cells; 6.5e-5 is unstable)")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--cpu-stages", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--slot-chunk", type=int, default=None,
                    help="library option slot_chunk (cell grouping by face signature)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run a plain device copy of this size 3 times (a kernel of "
                         "known HBM traffic, to scale FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from oceano_amd import problems
    from oceano_amd.integrator import DeviceIntegrator, butcher_rows

    if args.slot_chunk is not None:
        from oceano_amd import _capi
        _capi.check(_capi.lib().oc_set_option(b"slot_chunk", args.slot_chunk), "oc_set_option")
    P = build_problem()
    rows, _ = butcher_rows("stage_4_order_4", args.dt)
    S = len(rows)

    def run(precision):
        op, data, w_h, w_q = problems.setup_operator(P, precision=precision)
        rk = DeviceIntegrator(op, "stage_4_order_4")
        ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
        ki_q = [op.initialize_vector(1) for _ in range(S + 1)]
        r_h = [op.initialize_vector(0) for _ in range(2)]
        r_q = [op.initialize_vector(1) for _ in range(2)]

        def stage(k):  # eager stage k of the running step
            s = k % S
            ri = [w_h, w_q] if s == 0 else [r_h[s % 2], r_q[s % 2]]
            op.perform_stage_hydro(s, 0.0, rows[s][:s + 1], ri + [None, None], ki_h, ki_q,
                                   w_h, w_q, r_h[(s + 1) % 2], r_q[(s + 1) % 2])

        n_done = 0  # completed time steps
        for _ in range(max(1, -(-args.warmup // S))):
            rk.step(args.dt, w_h, w_q)
            n_done += 1
        # headline: the device-resident loop -- one captured RK step replayed, plus any
        # remaining stages eagerly, exactly args.steps stages between two synchronisations
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            rk.capture(args.dt, w_h, w_q)
            rk.replay(1)
        side.synchronize()
        n_done += 1
        full, rest = divmod(args.steps, S)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        with torch.cuda.stream(side):
            rk.replay(full)
            for k in range(rest):
                stage(k)
        side.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        n_done += full
        k = rest
        while k % S:
            stage(k)
            k += 1
        n_done += 1 if rest else 0
        # per-kernel durations: the same stages launched eagerly with HIP events around
        # each kernel on the launch stream
        op.set_timing(True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for k in range(args.steps):
            stage(k)
        torch.cuda.synchronize()
        eager = time.perf_counter() - t1
        rep = op.timing_report()
        op.set_timing(False)
        k = args.steps
        while k % S:
            stage(k)
            k += 1
        n_done += k // S
        torch.cuda.synchronize()
        err = op.compute_errors(w_h, w_q, P.exact, n_done * args.dt)
        static = {}  # vortex: all problem-data fields uniform, none stored
        per_stage = [algorithmic_bytes(op, s, static) for s in range(S)]
        kern = {}
        for name, (ms, n) in rep.items():
            if n == 0:
                continue
            b = sum(per_stage[j % S][name] for j in range(args.steps)) / args.steps
            mean_ms = ms / n
            kern[name] = {"mean_ms": mean_ms, "launches": n, "alg_bytes": b,
                          "alg_GBs": b / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else None}
        return {"op": op, "elapsed": elapsed, "eager": eager, "kernels": kern, "errors": err,
                "state": (op.unpack(w_h, 1).cpu().numpy(), op.unpack(w_q, 2).cpu().numpy()),
                "nonfinite": op.count_nonfinite(w_h) + op.count_nonfinite(w_q),
                "sweeps_last_stage": op.last_sweeps(), "t_end": n_done * args.dt}

    if args.copy_mib:
        n = args.copy_mib * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        del src, dst
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    runs = {p: run(p) for p in precisions}
    head = runs["mixed"] if "mixed" in runs else runs["fp64"]
    dofs = P.dofs()
    value = dofs * args.steps / head["elapsed"]
    dom = max(head["kernels"], key=lambda k: head["kernels"][k]["mean_ms"])
    achieved = head["kernels"][dom]["alg_GBs"]
    line = {
        "metric": "DoF-RK-stages/s on 1 GPU, mixed vs FP64 at matched L2 error; % HBM peak",
        "value": value,
        "unit": "DoF-RK-stages/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["elapsed"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if "mixed" in runs else "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "cells": P.mesh.n_cells,
                   "hanging_subfaces": P.mesh.n_hanging, "dofs": dofs, "degree": 3,
                   "scheme": "classic RK4 (Butcher), stage = step", "dt": args.dt,
                   "precision": "mixed" if "mixed" in runs else "fp64"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": MEASURED_TRAFFIC.get("mixed" if "mixed" in runs else "fp64")
                     if dom == "stage" else None},
    }
    for p, r in runs.items():
        line[p] = {"value": dofs * args.steps / r["elapsed"],
                   "ms_per_step": r["elapsed"] / args.steps * 1e3,
                   "eager_value": dofs * args.steps / r["eager"], "kernels": r["kernels"],
                   "l2_error_vs_exact": {k: float(v) for k, v in r["errors"].items()},
                   "t_end": r["t_end"], "nonfinite": r["nonfinite"],
                   "jacobi_sweeps_last_stage": r["sweeps_last_stage"]}
    if len(runs) == 2:
        (fh, fq), (mh, mq) = runs["fp64"]["state"], runs["mixed"]["state"]
        nd = runs["fp64"]["op"].dofs_per_cell

        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        line["mixed_vs_fp64"] = {
            "speedup": line["mixed"]["value"] / line["fp64"]["value"],
            "rel_l2_zeta": rel(mh, fh),
            "rel_l2_hu": rel(mq.reshape(-1, 2, nd)[:, 0], fq.reshape(-1, 2, nd)[:, 0]),
            "rel_l2_hv": rel(mq.reshape(-1, 2, nd)[:, 1], fq.reshape(-1, 2, nd)[:, 1])}
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(P, args.cpu_stages)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
