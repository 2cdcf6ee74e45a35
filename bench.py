"""Throughput of the per-stage DG right-hand side on one MI355X.

Default workload (the configuration the headline metric is quoted on): travelling vortex
at DG degree 3 on the seeded non-conforming mesh -- [0,2]x[0,1], 20x10 coarse cells refined
globally 4 times, then 30% of the cells (std::mt19937_64 seed 20261202) refined once more:
97,280 cells with 85,944 hanging subfaces, 4.67 M scalar DoF (zeta, hu, hv), periodic.
A "step" is one explicit RK stage (oc_dg_perform_stage_hydro) of classic RK4 over the
whole mesh; inputs are resident in HBM when timing starts.  The metric counts
DoF x stages / second.  The timed region runs the device-resident time loop: the whole RK4
steps of the region captured into one HIP graph (launched once before, untimed) and
replayed once, plus eager stages for a remainder: exactly --steps stages between two
synchronisations, timed by HIP events on the replay stream (`value`; the host clock around
the same region gives `wall_value`).

Other workloads (--workload): vortex_p1_160x80 (12,800 cells, p = 1, cache-resident),
vortex_p1_819k (the same refined to 819,200 cells, 9.8 M DoF), channel_p2_250x10 (2,500
cells, p = 2, discharge inflow / height outflow / walls / Manning / bumpy bed),
channel_p2_640k (the same channel at 4000 x 160 cells, 17.3 M DoF: the bed is stored at
every quadrature point and streamed every stage) and ensemble_64 (64 vortex instances of
1k-50k cells and degrees 1-4, each degree's instances advanced by one operator over the
union of their meshes; 25.6 M DoF).

Extra fields: `roofline` for the kernel that takes most of the stage time (algorithmic
HBM bytes per launch / mean launch time from HIP events recorded on the launch stream
around each kernel in a second pass over the same number of stages launched eagerly,
against 8 TB/s; `eager_value` is that pass's rate); `cpu_baseline` times the CPU
restatement of the same stage (its loops on the OpenMP threads of the host's CPU share) on
a bounded sample of the same workload.  `--ark` adds the additive (IMEX) stage's rate on
the same mesh; `--steady` adds the whole steady-state run (the channel's criterion: one
step changes the state by less than 1e-10 relative L2) per precision.
"""
import argparse
import gc
import json

import numpy as np
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DEFAULT = "vortex_p3_hanging_97k"
WORKLOADS = [DEFAULT, "vortex_p1_160x80", "vortex_p1_819k", "channel_p2_250x10",
             "channel_p2_640k", "ensemble_64"]


def measured_traffic(workload):
    """HBM bytes per launch of the stage kernel, per precision, as the committed profile
    summary of the workload states them (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE scaled by a
    copy kernel of known size): its line 'traffic per launch (bytes): fp64 N mixed M'.
    Returns ({precision: bytes}, path) or ({}, None) when no profile holds the figure."""
    import re
    path = os.path.join(ROOT, "profiles", "k_stage_%s.md" % workload)
    if not os.path.exists(path):
        return {}, None
    for line in open(path):
        m = re.search(r"traffic per launch \(bytes\):\s*fp64\s+([0-9.e+]+)\s+mixed\s+([0-9.e+]+)",
                      line)
        if m:
            return {"fp64": float(m.group(1)), "mixed": float(m.group(2))}, \
                os.path.relpath(path, ROOT)
    return {}, None


def static_bytes(op, data):
    """HBM bytes of the static stores one stage reads (OceanDG.h:498-569 read every stage):
    each stored field once (a field whose values are all equal is a kernel constant and
    costs nothing); z_b of cell_0 per Gauss point (discharge work), z_b of cell_1 per
    Gauss-Lobatto point (depth solve), the face z_b pair per face point, the boundary z_b
    and boundary data per boundary point."""
    def stored(v):
        v = np.asarray(v)
        return v.size > 0 and not np.all(v == v.reshape(-1)[0])
    b = 0
    c0 = np.asarray(data["cell_0"])
    for f in range(5):
        if stored(c0[:, f]):
            b += 8 * c0.shape[0]
    for k in ("cell_1", "face", "boundary"):
        if stored(data[k]):
            b += 8 * np.asarray(data[k]).size
    if op.n_boundary:
        b += 8 * 3 * op.n_boundary * op.n_q_points_1d
    return b


def workload(name):
    """-> (list of Problem, dt or ("cfl", C) per problem, degree label, note)"""
    from oceano_amd import problems
    if name == "vortex_p3_hanging_97k":
        return [problems.vortex(degree=3, refine=4, refine_fraction=0.3, seed=20261202)], \
            [2e-5], 3, "HBM-resident"
    if name == "vortex_p1_160x80":
        return [problems.vortex(degree=1, refine=3)], [2.8e-4], 1, \
            "cache-resident (1.2 MB per state)"
    if name == "vortex_p1_819k":
        return [problems.vortex(degree=1, refine=6)], [3.5e-5], 1, "HBM-resident"
    if name == "channel_p2_250x10":  # the channel deck's CFL
        return [problems.channel(degree=2, nx=250, ny=10)], [("cfl", 0.075)], 2, \
            "cache-resident (0.5 MB per state)"
    if name == "channel_p2_640k":  # the channel refined 4 times: 17.3 M DoF, stored bed
        return [problems.channel(degree=2, nx=4000, ny=160)], [("cfl", 0.075)], 2, \
            "HBM-resident; bed stored at every quadrature point (static data streamed)"
    if name == "ensemble_64":  # CFL 0.15 / (p + 1): RK4 stability of DG(p) in 2D
        groups, _ = problems.ensemble(n_instances=64, seed=11)
        return [groups[p] for p in sorted(groups)], \
            [("cfl", 0.15 / (p + 1)) for p in sorted(groups)], "1-4", \
            "HBM-resident; one operator per degree over the union of its instances"
    raise SystemExit("unknown workload " + name)


def loop_stage_bytes(op, stat, n_stages, k_bytes, c32):
    """Compulsory HBM bytes per stage of the device time loop, averaged over a step: as
    algorithmic_bytes' one-pass stage, except that with c32 (the mixed paired stage at
    p <= 3) the intermediate stage inputs are read and written in centred FP32 form (half
    the bytes of the FP64 vector plus one FP64 centre per cell)."""
    n, nd = op.n_cells, op.dofs_per_cell
    v = 24 * nd * n
    half = v / 2 + 8 * n
    tot = 0.0
    for s in range(n_stages):
        full = algorithmic_bytes(op, s, stat, k_bytes)["stage"]
        if c32:
            full -= (v - half) * ((s > 0) + (s + 1 < n_stages))
        tot += full
    return tot / n_stages


def algorithmic_bytes(op, stage, stat, k_bytes=8):
    """Compulsory HBM bytes of one stage (FP64 state vectors, k vectors of k_bytes per
    value -- 4 on the mixed path, whose k values are FP32 numbers --, int32 topology),
    per kernel.

    stage (one pass): reads zeta_r, q_r, k_h[1..stage], k_q[1..stage], w_h, w_q; writes
    k_h[stage+1], k_q[stage+1], r_h, r_q (w_h, w_q at the last stage); plus per cell the
    face tables (4 kinds + 8 neighbours + 8 face ids, int32), the cell sizes (2 doubles)
    and `stat` bytes of stored static fields (static_bytes; none for the vortex: every
    problem-data field is uniform and elided).  Neighbour traces are re-reads of zeta_r,
    q_r, not compulsory."""
    n, nd = op.n_cells, op.dofs_per_cell
    v1, v2 = 8 * nd * n, 16 * nd * n
    topo = (4 + 8 + 8) * 4 * n
    geom = 16 * n
    kf = k_bytes / 8.0
    one_pass = (v1 + v2) * (1 + 2) + (v1 + v2) * kf * (stage + 1) + topo + geom + stat
    rhs = v1 + v2 + v1 + topo + geom
    dis = v1 + v2 + v2 + stage * v2 + 2 * v2 + topo + geom
    jac = (stage + 1) * v1 + 2 * v1 + geom
    return {"stage": one_pass, "height_rhs": rhs, "discharge": dis,
            "rk_update": (stage + 3) * v2, "height_jacobi": jac}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def dealii_probe():
    """Whether the reference's own build could run on this host: deal.II's CMake package
    in the usual install prefixes (the reference needs deal.II >= 9.5 with p4est and MPI)."""
    import glob
    pats = ["/usr/lib/cmake/deal.II*/deal.IIConfig.cmake",
            "/usr/local/lib/cmake/deal.II*/deal.IIConfig.cmake",
            "/usr/lib/x86_64-linux-gnu/cmake/deal.II*/deal.IIConfig.cmake",
            "/opt/*/lib/cmake/deal.II*/deal.IIConfig.cmake",
            "/usr/share/cmake*/deal.II*/deal.IIConfig.cmake"]
    found = [f for p in pats for f in glob.glob(p)]
    return {"dealii_found": found, "searched": pats}


def cpu_baseline(problem, n_stages, threads=None, min_seconds=1.5, max_stages=256):
    """The CPU restatement (oracle/, its -O3 / FMA / AVX2 build) with its loops on `threads`
    OpenMP threads (default: the box's 16-core share, or fewer if the host has fewer), RK4
    stages from the projected initial state, at least n_stages and until min_seconds have
    passed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from oceano_amd.integrator import butcher_rows
    threads = threads or max(1, min(16, os.cpu_count() or 1))
    o = pyoracle.Oracle.from_mesh(problem.mesh, problem.degree, bcs=problem.bcs, fast=True)
    d = o.initialize_data_at_quadrature(problem.problem_data)
    h, q = o.project_hydro(problem.initial, d["cell_2"][:, 0])
    o.set_threads(threads)
    rows, _ = butcher_rows("stage_4_order_4", 1e-5)
    S = len(rows)
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(S + 1)]
    w_h, w_q, r_h, r_q = h.copy(), q.copy(), o.zeros(1), o.zeros(2)
    t0 = time.perf_counter()
    k = 0
    while k < max_stages and (k < n_stages or time.perf_counter() - t0 < min_seconds):
        s = k % S
        ri_h, ri_q = (w_h, w_q) if s == 0 else (r_h, r_q)
        o.perform_stage_hydro(s, rows[s][:s + 1], ri_h, ri_q, None, ki_h, ki_q, w_h, w_q, r_h,
                              r_q)
        k += 1
    dt = time.perf_counter() - t0
    dofs = problem.dofs()
    return {"value": dofs * k / dt, "unit": "DoF-RK-stages/s", "cores": threads,
            "kind": "port",
            "sample": "%d RK4 stages of the same workload (%d cells, %d DoF) from the "
                      "projected initial state, %.2f s wall on %d OpenMP threads (%s); "
                      "the restatement built -O3 -march=x86-64-v3 with FMA, its loops in "
                      "parallel, face sums atomic"
                      % (k, problem.mesh.n_cells, dofs, dt, threads, cpu_model()),
            "reference_build": dealii_probe()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed RK stages")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default=DEFAULT, choices=WORKLOADS)
    ap.add_argument("--dt", type=float, default=None,
                    help="RK4 step (default: the workload's; 2e-5 on the default mesh, "
                         "where 6.5e-5 is unstable)")
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--cpu-stages", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--slot-chunk", type=int, default=None,
                    help="library option slot_chunk (cell grouping by face signature)")
    ap.add_argument("--paired", type=int, default=None, choices=[0, 1],
                    help="library option paired_items (a batch's height and discharge "
                         "items in one workgroup, faces split between them)")
    ap.add_argument("--waves-per-block", type=int, default=None, choices=[1, 2, 4],
                    help="library option waves_per_block (work items per workgroup of the "
                         "unpaired stage kernel)")
    ap.add_argument("--ark", action="store_true",
                    help="also time the additive (IMEX) stage on the same workload")
    ap.add_argument("--steady", action="store_true",
                    help="also run the workload to steady state (the channel's criterion: one "
                         "step changes the state by < 1e-10 relative L2) and time the whole run")
    ap.add_argument("--autotune", default="auto", choices=["auto", "on", "off"],
                    help="time the work-layout options per operator and precision and keep "
                         "the fastest (auto: on for every workload; off: library defaults)")
    ap.add_argument("--ensemble-launch", default="auto", choices=["auto", "streams", "single"],
                    help="several operators: one stream per operator (per-degree launches) "
                         "or one launch per stage for all of them; auto times both on "
                         "graph-replayed steps and keeps the faster")
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

    if args.paired is not None:
        from oceano_amd import _capi
        _capi.check(_capi.lib().oc_set_option(b"paired_items", args.paired), "oc_set_option")
    if args.slot_chunk is not None:
        from oceano_amd import _capi
        _capi.check(_capi.lib().oc_set_option(b"slot_chunk", args.slot_chunk), "oc_set_option")
    if args.waves_per_block is not None:
        from oceano_amd import _capi
        _capi.check(_capi.lib().oc_set_option(b"waves_per_block", args.waves_per_block),
                    "oc_set_option")
    groups, dts, degree_label, note = workload(args.workload)
    if args.dt is not None:
        dts = [args.dt] * len(groups)
    S = 4
    multi = len(groups) > 1
    n_steps_timed = args.steps if not multi else S * (-(-args.steps // S))

    tune_on = args.autotune != "off"
    tuned = {}

    def run(precision):
        from oceano_amd import tune
        G = []
        for gi, (P, dt) in enumerate(zip(groups, dts)):
            if tune_on:
                chosen, table = tune.autotune(
                    P, precision, dt=None if isinstance(dt, tuple) else dt,
                    cfl=dt[1] if isinstance(dt, tuple) else None)
                tuned.setdefault(precision, []).append(
                    {"options": chosen, "us_per_step": {json.dumps(o): t * 1e6
                                                        for o, t in table}})
            op, data, w_h, w_q = problems.setup_operator(P, precision=precision)
            stat = static_bytes(op, data)
            if tune_on:
                tune.apply_options(tune.LIBRARY_DEFAULTS)
            if isinstance(dt, tuple):  # main.cc:1503-1505: dt = CFL / transport speed
                dt = dt[1] / float("%.3g" % op.transport_speed(w_h, w_q))
            rows, _ = butcher_rows("stage_4_order_4", dt)
            G.append({"P": P, "op": op, "w_h": w_h, "w_q": w_q, "dt": dt, "rows": rows,
                      "static_bytes": stat,
                      "rk": DeviceIntegrator(op, "stage_4_order_4"),
                      "ki_h": [op.initialize_vector(0) for _ in range(S + 1)],
                      "ki_q": [op.initialize_vector(1) for _ in range(S + 1)],
                      "r_h": [op.initialize_vector(0) for _ in range(2)],
                      "r_q": [op.initialize_vector(1) for _ in range(2)], "steps": 0})

        def stage(g, k):  # eager stage k of the running step of operator g
            s = k % S
            w_h, w_q, r_h, r_q = g["w_h"], g["w_q"], g["r_h"], g["r_q"]
            ri = [w_h, w_q] if s == 0 else [r_h[s % 2], r_q[s % 2]]
            g["op"].perform_stage_hydro(s, 0.0, g["rows"][s][:s + 1], ri + [None, None],
                                        g["ki_h"], g["ki_q"], w_h, w_q, r_h[(s + 1) % 2],
                                        r_q[(s + 1) % 2])

        def finish_step(g, k):
            while k % S:
                stage(g, k)
                k += 1
            g["steps"] += k // S

        for g in G:
            for _ in range(max(1, -(-args.warmup // S))):
                g["rk"].step(g["dt"], g["w_h"], g["w_q"])
                g["steps"] += 1
        # headline: the device-resident loop -- the timed steps captured into one HIP graph
        # per operator and replayed once, plus any remaining stages eagerly: exactly
        # n_steps_timed stages between two synchronisations, timed by HIP events on the
        # replay stream (the host clock around the region is reported beside it)
        side = torch.cuda.Stream()
        full, rest = divmod(n_steps_timed, S)
        streams = [side] + [torch.cuda.Stream() for _ in G[1:]]
        launch = None
        if multi:
            # several operators: one stream each, or all in one launch per stage
            from oceano_amd.integrator import EnsembleIntegrator
            ens = EnsembleIntegrator([g["op"] for g in G])
            dts_g = [g["dt"] for g in G]
            cand = {}
            for mode in (["streams", "single"] if args.ensemble_launch == "auto"
                         else [args.ensemble_launch]):
                torch.cuda.synchronize()
                with torch.cuda.stream(side):
                    if mode == "single":
                        ens.capture(dts_g, [g["w_h"] for g in G], [g["w_q"] for g in G],
                                    n_steps=2)
                        ens.replay(1)
                    else:
                        for g in G:
                            g["rk"].capture(g["dt"], g["w_h"], g["w_q"], n_steps=2)
                            g["rk"].replay(1)
                side.synchronize()
                for g in G:
                    g["steps"] += 2
                t0 = time.perf_counter()
                if mode == "single":
                    with torch.cuda.stream(side):
                        ens.replay(2)
                else:
                    for g, st in zip(G, streams):
                        with torch.cuda.stream(st):
                            g["rk"].replay(2)
                torch.cuda.synchronize()
                cand[mode] = (time.perf_counter() - t0) / 4 * 1e6
                for g in G:
                    g["steps"] += 4
            launch = min(cand, key=cand.get)
            tuned.setdefault("launch", {})[precision] = {"chosen": launch,
                                                         "us_per_step": cand}
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            if launch == "single":
                ens.capture(dts_g, [g["w_h"] for g in G], [g["w_q"] for g in G],
                            n_steps=max(1, full))
                ens.replay(1)  # first launch of the graph: outside the timed region
            else:
                for g in G:
                    g["rk"].capture(g["dt"], g["w_h"], g["w_q"], n_steps=max(1, full))
                    g["rk"].replay(1)  # first launch of the graph: outside the timed region
            for g in G:
                g["steps"] += max(1, full)
        side.synchronize()
        # independent operators (the ensemble's degree groups) replay on streams of their
        # own, so that one group's kernels overlap another's tail
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        quiesce()
        t0 = time.perf_counter()
        ev0.record(side)
        for st in streams[1:]:
            st.wait_event(ev0)
        if launch == "single":
            if full:
                with torch.cuda.stream(side):
                    ens.replay(1)
        else:
            for g, st in zip(G, streams):
                if full:
                    with torch.cuda.stream(st):
                        g["rk"].replay(1)
        with torch.cuda.stream(side):
            for k in range(rest):
                stage(G[0], k)
        for st in streams[1:]:
            e = torch.cuda.Event()
            e.record(st)
            side.wait_event(e)
        ev1.record(side)
        for st in streams:
            st.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        elapsed = ev0.elapsed_time(ev1) * 1e-3
        gc.enable()
        for g in G:
            g["steps"] += full
        if rest:
            finish_step(G[0], rest)
        # per-kernel durations: the same stages launched eagerly with HIP events around
        # each kernel on the launch stream
        for g in G:
            g["op"].set_timing(True)
        quiesce()
        t1 = time.perf_counter()
        for k in range(n_steps_timed):
            for g in G:
                stage(g, k)
        torch.cuda.synchronize()
        eager = time.perf_counter() - t1
        gc.enable()
        kern = {}
        per_op = []
        for g in G:
            op = g["op"]
            rep = op.timing_report()
            op.set_timing(False)
            finish_step(g, n_steps_timed)
            per_stage = [algorithmic_bytes(op, s, g["static_bytes"],
                                           4 if precision == "mixed" else 8)
                         for s in range(S)]
            for name, (ms, n) in rep.items():
                if n == 0:
                    continue
                e = kern.setdefault(name, {"ms": 0.0, "launches": 0, "bytes": 0.0})
                e["ms"] += ms
                e["launches"] += n
                b = sum(per_stage[j % S][name] for j in range(n_steps_timed))
                e["bytes"] += b
                if name == "stage":  # each operator's own stage kernel (per degree)
                    gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
                    per_op.append({"degree": g["P"].degree, "cells": g["P"].mesh.n_cells,
                                   "stage_us": ms / n * 1e3, "alg_bytes": b / n,
                                   "alg_GBs": gbs,
                                   "frac": gbs / HBM_PEAK_GBS if gbs else None})
        torch.cuda.synchronize()
        kernels = {}
        for name, e in kern.items():
            # per stage of the whole workload (over all its operators)
            mean_ms = e["ms"] / n_steps_timed
            b = e["bytes"] / n_steps_timed
            kernels[name] = {"mean_ms": mean_ms, "launches": e["launches"], "alg_bytes": b,
                             "alg_GBs": b / (mean_ms * 1e-3) / 1e9 if mean_ms > 0 else None}
        out = {"elapsed": elapsed, "wall": wall, "eager": eager, "kernels": kernels,
               "per_operator": per_op,
               "groups": G,
               "nonfinite": sum(g["op"].count_nonfinite(g["w_h"]) +
                                g["op"].count_nonfinite(g["w_q"]) for g in G),
               "sweeps_last_stage": sum(g["op"].last_sweeps() for g in G)}
        if all(g["P"].exact is not None for g in G):
            errs = [g["op"].compute_errors(g["w_h"], g["w_q"], g["P"].exact,
                                           g["steps"] * g["dt"]) for g in G]
            out["errors"] = {k: float(np.sqrt(sum(e[k] ** 2 for e in errs)))
                             if k.startswith("l2") else float(max(e[k] for e in errs))
                             for k in errs[0]}
            out["t_end"] = [g["steps"] * g["dt"] for g in G]
        out["state"] = [(g["op"].unpack(g["w_h"], 1).cpu().numpy(),
                         g["op"].unpack(g["w_q"], 2).cpu().numpy()) for g in G]
        return out

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
    head_p = "mixed" if "mixed" in runs else "fp64"
    head = runs[head_p]
    dofs = sum(P.dofs() for P in groups)
    value = dofs * n_steps_timed / head["elapsed"]
    dom = max(head["kernels"], key=lambda k: head["kernels"][k]["mean_ms"])
    achieved = head["kernels"][dom]["alg_GBs"]
    traffic_tab, traffic_src = measured_traffic(args.workload)
    traffic = traffic_tab.get(head_p) if dom == "stage" else None
    loop_roof = None
    if not multi:
        # the timed region is the device time loop: its stages' own compulsory bytes over its
        # time per stage (the FP32 paired stage at p <= 3 passes centred FP32 stage inputs)
        g0 = head["groups"][0]
        opts = tuned.get(head_p, [{}])[0].get("options", {}) if tune_on else {}
        paired = opts.get("paired_items", -1)
        paired = (g0["P"].degree <= 3) if paired == -1 else bool(paired)
        c32 = (head_p == "mixed" and paired and g0["P"].degree <= 3
               and g0["P"].viscosity == 0.0)
        lb = loop_stage_bytes(g0["op"], g0["static_bytes"], S,
                              4 if head_p == "mixed" else 8, c32)
        lach = lb / (head["elapsed"] / n_steps_timed) / 1e9
        loop_roof = {"achieved": lach, "frac": lach / HBM_PEAK_GBS,
                     "alg_bytes_per_stage": lb, "us_per_stage":
                     head["elapsed"] / n_steps_timed * 1e6,
                     "stage_inputs": "centred FP32" if c32 else "FP64"}
    line = {
        "metric": "DoF-RK-stages/s on 1 GPU, mixed vs FP64 at matched L2 error; % HBM peak",
        "value": value,
        "unit": "DoF-RK-stages/s",
        "n_gpus": 1,
        "steps": n_steps_timed,
        "warmup": args.warmup,
        "ms_per_step": head["elapsed"] / n_steps_timed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if "mixed" in runs else "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "cells": sum(P.mesh.n_cells for P in groups),
                   "hanging_subfaces": sum(getattr(P.mesh, "n_hanging", 0) for P in groups),
                   "dofs": dofs, "degree": degree_label, "operators": len(groups),
                   "scheme": "classic RK4 (Butcher), stage = step",
                   "dt": [g["dt"] for g in head["groups"]] if multi else head["groups"][0]["dt"],
                   "precision": head_p, "memory": note,
                   "autotuned": tuned if (tune_on or tuned) else None},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": traffic,
                     "traffic_source": traffic_src if traffic is not None else None,
                     "alg_bytes_per_launch": head["kernels"][dom]["alg_bytes"],
                     "loop": loop_roof},
        "timing": "HIP events on the replay stream around the timed region (one captured "
                  "graph of the timed steps per operator); host clock beside it",
    }
    for p, r in runs.items():
        line[p] = {"value": dofs * n_steps_timed / r["elapsed"],
                   "ms_per_step": r["elapsed"] / n_steps_timed * 1e3,
                   "wall_value": dofs * n_steps_timed / r["wall"],
                   "eager_value": dofs * n_steps_timed / r["eager"], "kernels": r["kernels"],
                   "nonfinite": r["nonfinite"],
                   "jacobi_sweeps_last_stage": r["sweeps_last_stage"]}
        if multi:  # each degree group's stage kernel against the HBM roofline
            line[p]["per_operator"] = r["per_operator"]
        if "errors" in r:
            line[p]["l2_error_vs_exact"] = r["errors"]
            line[p]["t_end"] = r["t_end"] if multi else r["t_end"][0]
    if len(runs) == 2:
        def rel(a, b):
            return float(np.linalg.norm(a - b) / np.linalg.norm(b))
        fs, ms = runs["fp64"]["state"], runs["mixed"]["state"]
        nds = [g["op"].dofs_per_cell for g in runs["fp64"]["groups"]]

        def comp(states, c):
            return np.concatenate([s[1].reshape(-1, 2, nd)[:, c].ravel()
                                   for s, nd in zip(states, nds)])
        qn = float(np.sqrt(np.linalg.norm(comp(fs, 0)) ** 2 + np.linalg.norm(comp(fs, 1)) ** 2))
        zf = np.concatenate([s[0] for s in fs])
        zm = np.concatenate([s[0] for s in ms])
        qmax = max(np.abs(comp(fs, 0)).max(), np.abs(comp(fs, 1)).max())
        line["mixed_vs_fp64"] = {
            "speedup": line["mixed"]["value"] / line["fp64"]["value"],
            "rel_l2_zeta": rel(zm, zf),
            # per component, relative to the norm of the whole discharge (hv ~ 0 in a channel)
            "rel_l2_hu": float(np.linalg.norm(comp(ms, 0) - comp(fs, 0)) / qn),
            "rel_l2_hv": float(np.linalg.norm(comp(ms, 1) - comp(fs, 1)) / qn),
            # nodal maxima, relative to the largest |zeta| and the largest |hu|, |hv|
            "rel_linf_zeta": float(np.abs(zm - zf).max() / np.abs(zf).max()),
            "rel_linf_hu": float(np.abs(comp(ms, 0) - comp(fs, 0)).max() / qmax),
            "rel_linf_hv": float(np.abs(comp(ms, 1) - comp(fs, 1)).max() / qmax)}
    if args.ark and not multi:
        line["ark"] = {p: run_ark(groups[0], dts[0], p, n_steps_timed) for p in precisions}
    if args.steady and not multi:
        line["steady"] = {p: run_steady(groups[0], dts[0], p) for p in precisions}
    if not args.no_cpu and not multi:
        line["cpu_baseline"] = cpu_baseline(groups[0], args.cpu_stages)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


def quiesce():
    """Before a timed region: collect garbage now and keep the collector off until the region
    ends (gc.enable()), so that no operator of an earlier leg is freed inside it -- freeing
    device memory synchronises the device and stalls the launch pipeline."""
    import torch
    gc.collect()
    gc.disable()
    torch.cuda.synchronize()


def run_steady(problem, dt, precision, tol=1e-10, chunk=2000, max_steps=400000):
    """The whole steady-state job (C3's criterion): graph-replayed RK4 steps from the
    initial state until one step changes (zeta, q) by less than tol in relative L2, checked
    every `chunk` steps; stops early when the change has not halved over 10 checks (a
    rounding floor).  Wall time includes the checks."""
    import torch
    from oceano_amd import problems
    from oceano_amd.integrator import DeviceIntegrator
    op, data, w_h, w_q = problems.setup_operator(problem, precision=precision)
    if isinstance(dt, tuple):
        dt = dt[1] / float("%.3g" % op.transport_speed(w_h, w_q))
    rk = DeviceIntegrator(op)
    history = []
    steps, change = 0, float("inf")
    quiesce()
    t0 = time.perf_counter()
    while steps < max_steps:
        n, change = rk.run_to_steady(w_h, w_q, dt, tol=tol, max_steps=chunk, chunk=chunk)
        steps += n
        history.append(change)
        if not change == change or change < tol:
            break
        if len(history) > 10 and change > 0.5 * history[-11]:
            break
    el = time.perf_counter() - t0
    gc.enable()
    return {"converged": bool(change < tol), "steps": steps, "dt": dt,
            "t_simulated": steps * dt, "last_change": change, "wall_s": el,
            "dof_stages_per_s": problem.dofs() * 4 * steps / el}


def run_ark(problem, dt, precision, n_stages):
    """The additive (IMEX) stage_3_order_2 scheme on the same workload with the library's
    default layout: eager stages (one launch each, k_stage_ark_pair, for an inviscid operator
    with paired work items; else five kernels: L_zeta, explicit L_q, stiff source, depth
    solve, modified-mass update) timed between two synchronisations; DoF-stages/s and the
    per-kernel event times."""
    import torch
    from oceano_amd import problems
    from oceano_amd.integrator import AdditiveRungeKuttaIntegrator
    op, data, w_h, w_q = problems.setup_operator(problem, precision=precision)
    if isinstance(dt, tuple):
        dt = dt[1] / float("%.3g" % op.transport_speed(w_h, w_q))
    rk = AdditiveRungeKuttaIntegrator()
    rk.reinit(op)
    S = rk.n_stages()
    steps = max(1, n_stages // S)
    for i in range(2):
        rk.perform_time_step(op, i * dt, dt, w_h, w_q)
    quiesce()
    t0 = time.perf_counter()
    for i in range(steps):
        rk.perform_time_step(op, i * dt, dt, w_h, w_q)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    gc.enable()
    op.set_timing(True)
    rk.perform_time_step(op, 0.0, dt, w_h, w_q)
    torch.cuda.synchronize()
    rep = op.timing_report()
    op.set_timing(False)
    return {"value": problem.dofs() * steps * S / el, "ms_per_stage": el / (steps * S) * 1e3,
            "stages": steps * S, "dt": dt, "nonfinite": op.count_nonfinite(w_h) +
            op.count_nonfinite(w_q),
            "kernel_ms_per_stage": {k: v[0] / S for k, v in rep.items() if v[1]}}


if __name__ == "__main__":
    main()
