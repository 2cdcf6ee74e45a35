"""Per-operator tuning of the stage kernels' work layout (no counterpart in the reference:
deal.II's MatrixFree picks its own cell batching).

The tunable parameters are the library options read at operator creation:
  slot_chunk       cell slots grouped by face signature within chunks of this many batches
                   (-1: the caller's order, 0: one XCD's share)
  xcd_swizzle      each XCD runs a contiguous batch range (1) or the round-robin order (0)
  waves_per_block  64-cell work items per workgroup of the one-pass stage kernel (1, 2, 4):
                   64, 128 or 256 cells per workgroup, the LDS of each wave its own slice
  paired_items     a batch's height and discharge items in one two-wave workgroup that
                   splits the faces between them (1), or independent items (0); -1 picks
                   by degree (paired for p <= 3)
Results do not depend on the first three (each cell's arithmetic is the same) and depend
on paired_items only at rounding level on hanging faces (another summation order); their
speed depends on the degree, the mesh and the precision.  `autotune` times graph-replayed RK4 steps of the
problem for each candidate and returns the fastest setting; `apply_options` sets it for
the operators created next.
"""
import time

from . import _capi

DEFAULT_CANDIDATES = [
    {"slot_chunk": -1, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 0},
    {"slot_chunk": 0, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 0},
    {"slot_chunk": -1, "xcd_swizzle": 0, "waves_per_block": 1, "paired_items": 0},
    {"slot_chunk": -1, "xcd_swizzle": 1, "waves_per_block": 2, "paired_items": 0},
    {"slot_chunk": -1, "xcd_swizzle": 1, "waves_per_block": 4, "paired_items": 0},
    {"slot_chunk": 0, "xcd_swizzle": 1, "waves_per_block": 2, "paired_items": 0},
    {"slot_chunk": 0, "xcd_swizzle": 1, "waves_per_block": 4, "paired_items": 0},
    {"slot_chunk": -1, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 1},
    {"slot_chunk": 0, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 1},
]
LIBRARY_DEFAULTS = {"slot_chunk": 0, "xcd_swizzle": 1, "waves_per_block": 1,
                    "paired_items": -1}


def apply_options(opts):
    for k, v in opts.items():
        _capi.check(_capi.lib().oc_set_option(k.encode(), int(v)), "oc_set_option")


def autotune(problem, precision="mixed", dt=None, cfl=None, candidates=None, steps=8,
             repeats=2, log=None):
    """-> (best options, [(options, seconds per RK4 step)]).  The library options are left
    at the best setting."""
    import torch
    from . import problems
    from .integrator import DeviceIntegrator
    table = []
    for opts in candidates or DEFAULT_CANDIDATES:
        apply_options(opts)
        op, _, w_h, w_q = problems.setup_operator(problem, precision=precision)
        step = dt
        if step is None:
            step = (cfl or 0.05) / float("%.3g" % op.transport_speed(w_h, w_q))
        rk = DeviceIntegrator(op)
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            rk.capture(step, w_h, w_q)
            rk.replay(2)
        side.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            with torch.cuda.stream(side):
                rk.replay(steps)
            side.synchronize()
            best = min(best, (time.perf_counter() - t0) / steps)
        table.append((dict(opts), best))
        if log:
            log("  tune %s: %.1f us per RK4 step" % (opts, best * 1e6))
        del rk, op
    chosen = min(table, key=lambda e: e[1])[0]
    apply_options(chosen)
    return chosen, table
