"""Parity at the benchmark configurations' full sizes: one classic RK4 step (4 stages, every
k vector) of C1 (vortex p1 160x80), C2 (vortex p3, 20x10 refined 4x + 30% seeded
refinement: 97,280 cells, 85,944 hanging subfaces) and C3 (channel p2 250x10) against the
oracle, the mixed path against FP64, and size-independent properties of the full C2 mesh
(free-stream preservation, conservation of the periodic volume)."""
import numpy as np
import pytest
import torch

from helpers import make_pair, rel_l2, to_dev, to_host
from oceano_amd import problems
from oceano_amd.integrator import butcher_rows

pytestmark = pytest.mark.gpu

CONFIGS = {
    "C1_vortex_p1_160x80": (lambda: problems.vortex(degree=1, refine=3), 2.8e-4),
    "C2_vortex_p3_hanging_97k": (lambda: problems.vortex(degree=3, refine=4,
                                                         refine_fraction=0.3), 2e-5),
    "C3_channel_p2_250x10": (lambda: problems.channel(degree=2, nx=250, ny=10), 1e-3),
}


def _gpu_step(op, h, q, rows):
    S = len(rows)
    w_h, w_q = to_dev(op, h, 1), to_dev(op, q, 2)
    ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
    ki_q = [op.initialize_vector(1) for _ in range(S + 1)]
    r_h = [op.initialize_vector(0) for _ in range(2)]
    r_q = [op.initialize_vector(1) for _ in range(2)]
    for s in range(S):
        ri = [w_h, w_q] if s == 0 else [r_h[s % 2], r_q[s % 2]]
        op.perform_stage_hydro(s, 0.0, rows[s][:s + 1], ri + [None, None], ki_h, ki_q, w_h,
                               w_q, r_h[(s + 1) % 2], r_q[(s + 1) % 2])
    torch.cuda.synchronize()
    return to_host(op, w_h, 1), to_host(op, w_q, 2), ki_h, ki_q


@pytest.mark.parametrize("name", list(CONFIGS))
def test_full_size_rk4_step_matches_oracle(cuda, name):
    make, dt = CONFIGS[name]
    P = make()
    print("[full-size] %s: %d cells, %d DoF" % (name, P.mesh.n_cells, P.dofs()), flush=True)
    op, o, h, q = make_pair(P)
    rows, _ = butcher_rows("stage_4_order_4", dt)
    S = len(rows)
    gh, gq, ki_h, ki_q = _gpu_step(op, h, q, rows)
    oh, oq = h.copy(), q.copy()
    oki_h = [o.zeros(1) for _ in range(S + 1)]
    oki_q = [o.zeros(2) for _ in range(S + 1)]
    o.time_step(rows, oh, oq, oki_h, oki_q, o.zeros(1), o.zeros(2))
    eh, eq = rel_l2(gh, oh), rel_l2(gq, oq)
    ek = max(max(rel_l2(to_host(op, ki_h[j], 1), oki_h[j]) for j in range(1, S + 1)),
             max(rel_l2(to_host(op, ki_q[j], 2), oki_q[j]) for j in range(1, S + 1)))
    kd = max(max(float(np.linalg.norm(to_host(op, ki_h[j], 1) - oki_h[j])) for j in
                 range(1, S + 1)),
             max(float(np.linalg.norm(to_host(op, ki_q[j], 2) - oki_q[j])) for j in
                 range(1, S + 1)))
    print("  FP64 step vs oracle: zeta %.2e q %.2e; k vectors: worst relative %.2e, "
          "worst absolute x dt / |w| %.2e" % (eh, eq, ek, kd * dt / np.linalg.norm(oh)))
    # the states to 1e-12.  A k vector's own norm can be tiny (L_zeta of a nearly uniform
    # discharge is a residual of cancelling fluxes), so its relative difference is no
    # measure; its contribution dt k to the state is bounded instead.
    assert eh < 1e-12 and eq < 1e-12
    assert kd * dt < 1e-13 * np.linalg.norm(oh)
    # mixed precision on the same step: within 1e-6 of FP64 (the change over the step is
    # resolved to 1e-4 relative)
    opm, _, _, _ = problems.setup_operator(P, precision="mixed")
    mh, mq, _, _ = _gpu_step(opm, h, q, rows)
    dh, dq = rel_l2(mh, gh), rel_l2(mq, gq)
    ch = rel_l2(mh - h, gh - h)
    print("  mixed vs FP64: zeta %.2e q %.2e, zeta increment %.2e" % (dh, dq, ch))
    assert dh < 1e-6 and dq < 1e-6


def test_full_c2_free_stream_and_volume(cuda):
    """free stream (zeta = 1, q = (1, 0), flat bed) on the full hanging C2 mesh: L = 0 to
    the round-off of each precision; the vortex's volume is conserved over 10 RK4 steps"""
    P = problems.vortex(degree=3, refine=4, refine_fraction=0.3)
    print("[full-size] C2 free stream, %d cells" % P.mesh.n_cells, flush=True)
    for precision in ("fp64", "mixed"):
        op, data, w_h, w_q = problems.setup_operator(P, precision=precision)
        n = P.mesh.n_cells * op.dofs_per_cell
        h = np.ones(n)
        q = np.zeros((P.mesh.n_cells, 2, op.dofs_per_cell))
        q[:, 0] = 1.0
        zh, zq = to_dev(op, h, 1), to_dev(op, q.ravel(), 2)
        dh, dq = op.initialize_vector(0), op.initialize_vector(1)
        op.rhs([zh, zq], [dh, dq])
        torch.cuda.synchronize()
        mh, mq = float(dh.abs().max()), float(dq.abs().max())
        print("  %s: max |L_zeta| %.2e, max |L_q| %.2e" % (precision, mh, mq))
        # FP64: round-off of the cancelling flux terms (each ~ |F| h ~ 6e-3); FP32
        # integrals: ~1e-7 of those terms (the lake at rest, not the free stream, is exact
        # in FP32 thanks to the surface deviation)
        bound = 1e-15 if precision == "fp64" else 1e-8
        assert mh < bound and mq < bound
        from oceano_amd.integrator import DeviceIntegrator
        rk = DeviceIntegrator(op)
        v0 = op.volume(w_h)
        for _ in range(10):
            rk.step(2e-5, w_h, w_q)
        v1 = op.volume(w_h)
        print("  %s: volume drift over 10 steps %.2e" % (precision, (v1 - v0) / v0))
        # FP64: exact to round-off (face fluxes cancel bitwise, every Jacobi sweep removes
        # the whole residual mass); FP32 integrals: the cell terms sum to zero over the test
        # functions only to FP32 rounding of the increments
        assert abs(v1 - v0) < (1e-13 if precision == "fp64" else 1e-8) * v0


def test_full_c2_long_run_mixed_within_bar(cuda):
    """2000 RK4 steps (t = 0.04) of the full C2 mesh, graph-replayed: the FP64 error against
    the exact vortex stays at its discretisation level; the mixed path's error, set by FP32
    rounding at this resolution (its discretisation error is ~100x smaller), stays below the
    1e-6 bar and within 1e-6 of FP64"""
    from oceano_amd.integrator import DeviceIntegrator
    P = problems.vortex(degree=3, refine=4, refine_fraction=0.3)
    out = {}
    for precision in ("fp64", "mixed"):
        print("[full-size] C2 long run %s, %d cells" % (precision, P.mesh.n_cells), flush=True)
        op, data, w_h, w_q = problems.setup_operator(P, precision=precision)
        rk = DeviceIntegrator(op)
        side = torch.cuda.Stream()
        torch.cuda.synchronize()
        with torch.cuda.stream(side):
            rk.capture(2e-5, w_h, w_q)
            rk.replay(2000)
        side.synchronize()
        e = op.compute_errors(w_h, w_q, P.exact, 2000 * 2e-5)
        print("  %s: l2 depth %.3e hu %.3e hv %.3e" % (precision, e["l2_depth"], e["l2_hu"],
                                                        e["l2_hv"]))
        out[precision] = (to_host(op, w_h, 1), to_host(op, w_q, 2), e)
    (fh, fq, fe), (mh, mq, me) = out["fp64"], out["mixed"]
    dh, dq = rel_l2(mh, fh), rel_l2(mq, fq)
    print("  mixed vs fp64: zeta %.3e q %.3e" % (dh, dq))
    assert fe["l2_depth"] < 1e-9 and fe["l2_hu"] < 1e-8 and fe["l2_hv"] < 1e-8
    assert all(me[k] < 1e-6 for k in ("l2_depth", "l2_hu", "l2_hv"))
    assert dh < 1e-6 and dq < 1e-6
