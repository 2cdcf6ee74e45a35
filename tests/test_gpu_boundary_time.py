"""Time-dependent boundary data (OceanDG.h:1798-1803: the inflow Functions get the stage
time before every stage) on the reference's own vortex set-up: supercritical inflow at
x = 0 whose data is the exact travelling vortex, supercritical outflow at x = 2, walls at
y = 0, 1 (Icbc_ShallowWaterVortex.h:262-271), u_inf = 6, radius 0.25.

The GPU runs 200 RK4 steps with a boundary table of the stage times; the oracle gets the
boundary values at each stage time before each stage.  Bar: 1e-12 relative L2, and the
result within the scheme's discretisation error of the exact vortex.
"""
import numpy as np
import pytest
import torch

from helpers import make_pair, rel_l2, to_dev, to_host
from oceano_amd import _capi, problems
from oceano_amd.integrator import DeviceIntegrator, butcher_rows

pytestmark = pytest.mark.gpu

N_STEPS = 200


def _problem(p=2):
    # centred on the inflow boundary: the inflow data change as the vortex enters
    return problems.vortex(degree=p, refine=1, periodic=False, u_inf=(6.0, 0.0), radius=0.25,
                           center=(0.1, 0.5))


def _dt(P):
    hmin = float(np.min(P.mesh.geom[:, 2:]))
    return 0.2 * hmin / ((6.0 + np.sqrt(9.81)) * (2 * P.degree + 1))


def _oracle_steps(P, op, o, h, q, dt, n_steps, frozen=False):
    rows, c = butcher_rows("stage_4_order_4", dt)
    S = len(rows)
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(S + 1)]
    r_h, r_q = o.zeros(1), o.zeros(2)
    w_h, w_q = h.copy(), q.copy()
    t = 0.0
    for _ in range(n_steps):
        for s in range(S):
            o.set_boundary_values(P.boundary_values(op, 0.0 if frozen else t + c[s] * dt))
            ri_h, ri_q = (w_h.copy(), w_q.copy()) if s == 0 else (r_h.copy(), r_q.copy())
            o.perform_stage_hydro(s, rows[s][:s + 1], ri_h, ri_q, None, ki_h, ki_q, w_h, w_q,
                                  r_h, r_q)
        t += dt
    return w_h, w_q


def _table(P, op, dt, n_steps):
    _, c = butcher_rows("stage_4_order_4", dt)
    return problems.boundary_table(P, op, 0.0, dt, n_steps, c)


def test_time_dependent_inflow_matches_oracle(cuda):
    print("[boundary-time] vortex p=2, %d RK4 steps, eager" % N_STEPS, flush=True)
    P = _problem()
    op, o, h, q = make_pair(P)
    dt = _dt(P)
    times, vals = _table(P, op, dt, N_STEPS)
    op.set_boundary_table(times, vals)
    rk = DeviceIntegrator(op)
    rk.set_time(0.0)
    gh, gq = to_dev(op, h, 1), to_dev(op, q, 2)
    for _ in range(N_STEPS):
        rk.step(dt, gh, gq)
    torch.cuda.synchronize()
    assert abs(rk.time() - N_STEPS * dt) < 1e-12
    w_h, w_q = _oracle_steps(P, op, o, h, q, dt, N_STEPS)
    eh, eq = rel_l2(to_host(op, gh, 1), w_h), rel_l2(to_host(op, gq, 2), w_q)
    dh = rel_l2(to_host(op, gh, 1) - h, w_h - h)
    print("  gpu vs oracle: zeta %.2e q %.2e d(zeta) %.2e" % (eh, eq, dh))
    assert eh <= 1e-12 and eq <= 1e-12 and dh <= 1e-9
    # the scheme's error against the exact vortex (the inflow is exact at every stage)
    err = op.compute_errors(gh, gq, P.exact, N_STEPS * dt)
    print("  vs exact at t=%.4f: depth %.2e discharge %.2e" % (N_STEPS * dt, err["l2_depth"],
                                                              err["l2_discharge"]))
    assert err["l2_depth"] < 2e-4 and err["l2_discharge"] < 3e-3
    # freezing the inflow at t = 0 (what a stage that ignores its time does) is visibly off
    f_h, f_q = _oracle_steps(P, op, o, h, q, dt, N_STEPS, frozen=True)
    assert rel_l2(f_q, w_q) > 1e-6


def test_replayed_steps_read_the_table(cuda):
    """A captured step replayed N times reads the table through the device entry counter:
    bitwise the eager steps."""
    print("[boundary-time] vortex p=2, capture + replay", flush=True)
    P = _problem()
    op, o, h, q = make_pair(P)
    dt = _dt(P)
    n = 60
    times, vals = _table(P, op, dt, n)
    op.set_boundary_table(times, vals)
    rk = DeviceIntegrator(op)
    eh, eq = to_dev(op, h, 1), to_dev(op, q, 2)
    rk.set_time(0.0)
    for _ in range(n):
        rk.step(dt, eh, eq)
    gh, gq = to_dev(op, h, 1), to_dev(op, q, 2)
    rk.set_time(0.0)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        rk.capture(dt, gh, gq)
        rk.replay(n // 2)
        rk.replay(n - n // 2)   # a second batch starts at the entry of its own time
    side.synchronize()
    assert abs(rk.time() - n * dt) < 1e-12
    assert torch.equal(gh, eh) and torch.equal(gq, eq)
    # a replay past the end of the table is refused
    with pytest.raises(RuntimeError, match="boundary table"):
        with torch.cuda.stream(side):
            rk.replay(1)
    side.synchronize()


def test_untabulated_time_and_stale_graph_errors(cuda):
    P = _problem(1)
    op, o, h, q = make_pair(P)
    dt = _dt(P)
    times, vals = _table(P, op, dt, 2)
    op.set_boundary_table(times, vals)
    gh, gq = to_dev(op, h, 1), to_dev(op, q, 2)
    Lh, Lq = op.initialize_vector(0), op.initialize_vector(1)
    op.rhs([gh, gq], [Lh, Lq], current_time=times[1])
    with pytest.raises(RuntimeError, match="no boundary data tabulated"):
        op.rhs([gh, gq], [Lh, Lq], current_time=0.3 * dt)
    with pytest.raises(RuntimeError, match="no boundary data tabulated"):
        op.apply(0.7 * dt, [gh, gq], [Lh, Lq])
    with pytest.raises(RuntimeError, match="non-decreasing"):
        op.set_boundary_table(times[::-1].copy(), vals)
    # a graph is invalidated by any later change of the data it holds
    rk = DeviceIntegrator(op)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        rk.capture(dt, gh, gq)
        rk.replay(1)
    side.synchronize()
    op.set_boundary_table(times, vals)
    rk.set_time(0.0)
    with pytest.raises(RuntimeError, match="capture again"):
        with torch.cuda.stream(side):
            rk.replay(1)
    with torch.cuda.stream(side):
        rk.capture(dt, gh, gq)
        rk.replay(1)
    side.synchronize()
    op.initialize_data_at_quadrature(P.problem_data)
    rk.set_time(0.0)
    with pytest.raises(RuntimeError, match="capture again"):
        with torch.cuda.stream(side):
            rk.replay(1)
    side.synchronize()


def test_periodic_and_wall_operators_accept_any_time(cuda):
    """no boundary face reads data: every time is accepted, also without a table"""
    P = problems.vortex(degree=1, refine=0)
    op, o, h, q = make_pair(P)
    gh, gq = to_dev(op, h, 1), to_dev(op, q, 2)
    Lh, Lq = op.initialize_vector(0), op.initialize_vector(1)
    op.rhs([gh, gq], [Lh, Lq], current_time=123.456)
    op.apply(-7.0, [gh, gq], [Lh, Lq])
    torch.cuda.synchronize()
    assert _capi is not None
