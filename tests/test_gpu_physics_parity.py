"""GPU (liboceano_dg.so, through the C ABI) vs the CPU oracle, FP64, on the branches the
vortex and channel cases do not reach: wind stress and Coriolis (stored fields and uniform
constants) with Manning friction over a non-flat bed, the absorbing outflow, a supercritical
outflow carrying backflow (the momentum-flux cut), and the viscous flux.

Bar: relative L2 <= 1e-12 per vector for L, M^-1 L and every vector of one RK4 step.
"""
import numpy as np
import pytest
import torch

from helpers import make_pair, rel_l2, to_dev, to_host
from oceano_amd import problems
from oceano_amd.integrator import butcher_rows

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _vortex_bc(p, bcs, **kw):
    return problems.vortex(degree=p, refine=1, refine_fraction=0.3, seed=7, periodic=False,
                           bcs=bcs, **kw)


CASES = {
    "forced_stored": lambda p: problems.forced(p),
    "forced_uniform": lambda p: problems.forced(p, uniform=True),
    "vortex_absorbing": lambda p: _vortex_bc(p, {0: "wall", 1: "supercritical_inflow",
                                                 2: "absorbing_outflow"}, u_inf=(2.0, 0.0)),
    # vortex centred at the outflow with a slow background flow: part of the outflow
    # boundary has q.n < 0 (backflow), where the momentum flux is cut
    "vortex_backflow": lambda p: _vortex_bc(p, None, u_inf=(0.1, 0.0), center=(1.8, 0.5)),
    "vortex_viscous": lambda p: _viscous(p),
    "height_discharge_inflow": lambda p: _vortex_bc(p, {0: "wall", 1: "discharge_inflow",
                                                        2: "height_inflow"}),
}


def _viscous(p):
    P = problems.vortex(degree=p, refine=1, refine_fraction=0.3, seed=8)
    P.viscosity = 2e-3
    return P


def _state(P, h, q, degree):
    rng = np.random.default_rng(200 + degree)
    return (h + 1e-3 * rng.standard_normal(h.shape), q + 1e-2 * rng.standard_normal(q.shape),
            rng)


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
@pytest.mark.parametrize("case", list(CASES))
def test_rhs_apply_physics(cuda, case, degree):
    print("[physics] case=%s degree=%d" % (case, degree), flush=True)
    P = CASES[case](degree)
    op, o, h, q = make_pair(P)
    h, q, rng = _state(P, h, q, degree)
    v = 0.5 * rng.standard_normal(q.shape) if P.viscosity else None
    Lh_ref, Lq_ref = o.rhs(h, q, v)
    src = [to_dev(op, h, 1), to_dev(op, q, 2)] + ([to_dev(op, v, 2)] if v is not None else [])
    Lh, Lq = op.initialize_vector(0), op.initialize_vector(1)
    op.rhs(src, [Lh, Lq])
    torch.cuda.synchronize()
    eh, eq = rel_l2(to_host(op, Lh, 1), Lh_ref), rel_l2(to_host(op, Lq, 2), Lq_ref)
    print("  rhs rel L2: L_zeta %.3e  L_q %.3e" % (eh, eq))
    assert eh <= TOL and eq <= TOL
    op.apply(0.0, src, [Lh, Lq])
    torch.cuda.synchronize()
    eh = rel_l2(to_host(op, Lh, 1), o.inverse_mass_tensor(Lh_ref, 1))
    eq = rel_l2(to_host(op, Lq, 2), o.inverse_mass_tensor(Lq_ref, 2))
    print("  apply rel L2: %.3e %.3e" % (eh, eq))
    assert eh <= TOL and eq <= TOL


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
@pytest.mark.parametrize("case", list(CASES))
def test_rk4_step_physics(cuda, case, degree):
    print("[physics] case=%s degree=%d" % (case, degree), flush=True)
    P = CASES[case](degree)
    op, o, h, q = make_pair(P)
    h, q, rng = _state(P, h, q, degree)
    v = 0.5 * rng.standard_normal(q.shape) if P.viscosity else None
    hmin = float(np.min(P.mesh.geom[:, 2:]))
    dt = 0.1 * hmin / (8.0 * (2 * degree + 1))
    rows, _ = butcher_rows("stage_4_order_4", dt)
    S = len(rows)
    w_h, w_q = h.copy(), q.copy()
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(S + 1)]
    o.time_step(rows, w_h, w_q, ki_h, ki_q, o.zeros(1), o.zeros(2), v=v)
    gv = to_dev(op, v, 2) if v is not None else None
    gw_h, gw_q = to_dev(op, h, 1), to_dev(op, q, 2)
    gki_h = [op.initialize_vector(0) for _ in range(S + 1)]
    gki_q = [op.initialize_vector(1) for _ in range(S + 1)]
    gr_h = [op.initialize_vector(0), op.initialize_vector(0)]
    gr_q = [op.initialize_vector(1), op.initialize_vector(1)]
    for s in range(S):
        ri = [gw_h, gw_q] if s == 0 else [gr_h[s % 2], gr_q[s % 2]]
        nxt = (s + 1) % 2
        op.perform_stage_hydro(s, 0.0, rows[s][:s + 1], ri + [None, gv], gki_h, gki_q, gw_h,
                               gw_q, gr_h[nxt], gr_q[nxt], write_scratch=True)
    torch.cuda.synchronize()
    errs = {"w_h": rel_l2(to_host(op, gw_h, 1), w_h), "w_q": rel_l2(to_host(op, gw_q, 2), w_q)}
    for j in range(S + 1):
        errs["ki_h[%d]" % j] = rel_l2(to_host(op, gki_h[j], 1), ki_h[j])
        errs["ki_q[%d]" % j] = rel_l2(to_host(op, gki_q[j], 2), ki_q[j])
    print("  " + " ".join("%s=%.2e" % kv for kv in errs.items()))
    assert max(errs.values()) <= TOL
    assert rel_l2(to_host(op, gw_q, 2) - q, w_q - q) <= 1e-9
