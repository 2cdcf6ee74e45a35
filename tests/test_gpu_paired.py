"""Paired work items (library option paired_items): each batch's height and discharge items
run as one two-wave workgroup that splits the faces between them (k_stage_pair).

FP64: one RK4 step against the oracle at the parity bar (1e-12 relative L2 per vector) on
conforming, hanging-node, boundary-condition and channel meshes for every degree.  Mixed
precision: two RK4 steps against the unpaired kernel (the face integrals are summed in
another order, so the results differ at FP32 rounding level; bar 1e-6 relative L2, the
mixed-vs-FP64 bar) and against FP64.
"""
import numpy as np
import pytest
import torch

from helpers import make_pair, rel_l2, to_host
from oceano_amd import problems, tune
from oceano_amd.integrator import DeviceIntegrator, butcher_rows
from test_gpu_parity import CASES, TOL, _run_steps

pytestmark = pytest.mark.gpu


@pytest.fixture
def paired():
    tune.apply_options({"paired_items": 1})
    yield
    tune.apply_options({"paired_items": -1})


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
@pytest.mark.parametrize("case", list(CASES))
def test_paired_rk4_step_fp64_parity(cuda, paired, case, degree):
    print("[paired] case=%s degree=%d" % (case, degree), flush=True)
    P = CASES[case](degree)
    op, o, h, q = make_pair(P)
    rng = np.random.default_rng(200 + degree)
    h = h + 1e-3 * rng.standard_normal(h.shape)
    q = q + 1e-2 * rng.standard_normal(q.shape)
    hmin = float(np.min(P.mesh.geom[:, 2:]))
    dt = 0.1 * hmin / (8.0 * (2 * degree + 1))
    rows, _ = butcher_rows("stage_4_order_4", dt)
    (gw_h, gw_q, gki_h, gki_q), (w_h, w_q, ki_h, ki_q) = _run_steps(op, o, h, q, rows, 1)
    errs = {"w_h": rel_l2(to_host(op, gw_h, 1), w_h), "w_q": rel_l2(to_host(op, gw_q, 2), w_q)}
    for j in range(1, len(ki_h)):
        errs["ki_h[%d]" % j] = rel_l2(to_host(op, gki_h[j], 1), ki_h[j])
        errs["ki_q[%d]" % j] = rel_l2(to_host(op, gki_q[j], 2), ki_q[j])
    print("  " + " ".join("%s=%.2e" % kv for kv in errs.items()))
    assert max(errs.values()) <= TOL


def _steps(P, precision, pair, n=2, dt=2e-4):
    tune.apply_options({"paired_items": pair})
    try:
        op, _, h, q = problems.setup_operator(P, precision=precision)
    finally:
        tune.apply_options({"paired_items": -1})
    rk = DeviceIntegrator(op)
    for _ in range(n):
        rk.step(dt, h, q)
    torch.cuda.synchronize()
    return op.unpack(h, 1).cpu().numpy(), op.unpack(q, 2).cpu().numpy()


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_paired_mixed_matches_unpaired(cuda, degree):
    P = problems.vortex(degree=degree, refine=1, refine_fraction=0.3, seed=7, periodic=False,
                        u_inf=(6.0, 0.0), radius=0.25)
    print("[paired] mixed hanging+bcs degree=%d" % degree, flush=True)
    ph, pq = _steps(P, "mixed", 1)
    uh, uq = _steps(P, "mixed", 0)
    fh, fq = _steps(P, "fp64", 0)
    e = {"paired-unpaired zeta": np.linalg.norm(ph - uh) / np.linalg.norm(uh),
         "paired-unpaired q": np.linalg.norm(pq - uq) / np.linalg.norm(uq),
         "paired-fp64 zeta": np.linalg.norm(ph - fh) / np.linalg.norm(fh),
         "paired-fp64 q": np.linalg.norm(pq - fq) / np.linalg.norm(fq)}
    print("  " + " ".join("%s=%.2e" % kv for kv in e.items()))
    assert max(e.values()) <= 1e-6


def test_paired_fp64_matches_unpaired_without_hanging_faces(cuda):
    """Without hanging faces every DoF receives its face integrals in the same order as the
    unpaired kernel (a DoF on one face: cell + that face; a corner: cell + x-face + y-face),
    and the flux formulas are the same functions: the two kernels agree to rounding."""
    P = problems.vortex(degree=3, refine=1)
    ph, pq = _steps(P, "fp64", 1)
    uh, uq = _steps(P, "fp64", 0)
    eh = np.linalg.norm(ph - uh) / np.linalg.norm(uh)
    eq = np.linalg.norm(pq - uq) / np.linalg.norm(uq)
    print("  fp64 paired vs unpaired: zeta %.2e q %.2e" % (eh, eq))
    assert eh <= 1e-14 and eq <= 1e-14
