"""The device-resident time loop (velocity update, transport speed, mass check, the
library's RK driver and its HIP-graph replay) against the CPU oracle and the eager path.

OceanDG.h:2506-2662 (transport speed, velocity field), 2138-2184 (mass check),
ExplicitRungeKuttaIntegrator.h:179-298 with main.cc:1546-1577 (time loop).
"""
import numpy as np
import pytest
import torch

from helpers import make_pair, rel_l2, to_dev, to_host
from oceano_amd import problems
from oceano_amd.integrator import DeviceIntegrator, butcher_rows

pytestmark = pytest.mark.gpu


def _bumpy_vortex(degree):
    """hanging-node vortex over a smooth bed (z_b enters every diagnostic)"""
    P = problems.vortex(degree=degree, refine=1, refine_fraction=0.3, seed=5)
    base = P.problem_data

    def data(xy):
        d = base(xy)
        d[:, 0] = 0.05 * np.sin(np.pi * xy[:, 0]) * np.cos(np.pi * xy[:, 1])
        return d
    P.problem_data = data
    return P


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_velocity_speed_volume_match_oracle(cuda, degree):
    P = _bumpy_vortex(degree)
    print("[timeloop] diagnostics p=%d cells=%d" % (degree, P.mesh.n_cells), flush=True)
    op, o, h, q = make_pair(P)
    rng = np.random.default_rng(degree)
    h = h + 1e-3 * rng.standard_normal(h.shape)
    q = q + 1e-2 * rng.standard_normal(q.shape)
    dh, dq = to_dev(op, h, 1), to_dev(op, q, 2)
    vel = op.velocity(dh, dq)
    e = rel_l2(to_host(op, vel, 2), o.velocity_field(h, q))
    s_gpu, s_ref = op.transport_speed(dh, dq), o.transport_speed(h, q)
    v_gpu, v_ref = op.volume(dh), o.volume(h)
    print("  velocity rel L2 %.2e, speed %.15e vs %.15e, volume %.15e vs %.15e"
          % (e, s_gpu, s_ref, v_gpu, v_ref))
    assert e < 1e-13
    assert abs(s_gpu - s_ref) < 1e-13 * s_ref
    assert abs(v_gpu - v_ref) < 1e-13 * abs(v_ref)


def _python_steps(op, rows, w_h, w_q, n):
    S = len(rows)
    ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
    ki_q = [op.initialize_vector(1) for _ in range(S + 1)]
    r_h = [op.initialize_vector(0) for _ in range(2)]
    r_q = [op.initialize_vector(1) for _ in range(2)]
    for _ in range(n):
        for s in range(S):
            ri = [w_h, w_q] if s == 0 else [r_h[s % 2], r_q[s % 2]]
            op.perform_stage_hydro(s, 0.0, rows[s][:s + 1], ri + [None, None], ki_h, ki_q,
                                   w_h, w_q, r_h[(s + 1) % 2], r_q[(s + 1) % 2])
    return r_h, r_q


def _device_to_host(ptr, n_doubles):
    """n_doubles of device memory at ptr (a library-owned buffer) into a numpy array"""
    import ctypes as C
    import glob
    import os
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:  # the HIP runtime torch ships with
        hip = C.CDLL(glob.glob(os.path.join(os.path.dirname(torch.__file__), "lib",
                                            "libamdhip64.so*"))[0])
    hip.hipMemcpy.restype = C.c_int
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.empty(n_doubles, np.float64)
    torch.cuda.synchronize()
    assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
    return out


def test_mixed_loop_registers_hold_centred_fp32_stage_inputs(cuda):
    """The layout oc_rk_registers documents for a mixed operator with paired work items:
    after one step, the r registers that fed the last two stages decode (floats of the
    first half, plus each cell's FP64 centre after them for zeta) to those stage inputs of
    the stage-by-stage loop over FP64 vectors, within FP32 rounding of the centred values."""
    import ctypes as C
    from oceano_amd import _capi
    P = problems.vortex(degree=3, refine=1, refine_fraction=0.3, seed=9)
    op, _, h, q = make_pair(P, precision="mixed")
    dt = 2e-4
    rows, _ = butcher_rows("stage_4_order_4", dt)
    S = len(rows)
    r_h, r_q = _python_steps(op, rows, to_dev(op, h, 1), to_dev(op, q, 2), 1)
    rk = DeviceIntegrator(op, "stage_4_order_4")
    b_h, b_q = to_dev(op, h, 1), to_dev(op, q, 2)
    rk.step(dt, b_h, b_q)
    kh, kq = (C.c_void_p * (S + 1))(), (C.c_void_p * (S + 1))()
    rh, rq = (C.c_void_p * 2)(), (C.c_void_p * 2)()
    _capi.check(_capi.lib().oc_rk_registers(rk._h, kh, kq, rh, rq), "oc_rk_registers")
    nd, NB = op.dofs_per_cell, op.n_cells_padded // 64
    live = (np.arange(NB * 64) < op.n_cells).reshape(NB, 64)
    for j in range(2):
        bh = _device_to_host(rh[j], NB * nd * 64)
        bq = _device_to_host(rq[j], NB * 2 * nd * 64)
        vals = bh.view(np.float32)[:NB * nd * 64].reshape(NB, nd, 64).astype(np.float64)
        cen = bh[NB * nd * 64 // 2:NB * nd * 64 // 2 + NB * 64].reshape(NB, 64)
        ref = r_h[j].cpu().numpy().reshape(NB, nd, 64)
        # float(zeta - centre) to half an FP32 ulp, the sum back in FP64 to half an FP64 ulp
        dz = np.abs(vals + cen[:, None, :] - ref)
        bound = 2.0 ** -24 * np.abs(ref - cen[:, None, :]) + 2.0 ** -52 * np.abs(ref) + 1e-300
        mz = np.broadcast_to(live[:, None, :], dz.shape)
        vq = bq.view(np.float32)[:NB * 2 * nd * 64].reshape(NB, 2, nd, 64).astype(np.float64)
        refq = r_q[j].cpu().numpy().reshape(NB, 2, nd, 64)
        dq = np.abs(vq - refq)
        mq = np.broadcast_to(live[:, None, None, :], dq.shape)
        print("  register %d: max |zeta err| / bound %.3f, max |q err| / bound %.3f"
              % (j, (dz / bound)[mz].max(), (dq / (2.0 ** -24 * np.abs(refq) + 1e-300))[mq].max()))
        assert np.all((dz <= bound)[mz])
        assert np.all((dq <= 2.0 ** -24 * np.abs(refq) + 1e-300)[mq])


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("degree", [1, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_integrator_and_graph_replay(cuda, precision, degree, periodic):
    """oc_rk_step equals the stage-by-stage loop bitwise (in mixed precision the loop's
    intermediate stage inputs travel in their centred FP32 form, the stage-by-stage loop's
    in FP64); a captured step replayed three times, with eager work and synchronisations
    between the replays, equals three eager steps bitwise after every replay; and the FP64
    result matches the oracle's RK4."""
    P = problems.vortex(degree=degree, refine=1, refine_fraction=0.3, seed=9,
                        periodic=periodic)
    print("[timeloop] integrator %s p=%d periodic=%d cells=%d"
          % (precision, degree, periodic, P.mesh.n_cells), flush=True)
    op, o, h, q = make_pair(P, precision=precision)
    dt = 2e-4
    rows, _ = butcher_rows("stage_4_order_4", dt)
    a_h, a_q = to_dev(op, h, 1), to_dev(op, q, 2)
    _python_steps(op, rows, a_h, a_q, 1)
    rk = DeviceIntegrator(op, "stage_4_order_4")
    b_h, b_q = to_dev(op, h, 1), to_dev(op, q, 2)
    rk.step(dt, b_h, b_q)
    torch.cuda.synchronize()
    assert torch.equal(a_h, b_h) and torch.equal(a_q, b_q)
    # graph: captured on a side stream, state vectors fixed at capture
    g_h, g_q = to_dev(op, h, 1), to_dev(op, q, 2)
    e_h, e_q = g_h.clone(), g_q.clone()
    rk_e = DeviceIntegrator(op, "stage_4_order_4")
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        rk.capture(dt, g_h, g_q)
    side.synchronize()
    assert torch.equal(g_h, to_dev(op, h, 1)), "capture must not execute the step"
    for rep in range(3):
        with torch.cuda.stream(side):
            rk.replay(1)
        side.synchronize()
        rk_e.step(dt, e_h, e_q)
        _ = (e_h * 2.0).sum().item()  # unrelated eager work between replays
        torch.cuda.synchronize()
        print("  replay %d: max|diff| %.3e" % (rep, float((g_h - e_h).abs().max())))
        assert torch.equal(g_h, e_h) and torch.equal(g_q, e_q)
    if precision == "fp64":
        S = len(rows)
        w_h, w_q = h.copy(), q.copy()
        ki_h = [o.zeros(1) for _ in range(S + 1)]
        ki_q = [o.zeros(2) for _ in range(S + 1)]
        r_h, r_q = o.zeros(1), o.zeros(2)
        for _ in range(3):
            o.time_step(rows, w_h, w_q, ki_h, ki_q, r_h, r_q)
        eh, eq = rel_l2(to_host(op, g_h, 1), w_h), rel_l2(to_host(op, g_q, 2), w_q)
        print("  3 replayed steps vs oracle: zeta %.2e q %.2e" % (eh, eq))
        assert eh < 1e-12 and eq < 1e-12


def test_viscous_step_updates_velocity(cuda):
    """With viscosity the step ends with the velocity update of the new state
    (main.cc:1573-1575); the stage reads the previous velocity."""
    P = problems.vortex(degree=2, refine=1)
    P.viscosity = 1e-3
    print("[timeloop] viscous cells=%d" % P.mesh.n_cells, flush=True)
    op, o, h, q = make_pair(P)
    w_h, w_q = to_dev(op, h, 1), to_dev(op, q, 2)
    vel = op.velocity(w_h, w_q)
    v0 = to_host(op, vel, 2)
    assert rel_l2(v0, o.velocity_field(h, q)) < 1e-13
    rows, _ = butcher_rows("stage_4_order_4", 1e-4)
    S = len(rows)
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(S + 1)]
    r_h, r_q = o.zeros(1), o.zeros(2)
    oh, oq = h.copy(), q.copy()
    o.time_step(rows, oh, oq, ki_h, ki_q, r_h, r_q, v=v0)
    rk = DeviceIntegrator(op, "stage_4_order_4")
    rk.step(1e-4, w_h, w_q, vel)
    torch.cuda.synchronize()
    assert rel_l2(to_host(op, w_h, 1), oh) < 1e-12
    assert rel_l2(to_host(op, w_q, 2), oq) < 1e-12
    assert rel_l2(to_host(op, vel, 2), o.velocity_field(oh, oq)) < 1e-12


def test_mass_conservation(cuda):
    """check_mass (OceanDG.h:2138-2184): on the periodic vortex the volume stays constant to
    round-off over 20 steps (each lumped Jacobi sweep removes the whole residual mass, so
    the 15-sweep cap does not leak mass); on the channel (discharge inflow, height outflow)
    the volume changes by the time integral of the boundary flux, sum_s b_s dt sum(L_zeta,s)."""
    P = problems.vortex(degree=3, refine=1, refine_fraction=0.3, seed=4)
    print("[timeloop] mass vortex cells=%d" % P.mesh.n_cells, flush=True)
    op, data, w_h, w_q = problems.setup_operator(P)
    rk = DeviceIntegrator(op)
    v0 = op.volume(w_h)
    for _ in range(20):
        rk.step(2e-4, w_h, w_q)
    v1 = op.volume(w_h)
    print("  vortex volume drift %.3e" % ((v1 - v0) / v0))
    assert abs(v1 - v0) < 1e-13 * v0

    C = problems.channel(degree=2, nx=50, ny=4)
    print("[timeloop] mass channel cells=%d" % C.mesh.n_cells, flush=True)
    op, data, w_h, w_q = problems.setup_operator(C)
    dt = 2e-3
    rows, _ = butcher_rows("stage_4_order_4", dt)
    S = len(rows)
    ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
    ki_q = [op.initialize_vector(1) for _ in range(S + 1)]
    r_h = [op.initialize_vector(0) for _ in range(2)]
    r_q = [op.initialize_vector(1) for _ in range(2)]
    v0 = op.volume(w_h)
    flux = 0.0
    for _ in range(10):
        for s in range(S):
            ri = [w_h, w_q] if s == 0 else [r_h[s % 2], r_q[s % 2]]
            op.perform_stage_hydro(s, 0.0, rows[s][:s + 1], ri + [None, None], ki_h, ki_q,
                                   w_h, w_q, r_h[(s + 1) % 2], r_q[(s + 1) % 2])
        flux += sum(rows[-1][s] * op.sum(ki_h[s + 1], 1) for s in range(S))
    v1 = op.volume(w_h)
    print("  channel: volume change %.6e, boundary flux integral %.6e" % (v1 - v0, flux))
    assert abs((v1 - v0) - flux) < 1e-10 * v0
