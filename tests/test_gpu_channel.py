"""C3: the channel (discharge inflow Q = 4.42, height outflow zeta = 2, walls, Manning
n = 0.02, SWASHES bump plus seeded Gaussian bumps) driven towards its steady state by the
device-resident loop, in FP64 and in mixed precision: the step-to-step change falls, the
discharge stays at the inflow value along the channel, and the two precisions agree."""
import numpy as np
import pytest
import torch

from oceano_amd import problems
from oceano_amd.integrator import DeviceIntegrator

pytestmark = pytest.mark.gpu


def _steady(precision, steps):
    C = problems.channel(degree=2, nx=50, ny=2)
    op, data, w_h, w_q = problems.setup_operator(C, precision=precision)
    speed = op.transport_speed(w_h, w_q)
    dt = 0.075 / float("%.3g" % speed)  # main.cc:1549-1555, CFL of the channel deck
    rk = DeviceIntegrator(op)
    n0, change0 = rk.run_to_steady(w_h, w_q, dt, tol=1e-10, max_steps=250, chunk=250)
    n, change = rk.run_to_steady(w_h, w_q, dt, tol=1e-10, max_steps=steps, chunk=250)
    return op, w_h, w_q, n0 + n, (change0, change), dt


def test_channel_towards_steady_state(cuda):
    print("[channel] fp64", flush=True)
    op, h64, q64, n, ch64, dt = _steady("fp64", 6000)
    print("  fp64: %d steps of %.3e s, relative change per step %.3e -> %.3e"
          % (n, dt, ch64[0], ch64[1]))
    print("[channel] mixed", flush=True)
    op32, h32, q32, n32, ch32, _ = _steady("mixed", 6000)
    print("  mixed: %d steps, relative change per step %.3e -> %.3e" % (n32, ch32[0], ch32[1]))
    assert np.isfinite(ch64[1]) and np.isfinite(ch32[1])
    # weak Manning damping (rate ~ g n^2 |u| / h^(4/3) ~ 3e-3 /s): 35 s of flow take the
    # gravity-wave transient down by an order of magnitude
    assert ch64[1] < 0.1 * ch64[0] and ch32[1] < 0.1 * ch32[0]
    # mean x-discharge ~ Q / width everywhere once the transient has left
    qx = op.unpack(q64, 2).cpu().numpy().reshape(-1, 2, op.dofs_per_cell)[:, 0]
    assert abs(qx.mean() - 4.42) < 0.05 * 4.42
    dh = (op.unpack(h64, 1) - op32.unpack(h32, 1)).norm() / op.unpack(h64, 1).norm()
    dq = (op.unpack(q64, 2) - op32.unpack(q32, 2)).norm() / op.unpack(q64, 2).norm()
    print("  mixed vs fp64: zeta %.2e q %.2e" % (float(dh), float(dq)))
    assert float(dh) < 1e-5 and float(dq) < 1e-4
