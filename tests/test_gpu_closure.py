"""The closure network's BF16 matrix-core kernels against the CPU oracle
(oracle/mlp_oracle.py), and the network inside the DG stage."""
import os

import numpy as np
import pytest
import torch

import mlp_oracle as mo
from helpers import rel_l2, to_dev, to_host
from oceano_amd import closure, problems
from oceano_amd.mlp import ClosureNetwork

pytestmark = pytest.mark.gpu


def _net(n_in, n_out, seed, norm=True):
    net = ClosureNetwork(n_in, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    if norm:
        net.set_normalization(rng.standard_normal(n_in) * 0.2, 0.5 + rng.random(n_in),
                              rng.standard_normal(n_out) * 0.1, 0.5 + rng.random(n_out))
    return net, rng


def _oracle(net, quantize=True):
    return mo.Network(net.get_params().astype(np.float64), net.n_in, net.n_out,
                      norm=net.normalization, quantize=quantize)


@pytest.mark.parametrize("n", [1, 37, 1000])
def test_forward_matches_oracle(cuda, n):
    net, rng = _net(30, 18, 5)
    print("[closure] forward n=%d" % n, flush=True)
    x = rng.standard_normal((n, 30)).astype(np.float32)
    y = net.forward(torch.tensor(x, device=cuda)).cpu().numpy()
    ref = _oracle(net).predict(x.astype(np.float64))
    e = rel_l2(y, ref)
    print("  rel L2 vs oracle %.2e" % e)
    assert e < 2e-3
    # and the BF16 network stays within BF16 accuracy of the FP64 one
    assert rel_l2(y, _oracle(net, quantize=False).predict(x.astype(np.float64))) < 3e-2


@pytest.mark.parametrize("n", [16, 333, 9000])
def test_gradients_match_oracle(cuda, n):
    """loss and every parameter gradient (n = 9000 spreads over many blocks)"""
    net, rng = _net(30, 18, 7)
    print("[closure] gradients n=%d" % n, flush=True)
    x = rng.standard_normal((n, 30)).astype(np.float32)
    t = (rng.standard_normal((n, 18)) * 0.3).astype(np.float32)
    loss, g = net.gradients(torch.tensor(x, device=cuda), torch.tensor(t, device=cuda))
    lref, gref = _oracle(net).loss_and_gradient(x.astype(np.float64), t.astype(np.float64))
    print("  loss %.8e vs %.8e, gradient rel L2 %.2e" % (loss, lref, rel_l2(g, gref)))
    assert abs(loss - lref) < 1e-4 * lref
    assert rel_l2(g, gref) < 5e-3
    for a, b in ((mo.W1, mo.B1), (mo.W2, mo.B2), (mo.W3, mo.B3), (mo.W4, mo.B4),
                 (mo.B1, mo.W2), (mo.B4, mo.NPARAM)):
        assert rel_l2(g[a:b], gref[a:b]) < 1e-2, (a, b)


def test_adam_step_matches_oracle(cuda):
    net, rng = _net(30, 18, 9)
    x = torch.tensor(rng.standard_normal((500, 30)), dtype=torch.float32, device=cuda)
    t = torch.tensor(rng.standard_normal((500, 18)), dtype=torch.float32, device=cuda)
    print("[closure] adam", flush=True)
    p0 = net.get_params()
    m = np.zeros_like(p0)
    v = np.zeros_like(p0)
    for step in (1, 2):
        _, g = net.gradients(x, t)
        net.adam(1e-3)
        torch.cuda.synchronize()
        p0, m, v = mo.adam(p0, m, v, g, step, 1e-3)
        p1 = net.get_params()
        assert np.abs(p1 - p0).max() < 1e-6
    # the BF16 copies follow the master weights: forward equals the oracle on the new ones
    y = net.forward(x[:64]).cpu().numpy()
    assert rel_l2(y, _oracle(net).predict(x[:64].cpu().numpy().astype(np.float64))) < 2e-3


def test_training_reduces_loss(cuda):
    """a smooth synthetic regression: 400 Adam steps cut the loss by more than 20x"""
    net, rng = _net(30, 18, 11, norm=False)
    n = 8192
    x = rng.standard_normal((n, 30)).astype(np.float32)
    A = rng.standard_normal((30, 18)) / np.sqrt(30)
    t = np.tanh(x @ A).astype(np.float32)
    X, T = torch.tensor(x, device=cuda), torch.tensor(t, device=cuda)
    print("[closure] training", flush=True)
    l0 = net.train_step(X, T, 2e-3, want_loss=True)
    g = torch.Generator(device="cpu").manual_seed(0)
    for it in range(400):
        idx = torch.randint(0, n, (1024,), generator=g).to(cuda)
        net.train_step(X[idx].contiguous(), T[idx].contiguous(), 2e-3)
    l1, _ = net.gradients(X, T)
    print("  loss %.4e -> %.4e" % (l0, l1))
    assert l1 < l0 / 20


def _channel_with_stats(seed=1003):
    S = closure.SubgridChannel(seed, nx=50, ny=2, factor=2)
    op, data, h, q = problems.setup_operator(S.plain)
    op.set_subgrid_stats(S.cell_stats(S.plain.mesh))
    return S, op, h, q


def test_features_match_host(cuda):
    S, op, h, q = _channel_with_stats()
    print("[closure] features cells=%d" % op.n_cells, flush=True)
    X = op.closure_features(h, q).cpu().numpy()
    Vi, _ = closure.nodal_to_modal(2)
    zh = to_host(op, h, 1).reshape(-1, 9)
    zq = to_host(op, q, 2).reshape(-1, 2, 9)
    ref = np.concatenate([zh @ Vi.T, zq[:, 0] @ Vi.T, zq[:, 1] @ Vi.T,
                          S.cell_stats(S.plain.mesh)], 1)
    assert X.shape == ref.shape
    assert np.abs(X - ref).max() <= 1e-6 * np.abs(ref).max()


def test_closure_in_stage(cuda):
    """with a network attached the stage's k_q gains exactly the correction of the stage
    input; the correction equals the oracle network on the host features; a network with
    non-finite output leaves the plain model"""
    S, op, h, q = _channel_with_stats()
    print("[closure] in-stage cells=%d" % op.n_cells, flush=True)
    net, rng = _net(30, 18, 13)
    op.set_closure(net)
    tau = op.closure_source(h, q)
    X = op.closure_features(h, q).cpu().numpy().astype(np.float64)
    _, V = closure.nodal_to_modal(2)
    tm = _oracle(net).predict(X)
    ref = np.stack([tm[:, :9] @ V.T, tm[:, 9:] @ V.T], 1).ravel()
    e = rel_l2(to_host(op, tau, 2), ref)
    print("  correction vs oracle network: rel L2 %.2e" % e)
    assert e < 2e-3
    dt = 1e-3
    factor = [dt]

    def one_stage(with_net):
        op.set_closure(net if with_net else None)
        ki_h = [op.initialize_vector(0) for _ in range(3)]
        ki_q = [op.initialize_vector(1) for _ in range(3)]
        r_h, r_q = op.initialize_vector(0), op.initialize_vector(1)
        wh, wq = h.clone(), q.clone()
        op.perform_stage_hydro(0, 0.0, factor, [h, q, None, None], ki_h, ki_q, wh, wq, r_h,
                               r_q)
        torch.cuda.synchronize()
        return ki_q[1], r_q
    k0, r0 = one_stage(False)
    k1, r1 = one_stage(True)
    scale = float(k0.abs().max()) + float(tau.abs().max())
    assert torch.allclose(k1 - k0, tau, rtol=0, atol=1e-14 * scale)
    assert torch.allclose(r1 - r0, dt * tau, rtol=0, atol=1e-14 * (scale + float(q.abs().max())))
    # fallback: NaN weights give NaN outputs, hence no correction
    p = net.get_params()
    p[mo.B4] = np.nan
    net.set_params(p)
    k2, _ = one_stage(True)
    assert torch.equal(k2, k0)
    op.set_closure(None)


TRAINED = os.path.join(os.path.dirname(closure.__file__), "data", "closure_channel_p2.npz")


def test_trained_closure_gain_and_fallback(cuda):
    """The network trained by `python -m oceano_amd.closure` (256 seeded bathymetries,
    fine runs of the solver; oceano_amd/data/closure_channel_p2.npz) on a held-out
    bathymetry.  A priori it removes part of the plain sub-grid model's momentum-tendency
    error on unseen data; a posteriori the coarse-run errors against the restricted fine
    solution are printed for the plain model, plain + network, the sampled bed and the
    exact-correction ceiling (on this channel the coupled runs are very sensitive to
    source errors, and even the exact correction gains well under 2x).  Detaching the
    network gives the plain model bitwise."""
    net = ClosureNetwork.load(TRAINED)
    ev = closure.Sample(2000, n_snap=2)
    print("[closure] trained closure on held-out seed 2000: %d active cells"
          % int(ev.active.sum()), flush=True)
    gain = net.output_gain
    net.set_output_gain(1.0)  # a priori: the network's own output
    loss, _ = net.gradients(ev.X, ev.T)
    net.set_output_gain(gain)
    scale = torch.tensor(net.normalization[3], device=cuda, dtype=torch.float32)
    zero = float(((ev.T / scale) ** 2).mean())
    print("  held-out a-priori loss %.3e, no correction %.3e" % (loss, zero))
    assert loss < zero
    r = closure.evaluate(ev, net, gains=(1.0,))
    for k, v in r.items():
        print("  %-10s zeta %.3e hu %.3e" % (k, v[0], v[1]))
    assert all(np.isfinite(v).all() for v in r.values())
    # the exact correction of the initial state improves on the plain model
    assert r["exact"][0] < r["plain"][0]
    # a posteriori, with the output gain calibrated on coupled runs of training
    # bathymetries (closure.calibrate_gain), the network does not degrade the plain model
    # on a held-out one (gain 0 would be the plain model itself)
    print("  calibrated output gain %g" % net.output_gain)
    assert r["learned"][0] <= r["plain"][0] * (1 + 1e-9)
    assert r["learned"][1] <= r["plain"][1] * (1 + 1e-9)
    # fallback: with the network detached the stage is the plain model
    S = ev.S
    op, _, h, q = problems.setup_operator(S.plain)
    op.set_subgrid_stats(S.cell_stats(S.plain.mesh), ev.active)
    a_h, a_q = h.clone(), q.clone()
    op.set_closure(net)
    op.set_closure(None)
    from oceano_amd.integrator import DeviceIntegrator
    DeviceIntegrator(op).step(1e-3, a_h, a_q)
    op2, _, b_h, b_q = problems.setup_operator(S.plain)
    DeviceIntegrator(op2).step(1e-3, b_h, b_q)
    torch.cuda.synchronize()
    assert torch.equal(a_h, b_h) and torch.equal(a_q, b_q)
