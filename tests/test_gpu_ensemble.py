"""C4 batching: a degree's instances of the ensemble run as one operator over the disjoint
union of their meshes (one launch per stage for all of them).  Every cell's arithmetic is
independent of the other instances, so two RK4 steps of the union reproduce the separate
runs of each instance bit for bit, in FP64 and in mixed precision."""
import numpy as np
import pytest
import torch

from oceano_amd import problems
from oceano_amd.integrator import DeviceIntegrator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_union_equals_separate_instances(cuda, precision):
    groups, inst = problems.ensemble(n_instances=8, seed=3, n_min=200, n_max=3000)
    dt = 1e-4
    for p, G in groups.items():
        print("[ensemble] %s p=%d union of %d instances, %d cells"
              % (precision, p, len(G.mesh.starts) - 1, G.mesh.n_cells), flush=True)
        op, data, w_h, w_q = problems.setup_operator(G, precision=precision)
        rk = DeviceIntegrator(op)
        for _ in range(2):
            rk.step(dt, w_h, w_q)
        uh = op.unpack(w_h, 1).cpu().numpy().reshape(G.mesh.n_cells, -1)
        uq = op.unpack(w_q, 2).cpu().numpy().reshape(G.mesh.n_cells, -1)
        members = [d for d in inst if d["degree"] == p]
        for i, d in enumerate(members):
            P = problems.vortex(degree=p, refine=0, nx=d["nx"], ny=d["ny"], center=d["center"],
                                radius=d["radius"], amplitude=d["amplitude"])
            op1, _, h1, q1 = problems.setup_operator(P, precision=precision)
            rk1 = DeviceIntegrator(op1)
            for _ in range(2):
                rk1.step(dt, h1, q1)
            a, b = G.mesh.starts[i], G.mesh.starts[i + 1]
            sh = op1.unpack(h1, 1).cpu().numpy().reshape(b - a, -1)
            sq = op1.unpack(q1, 2).cpu().numpy().reshape(b - a, -1)
            assert np.array_equal(uh[a:b], sh) and np.array_equal(uq[a:b], sq), (p, i)


def test_full_c4_union(cuda):
    """the benchmark's full ensemble (64 instances, seed 11, 1k-50k cells each, ~734k
    cells): per degree, the union's smallest and largest members reproduce their separate
    runs bit for bit in both precisions, and the mixed union is within 1e-6 of FP64"""
    groups, inst = problems.ensemble()
    dt = 5e-5
    for p, G in groups.items():
        members = [d for d in inst if d["degree"] == p]
        sizes = [d["nx"] * d["ny"] for d in members]
        pick = sorted({int(np.argmin(sizes)), int(np.argmax(sizes))})
        out = {}
        for precision in ("fp64", "mixed"):
            print("[ensemble] full C4 %s p=%d: %d instances, %d cells"
                  % (precision, p, len(members), G.mesh.n_cells), flush=True)
            op, data, w_h, w_q = problems.setup_operator(G, precision=precision)
            rk = DeviceIntegrator(op)
            for _ in range(2):
                rk.step(dt, w_h, w_q)
            uh = op.unpack(w_h, 1).cpu().numpy().reshape(G.mesh.n_cells, -1)
            uq = op.unpack(w_q, 2).cpu().numpy().reshape(G.mesh.n_cells, -1)
            out[precision] = (uh, uq)
            del op, w_h, w_q, rk
            for i in pick:
                d = members[i]
                P = problems.vortex(degree=p, refine=0, nx=d["nx"], ny=d["ny"],
                                    center=d["center"], radius=d["radius"],
                                    amplitude=d["amplitude"])
                op1, _, h1, q1 = problems.setup_operator(P, precision=precision)
                rk1 = DeviceIntegrator(op1)
                for _ in range(2):
                    rk1.step(dt, h1, q1)
                a, b = G.mesh.starts[i], G.mesh.starts[i + 1]
                sh = op1.unpack(h1, 1).cpu().numpy().reshape(b - a, -1)
                sq = op1.unpack(q1, 2).cpu().numpy().reshape(b - a, -1)
                assert np.array_equal(uh[a:b], sh) and np.array_equal(uq[a:b], sq), \
                    (precision, p, i)
        (fh, fq), (mh, mq) = out["fp64"], out["mixed"]
        eh = np.linalg.norm(mh - fh) / np.linalg.norm(fh)
        eq = np.linalg.norm(mq - fq) / np.linalg.norm(fq)
        print("  p=%d mixed vs fp64 over the union: zeta %.2e q %.2e" % (p, eh, eq))
        assert eh < 1e-6 and eq < 1e-6

@pytest.mark.parametrize("precision,degrees", [("fp64", (1, 2, 4)), ("mixed", (1, 2, 4)),
                                              ("fp64", (1, 2)), ("mixed", (1, 2))])
def test_single_launch_equals_per_degree_launches(cuda, precision, degrees):
    """All degree groups in one launch per stage (oc_rk_*_multi, the kernel k_stage_multi):
    eager steps reproduce the per-degree operators' own steps to rounding, and a graph
    replay of the same steps reproduces the eager multi-operator steps bit for bit.

    The multi-operator kernel runs the same device functions on the same inputs, but it is
    a different compiled kernel: the backend pairs multiplies with adds into FMAs per
    basic block, and the pairing it picks can differ between the two kernels (measured:
    identical at degrees 1 and 2; at degree 4, 1-ulp differences in about 4% of the FP64
    values after one step, and 4.7e-8 of the largest value, under one FP32 unit in the last
    place, after three mixed-precision steps), so the comparison with the per-degree kernels
    is relative to the largest value, at 1e-13 in FP64 and 5e-7 (8 FP32 units in the last
    place) in mixed precision.

    With the p = 4 group the operators do not all pair their items (paired_items defaults to
    p <= 3), so the launch is k_stage_multi; groups of p <= 3 only take k_stage_multi_pair."""
    from oceano_amd.integrator import EnsembleIntegrator
    groups, inst = problems.ensemble(n_instances=8, seed=3, n_min=200, n_max=3000)
    assert all(p in groups for p in degrees)
    dt = 1e-4
    ops, hs, qs, ref = [], [], [], []
    for p in degrees:
        G = groups[p]
        op, data, w_h, w_q = problems.setup_operator(G, precision=precision)
        r_h, r_q = w_h.clone(), w_q.clone()
        rk = DeviceIntegrator(op)
        for _ in range(3):
            rk.step(dt, r_h, r_q)
        ops.append(op)
        hs.append(w_h)
        qs.append(w_q)
        ref.append((r_h, r_q))
    print("[ensemble] %s single launch over %d degree groups" % (precision, len(ops)),
          flush=True)
    multi = EnsembleIntegrator(ops)
    e_h = [h.clone() for h in hs]
    e_q = [q.clone() for q in qs]
    for _ in range(3):
        multi.step(dt, e_h, e_q)
    torch.cuda.synchronize()
    tol = 1e-13 if precision == "fp64" else 5e-7
    for g in range(len(ops)):
        for x, y in ((e_h[g], ref[g][0]), (e_q[g], ref[g][1])):
            rel = float((x - y).abs().max() / y.abs().max())
            print("  group %d: max difference / max value %.2e" % (g, rel), flush=True)
            assert rel <= tol, (g, rel)
    # the same three steps captured into one graph and replayed once
    g_h = [h.clone() for h in hs]
    g_q = [q.clone() for q in qs]
    multi.set_time(0.0)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        multi.capture(dt, g_h, g_q, n_steps=3)
        multi.replay(1)
    side.synchronize()
    for g in range(len(ops)):
        assert torch.equal(g_h[g], e_h[g]) and torch.equal(g_q[g], e_q[g]), g


@pytest.mark.parametrize("degree", [2, 3, 4])
@pytest.mark.parametrize("wpb", [2, 4])
def test_waves_per_block_does_not_change_results(cuda, wpb, degree):
    """64-cell work items per workgroup of the one-pass stage (library option
    waves_per_block; each batch's height and discharge items on LDS slices of their own
    sizes): the same steps bit for bit"""
    from oceano_amd import tune
    print("[wpb] degree=%d waves_per_block=%d" % (degree, wpb), flush=True)
    P = problems.vortex(degree=degree, refine=1, refine_fraction=0.3, seed=5)
    out = []
    try:
        for w in (1, wpb):
            # unpaired items: with paired_items a workgroup is always one batch's two items
            tune.apply_options({"waves_per_block": w, "paired_items": 0})
            for precision in ("fp64", "mixed"):
                op, _, h, q = problems.setup_operator(P, precision=precision)
                rk = DeviceIntegrator(op)
                for _ in range(2):
                    rk.step(1e-4, h, q)
                out.append((h, q))
    finally:
        tune.apply_options(tune.LIBRARY_DEFAULTS)
    for a, b in zip(out[:2], out[2:]):
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])
