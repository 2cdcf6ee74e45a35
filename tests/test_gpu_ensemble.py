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
