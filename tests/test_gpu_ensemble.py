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
