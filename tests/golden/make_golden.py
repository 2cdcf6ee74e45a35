"""Writes the golden stage vectors of tests/golden/stage_vectors.npz from the CPU oracle.

The reference ships no golden vectors for this path, so these fixtures come from the
oracle (oracle/oracle.cpp), which tests/test_oracle_pins.py pins against closed forms and
invariants.  They freeze its results: test_golden.py checks that the oracle still
reproduces them and that the GPU stage (FP64) matches them to 1e-12.

Cases (small, every degree): the travelling vortex on a periodic 8x4 mesh with 30% of the
cells refined (hanging faces), and the channel with discharge inflow, height outflow,
walls, Manning friction and a bumpy bed on 20x2 cells.  Each case stores the mesh, the
static data, the perturbed state, L(state) and one classic-RK4 step.

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from oceano_amd import problems  # noqa: E402
from oceano_amd.integrator import butcher_rows  # noqa: E402
from oceano_amd.mesh import Mesh  # noqa: E402


class _Points:
    """what Problem.boundary_values reads from an operator, served by the oracle"""

    def __init__(self, o):
        self.quadrature_points = o.quadrature_points
        self.n_boundary = o.boundary.shape[0]
        self.n_q_points_1d = o.nq


def cases():
    for p in (1, 2, 3, 4):
        m = Mesh.rectangle(0.0, 2.0, 0.0, 1.0, 8, 4, periodic=(True, True),
                           boundary_ids=(1, 2, 0, 0))
        m.refine_fraction(0.3, 20261202)
        P = problems.vortex(degree=p, refine=0)
        P.mesh = m
        yield "vortex_p%d" % p, P, 2e-3
        yield "channel_p%d" % p, problems.channel(degree=p, nx=20, ny=2), 2e-3


def build():
    out = {}
    for name, P, dt in cases():
        o = pyoracle.Oracle.from_mesh(P.mesh, P.degree, bcs=P.bcs, g=P.g)
        data = o.initialize_data_at_quadrature(P.problem_data)
        bvals = None
        if P.boundary_values is not None and P.mesh.boundary.shape[0]:
            bvals = P.boundary_values(_Points(o), 0.0)
            o.set_boundary_values(bvals)
        h, q = o.project_hydro(P.initial, data["cell_2"][:, 0])
        rng = np.random.default_rng(P.degree + 100 * (name[0] == "c"))
        h = h + 1e-3 * rng.standard_normal(h.shape)
        q = q + 1e-2 * rng.standard_normal(q.shape)
        Lh, Lq = o.rhs(h, q)
        rows, _ = butcher_rows("stage_4_order_4", dt)
        S = len(rows)
        w_h, w_q = h.copy(), q.copy()
        ki_h = [o.zeros(1) for _ in range(S + 1)]
        ki_q = [o.zeros(2) for _ in range(S + 1)]
        o.time_step(rows, w_h, w_q, ki_h, ki_q, o.zeros(1), o.zeros(2))
        d = {"degree": P.degree, "dt": dt, "bc_ids": np.array(list(P.bcs.keys()), np.int32),
             "bc_kinds": np.array([pyoracle.BC[k] for k in P.bcs.values()], np.int32),
             "geom": P.mesh.geom, "inner": P.mesh.inner,
             "boundary": P.mesh.boundary, "h": h, "q": q, "Lh": Lh, "Lq": Lq, "w_h": w_h,
             "w_q": w_q}
        for k, v in data.items():
            d["data_" + k] = v
        if bvals is not None:
            d["bvals"] = bvals
        for k, v in d.items():
            out[name + "/" + k] = np.asarray(v)
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "stage_vectors.npz"), **build())
    print("wrote", os.path.join(HERE, "stage_vectors.npz"))
