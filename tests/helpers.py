"""Shared set-up for parity tests: one problem, the GPU operator and the CPU oracle fed
with identical static data and initial state."""
import numpy as np

import pyoracle
from oceano_amd import problems


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def make_oracle(problem):
    o = pyoracle.Oracle.from_mesh(problem.mesh, problem.degree, bcs=problem.bcs, g=problem.g,
                                  viscosity=problem.viscosity)
    data = o.initialize_data_at_quadrature(problem.problem_data)
    h, q = o.project_hydro(problem.initial, data["cell_2"][:, 0])
    return o, data, h, q


def make_pair(problem, precision="fp64"):
    """GPU operator and oracle on the same problem; returns (op, oracle, h0, q0) with the
    initial state as cell-major numpy arrays (identical for both)."""
    op, data, h_dev, q_dev = problems.setup_operator(problem, precision=precision)
    o, odata, h, q = make_oracle(problem)
    if problem.boundary_values is not None and op.n_boundary:
        o.set_boundary_values(problem.boundary_values(op, 0.0))
    # same initial state on both sides: take the oracle's projection, pack to the GPU
    return op, o, h, q


def to_dev(op, x, ncomp):
    return op.pack(x, ncomp)


def to_host(op, v, ncomp):
    return op.unpack(v, ncomp).cpu().numpy()
