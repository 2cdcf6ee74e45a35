"""Test-case set-ups (ICBC layer): meshes, problem data, boundary data, initial states.

Vortex (Icbc_ShallowWaterVortex.h): [0,2]x[0,1], 20x10 coarse cells (the fixture
tests/shallowWaterVortex/shallowWaterVortex_20x10cells.msh is this structured grid with
boundary ids 0 = walls y=0,1, 1 = x=0, 2 = x=2), refined globally, optionally with a seeded
fraction refined once more (hanging nodes); periodic, or with the reference's
supercritical inflow / outflow / wall conditions (Icbc_ShallowWaterVortex.h:262-271).
Channel (Icbc_ChannelFlow.h BC kinds): discharge inflow at x=0, height condition at the
outflow, walls at y=0,1, Manning friction, a bump plus seeded sub-cell bumps.
"""
import numpy as np

from . import _capi
from ._capi import check, dptr, lib
from .mesh import Mesh


def vortex_exact(xy, t, prm):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros((xy.shape[0], 3))
    p = np.ascontiguousarray(prm, dtype=np.float64)
    check(lib().oc_icbc_vortex(dptr(p), t, xy.shape[0], dptr(xy), dptr(out)), "oc_icbc_vortex")
    return out


def channel_data(xy, prm, seed):
    xy = np.ascontiguousarray(xy, dtype=np.float64)
    out = np.zeros((xy.shape[0], 5))
    p = np.ascontiguousarray(prm, dtype=np.float64)
    check(lib().oc_icbc_channel(dptr(p), seed, xy.shape[0], dptr(xy), dptr(out)),
          "oc_icbc_channel")
    return out


class Problem:
    """mesh, operator parameters, problem data, boundary data and initial condition."""

    def __init__(self, mesh, degree, bcs, problem_data, initial, boundary_values=None,
                 exact=None, g=9.81, viscosity=0.0, name=""):
        self.mesh = mesh
        self.degree = degree
        self.bcs = bcs
        self.problem_data = problem_data    # xy -> [n,5]
        self.initial = initial              # xy -> [n,3] free surface, u, v
        self.boundary_values = boundary_values  # (op, t) -> [n_bdry, nq, 3]
        self.exact = exact                  # (xy, t) -> [n,3] or None
        self.g = g
        self.viscosity = viscosity
        self.name = name

    def dofs(self):
        return 3 * self.mesh.n_cells * (self.degree + 1) ** 2


def vortex(degree=1, refine=3, refine_fraction=None, seed=20261202, periodic=True,
           u_inf=(1.0, 0.0), amplitude=0.1, radius=0.35, center=(0.5, 0.5), h0=1.0,
           g=9.81, nx=20, ny=10, bcs=None):
    """Travelling vortex; non-periodic: boundary ids 1 (x = 0), 2 (x = 2), 0 (y = 0, 1) with
    the reference's supercritical inflow / outflow / wall conditions unless `bcs` maps the
    ids to other kinds.  Boundary data are the exact solution at the stage time (the
    reference's ExactSolution boundary functions, Icbc_ShallowWaterVortex.h:262-271)."""
    mesh = Mesh.rectangle(0.0, 2.0, 0.0, 1.0, nx, ny, periodic=(periodic, periodic),
                          boundary_ids=(1, 2, 0, 0))
    if refine:
        mesh.refine_global(refine)
    if refine_fraction:
        mesh.refine_fraction(refine_fraction, seed)
    prm = [h0, u_inf[0], u_inf[1], amplitude, radius, center[0], center[1], g,
           2.0 if periodic else 0.0, 1.0 if periodic else 0.0]

    def exact(xy, t):
        return vortex_exact(xy, t, prm)

    def problem_data(xy):
        return np.zeros((xy.shape[0], 5))

    if periodic:
        bcs = {}
    elif bcs is None:
        bcs = {0: "wall", 1: "supercritical_inflow", 2: "supercritical_outflow"}

    def boundary_values(op, t):
        xy = op.quadrature_points("boundary")
        return exact(xy, t).reshape(op.n_boundary, op.n_q_points_1d, 3)

    return Problem(mesh, degree, bcs, problem_data, lambda xy: exact(xy, 0.0),
                   None if periodic else boundary_values, exact, g=g,
                   name="vortex_p%d" % degree)


def channel(degree=2, nx=250, ny=10, seed=7, manning=0.02, discharge=4.42, h_out=2.0,
            g=9.81, n_bumps=-1):
    mesh = Mesh.rectangle(0.0, 25.0, 0.0, 1.0, nx, ny, boundary_ids=(1, 2, 0, 0))
    prm = [manning, n_bumps, 0.02, 0.05, 0.02, 0.05, 1.0, 24.0]

    def problem_data(xy):
        return channel_data(xy, prm, seed)

    def initial(xy):
        zb = problem_data(xy)[:, 0]
        h = np.maximum(h_out + zb, 1e-12)
        return np.stack([np.full(xy.shape[0], h_out), discharge / h, np.zeros(xy.shape[0])], 1)

    def boundary_values(op, t):
        v = np.zeros((op.n_boundary, op.n_q_points_1d, 3))
        v[:, :, 0] = h_out
        v[:, :, 1] = discharge
        return v

    bcs = {0: "wall", 1: "discharge_inflow", 2: "height_inflow"}
    return Problem(mesh, degree, bcs, problem_data, initial, boundary_values, None, g=g,
                   name="channel_p%d" % degree)


def forced(degree=2, refine=1, refine_fraction=0.3, seed=5, uniform=False, g=9.81):
    """A periodic basin driven by wind stress and Coriolis over a smooth bed with Manning
    friction: every term of ShallowWater::source (ShallowWater.h:408-432) is active.
    uniform=False stores non-uniform wind and Coriolis fields (a beta plane f = f0 + beta y)
    and a non-flat bed; uniform=True makes every problem-data field a constant (the kernels
    then read each as a scalar)."""
    mesh = Mesh.rectangle(0.0, 2.0, 0.0, 1.0, 10, 5, periodic=(True, True))
    if refine:
        mesh.refine_global(refine)
    if refine_fraction:
        mesh.refine_fraction(refine_fraction, seed)

    def problem_data(xy):
        x, y = xy[:, 0], xy[:, 1]
        out = np.zeros((xy.shape[0], 5))
        if uniform:
            out[:, 0] = -0.1
            out[:, 1] = 0.03
            out[:, 2] = 0.06
            out[:, 3] = -0.04
            out[:, 4] = 0.9
        else:
            out[:, 0] = -0.1 + 0.05 * np.sin(np.pi * x) * np.cos(2 * np.pi * y)
            out[:, 1] = 0.03
            out[:, 2] = 0.05 + 0.02 * np.sin(2 * np.pi * y)
            out[:, 3] = -0.03 * np.cos(np.pi * x)
            out[:, 4] = 0.5 + 0.3 * np.sin(2 * np.pi * y)
        return out

    def initial(xy):
        x, y = xy[:, 0], xy[:, 1]
        return np.stack([1.0 + 0.05 * np.sin(np.pi * x) * np.sin(2 * np.pi * y),
                         0.3 + 0.1 * np.cos(2 * np.pi * y), 0.1 * np.sin(np.pi * x)], 1)

    return Problem(mesh, degree, {}, problem_data, initial, None, None, g=g,
                   name="forced_%s_p%d" % ("uniform" if uniform else "stored", degree))


def lake_at_rest(degree=2, nx=8, ny=4, level=1.0, seed=3):
    """zeta = const, q = 0 over a bed with smooth and discontinuous parts
    (Icbc_LakeAtRest.h:45-77): the well-balanced operator returns L = 0."""
    mesh = Mesh.rectangle(0.0, 2.0, 0.0, 1.0, nx, ny, boundary_ids=(0, 0, 0, 0))
    rng = np.random.default_rng(seed)
    steps = rng.uniform(0.2, 0.8, size=4)

    def problem_data(xy):
        x, y = xy[:, 0], xy[:, 1]
        zb = 0.5 + 0.2 * np.sin(3 * x) * np.cos(2 * y) + np.where(x > 0.73, 0.1, 0.0)
        out = np.zeros((xy.shape[0], 5))
        out[:, 0] = zb
        out[:, 1] = 0.03
        return out

    def initial(xy):
        return np.stack([np.full(xy.shape[0], level), np.zeros(xy.shape[0]),
                         np.zeros(xy.shape[0])], 1)

    return Problem(mesh, degree, {0: "wall"}, problem_data, initial, None, None,
                   name="lake_at_rest")


def boundary_table(problem, op, t0, dt, n_steps, c):
    """Boundary data of n_steps RK steps from t0 in the order a replayed time loop reads
    them: entry r S + s holds the data at t0 + r dt + c[s] dt (the stage times of
    ExplicitRungeKuttaIntegrator.h:212-271).  Returns (times, values[n, n_bdry, nq, 3])."""
    times = np.array([t0 + r * dt + cs * dt for r in range(n_steps) for cs in c])
    # a step's last stage time repeats as the next step's first (or as another stage's):
    # evaluate each time once, so that repeated entries hold identical data whichever of
    # them a call picks
    values, first = [], []
    for t in times:
        tol = 1e-12 * max(1.0, abs(t))
        j = next((k for k in first[-2 * len(c):] if abs(times[k] - t) <= tol), None)
        if j is None:
            first.append(len(values))
            values.append(problem.boundary_values(op, t))
        else:
            values.append(values[j])
    return times, np.stack(values)


def setup_operator(problem, precision="fp64", device="cuda", max_iteration_height=15):
    """Build the GPU operator, tabulate the static data and project the initial state."""
    from .operator import OceanoOperator
    op = OceanoOperator(problem.mesh, problem.degree, g=problem.g,
                        viscosity=problem.viscosity, precision=precision,
                        boundary_conditions=problem.bcs, device=device,
                        max_iteration_height=max_iteration_height)
    data = op.initialize_data_at_quadrature(problem.problem_data)
    if problem.boundary_values is not None and op.n_boundary:
        op.set_boundary_values(problem.boundary_values(op, 0.0))
    zb2 = data["cell_2"][:, 0]
    h, q = op.project_hydro(problem.initial, zb2)
    return op, data, h, q


class MeshUnion:
    """Disjoint union of meshes as one operator mesh: the cell and face indices of instance
    i are offset by the cells before it, so that one operator (one launch per stage)
    advances every instance; no face joins two instances.  The instances keep their own
    coordinates (the operator reads only cell sizes and connectivity); functions of the
    union tell instances apart by the cell-major order of the points they receive."""

    def __init__(self, meshes, x_shift=0.0):
        geoms, inners, bdrys, starts = [], [], [], [0]
        for i, m in enumerate(meshes):
            g = m.geom.copy()
            g[:, 0] += x_shift * i
            geoms.append(g)
            inn = m.inner.copy()
            inn[:, 0] += starts[-1]
            inn[:, 3] += starts[-1]
            inners.append(inn)
            b = m.boundary.copy()
            b[:, 0] += starts[-1]
            bdrys.append(b)
            starts.append(starts[-1] + m.n_cells)
        self.geom = np.ascontiguousarray(np.concatenate(geoms), np.float64)
        self.inner = np.ascontiguousarray(np.concatenate(inners), np.int32)
        self.boundary = np.ascontiguousarray(np.concatenate(bdrys), np.int32).reshape(-1, 3)
        self.n_cells = starts[-1]
        self.n_hanging = sum(getattr(m, "n_hanging", 0) for m in meshes)
        self.starts = starts
        self.x_shift = x_shift
        self._keep = meshes

    def ranges(self, n_points):
        """row ranges of each instance in a cell-major array of n_points points"""
        per = n_points // self.n_cells
        return [(per * a, per * b) for a, b in zip(self.starts[:-1], self.starts[1:])]


def ensemble(n_instances=64, seed=11, n_min=1000, n_max=50000, degrees=(1, 2, 3, 4),
             g=9.81):
    """C4: n_instances travelling vortices.  Per instance: N ~ log-uniform[n_min, n_max]
    realised as the nearest 2:1 grid (2m x m cells) on [0,2]x[0,1], periodic; p ~ U(degrees);
    centre ~ U([0.5,1.5]x[0.4,0.6]), radius ~ U[0.2,0.35], amplitude ~ U[0.05,0.15]
    (numpy default_rng(seed)).  Returns {degree: Problem over the union of that degree's
    instances} and the per-instance parameters."""
    rng = np.random.default_rng(seed)
    inst = []
    for i in range(n_instances):
        n = np.exp(rng.uniform(np.log(n_min), np.log(n_max)))
        m = max(1, int(round(np.sqrt(n / 2.0))))
        inst.append({"ny": m, "nx": 2 * m, "degree": int(rng.choice(degrees)),
                     "center": (rng.uniform(0.5, 1.5), rng.uniform(0.4, 0.6)),
                     "radius": rng.uniform(0.2, 0.35), "amplitude": rng.uniform(0.05, 0.15)})
    groups = {}
    for p in sorted({d["degree"] for d in inst}):
        members = [d for d in inst if d["degree"] == p]
        meshes = [Mesh.rectangle(0.0, 2.0, 0.0, 1.0, d["nx"], d["ny"], periodic=(True, True))
                  for d in members]
        U = MeshUnion(meshes)
        prms = [[1.0, 1.0, 0.0, d["amplitude"], d["radius"], d["center"][0], d["center"][1], g,
                 2.0, 1.0] for d in members]

        def exact(xy, t, U=U, prms=prms):
            out = np.zeros((xy.shape[0], 3))
            for (a, b), prm in zip(U.ranges(xy.shape[0]), prms):
                out[a:b] = vortex_exact(xy[a:b], t, prm)
            return out

        groups[p] = Problem(U, p, {}, lambda xy: np.zeros((xy.shape[0], 5)),
                            lambda xy, exact=exact: exact(xy, 0.0), None, exact, g=g,
                            name="ensemble_p%d" % p)
    return groups, inst
