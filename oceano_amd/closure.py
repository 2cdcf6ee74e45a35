"""Learned sub-grid closure of the coarse channel (project-original: the reference has no
closure; it enters the stage as an extra momentum source, the slot of
ShallowWater.h:425-429 / OceanDG.h:734-736).

Set-up of one bathymetry sample (`SubgridChannel`):
  * fine model   the channel [0,25]x[0,1] at `factor` times the coarse resolution per
                 direction, bathymetry sampled pointwise (the bumps are resolved);
  * plain model  the coarse channel whose bathymetry is the cell-wise L2 projection of the
                 fine bathymetry onto Q_p (the sub-grid mean in the polynomial sense) - the
                 plain sub-grid model the learned correction is added to;
  * restriction  R: fine DG state -> coarse DG state, the exact L2 projection (the fine
                 cells nest in the coarse ones; 1-D factors integrated exactly by Gauss).

The closure target of a fine state u_f (on the coarse cells) is the momentum tendency the
plain coarse operator misses:
    tau = R(M_f^-1 L_f(u_f)) - M_c^-1 L_c(R u_f)          (discharge part)
so that d/dt (R u_f) = M_c^-1 L_c(R u_f) + tau exactly.  The network predicts tau's modal
coefficients from the cell's modal state and the sub-grid bathymetry statistics.
"""
import numpy as np

from . import problems
from .mesh import Mesh


# ----------------------------------------------------------------- 1-D tools -----------
def gauss(n):
    """Gauss-Legendre points and weights on [0, 1]"""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def gauss_lobatto_nodes(n):
    """n Gauss-Lobatto points on [0, 1]: the FE_DGQ(n-1) support points"""
    if n == 2:
        return np.array([0.0, 1.0])
    c = np.zeros(n)
    c[-1] = 1.0
    inner = np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c))
    return 0.5 * (np.concatenate([[-1.0], np.sort(inner), [1.0]]) + 1.0)


def lagrange(nodes, t):
    """[len(t), len(nodes)] values of the Lagrange basis on `nodes` at `t`"""
    t = np.atleast_1d(t)
    out = np.ones((t.size, nodes.size))
    for i in range(nodes.size):
        for j in range(nodes.size):
            if j != i:
                out[:, i] *= (t - nodes[j]) / (nodes[i] - nodes[j])
    return out


def legendre_modes(n, t):
    """[len(t), n] orthonormal Legendre polynomials on [0, 1]"""
    t = np.atleast_1d(t)
    out = np.zeros((t.size, n))
    for m in range(n):
        c = np.zeros(m + 1)
        c[m] = 1.0
        out[:, m] = np.sqrt(2 * m + 1) * np.polynomial.legendre.legval(2 * t - 1, c)
    return out


def nodal_to_modal(degree):
    """[nd, nd] matrix taking lexicographic GL nodal coefficients (k = i + n1 j) to the
    coefficients of the tensor Legendre modes (m = a + n1 b, a the x degree)."""
    n1 = degree + 1
    V1 = legendre_modes(n1, gauss_lobatto_nodes(n1))  # V1[node, mode]
    V = np.kron(V1, V1)  # rows k = i + n1 j  (kron(A_y, A_x) ordering)
    return np.linalg.inv(V), V


def restriction_1d(degree, factor):
    """Rm[s] (n1 x n1): coarse nodal coefficients of the L2 projection of a fine 1-D
    polynomial living on sub-interval s of [0, 1]: M^-1 B[s] with
    B[s][a][b] = int_{s/f}^{(s+1)/f} l_a(t) l_b(f t - s) dt (exact)."""
    n1 = degree + 1
    nodes = gauss_lobatto_nodes(n1)
    xq, wq = gauss(n1 + 1)
    M = np.einsum("q,qa,qb->ab", wq, lagrange(nodes, xq), lagrange(nodes, xq))
    Mi = np.linalg.inv(M)
    R = np.zeros((factor, n1, n1))
    for s in range(factor):
        t = (xq + s) / factor
        B = np.einsum("q,qa,qb->ab", wq / factor, lagrange(nodes, t), lagrange(nodes, xq))
        R[s] = Mi @ B
    return R, M


# -------------------------------------------------------------- the sample -------------
class SubgridChannel:
    """One seeded bathymetry of the channel at coarse and fine resolution."""

    def __init__(self, seed, degree=2, nx=250, ny=10, factor=8, manning=0.02,
                 discharge=4.42, h_out=2.0, length=25.0, n_quad=48, factor_y=None,
                 steady_start=True):
        self.seed, self.degree, self.nx, self.ny, self.factor = seed, degree, nx, ny, factor
        self.factor_y = factor if factor_y is None else factor_y
        self.manning, self.discharge, self.h_out, self.length = manning, discharge, h_out, length
        self.dx = length / nx
        self.fine = problems.channel(degree=degree, nx=nx * factor, ny=ny * self.factor_y,
                                     seed=seed, manning=manning, discharge=discharge,
                                     h_out=h_out)
        self._fine_data = self.fine.problem_data
        # the Q_p projection of the bed and the statistics of its sub-grid part, per coarse
        # column: mean of b, variance and skewness of b - P_p b (what the coarse bed misses;
        # zero where the bed is resolved, e.g. on the parabolic bump)
        xq, wq = gauss(n_quad)
        x = (np.arange(nx)[:, None] + xq[None, :]) * self.dx
        b = self.bed(x.ravel()).reshape(nx, n_quad)
        P = legendre_modes(degree + 1, xq)  # [q, m]
        self.bed_modes = (b * wq[None, :]) @ P  # [nx, n1]
        mean = b @ wq
        dev = b - self.bed_modes @ P.T
        var = (dev ** 2) @ wq
        third = (dev ** 3) @ wq
        skew = np.zeros(nx)
        big = var > 1e-14
        skew[big] = third[big] / var[big] ** 1.5
        self.stats = np.stack([mean, var, skew], 1)  # [nx, 3]
        self.plain = self._plain_problem()
        if steady_start:
            self.profile = SteadyProfile(self.bed, discharge, manning, self.fine.g, h_out,
                                         length)
            self.fine.initial = self.profile.initial

    def bed(self, x):
        """b(x) = -z_b of the fine model (bump + seeded Gaussian ridges)"""
        xy = np.stack([np.asarray(x, np.float64), np.full(np.size(x), 0.5)], 1)
        return -self._fine_data(xy)[:, 0]

    def plain_bed(self, x):
        x = np.asarray(x, np.float64)
        ix = np.clip(np.floor(x / self.dx).astype(np.int64), 0, self.nx - 1)
        t = x / self.dx - ix
        P = legendre_modes(self.degree + 1, t)
        return np.einsum("nm,nm->n", self.bed_modes[ix], P)

    def _plain_problem(self):
        P = problems.channel(degree=self.degree, nx=self.nx, ny=self.ny, seed=self.seed,
                             manning=self.manning, discharge=self.discharge,
                             h_out=self.h_out)

        def data(xy):
            out = np.zeros((xy.shape[0], 5))
            out[:, 0] = -self.plain_bed(xy[:, 0])
            out[:, 1] = self.manning
            return out
        P.problem_data = data
        P.name = "channel_plain_subgrid_p%d" % self.degree
        return P

    def active_cells(self, mesh, min_variance=1e-8, hops=2):
        """cells within `hops` face neighbours of a cell whose sub-grid variance exceeds
        min_variance: the correction of a ridge spans its neighbourhood (its lobes balance;
        cutting one leaves a net force), elsewhere the plain model is kept"""
        st = self.cell_stats(mesh)
        act = st[:, 1] > min_variance
        inner = mesh.inner
        for _ in range(hops):
            nxt = act.copy()
            a, b = inner[:, 0], inner[:, 3]
            nxt[b[act[a]]] = True
            nxt[a[act[b]]] = True
            act = nxt
        return act.astype(np.int32)

    def cell_stats(self, mesh):
        """[n_cells, 3] sub-grid (mean, var, skew) of the bed for the cells of the coarse
        mesh (the bed depends on x only)."""
        ix = np.clip(np.round(mesh.geom[:, 0] / self.dx).astype(np.int64), 0, self.nx - 1)
        return self.stats[ix]


class Restriction:
    """Exact L2 projection from the fine to the coarse channel mesh (torch, on the device,
    cell-major vectors as produced by OceanoOperator.unpack)."""

    def __init__(self, coarse_mesh, fine_mesh, degree, factor, device="cuda", factor_y=None):
        import torch
        self.n1 = degree + 1
        self.f = factor
        self.fy = factor if factor_y is None else factor_y
        self.Rm = torch.tensor(restriction_1d(degree, factor)[0], dtype=torch.float64,
                               device=device)
        self.Rmy = torch.tensor(restriction_1d(degree, self.fy)[0], dtype=torch.float64,
                                device=device)
        gc, gf = coarse_mesh.geom, fine_mesh.geom
        hx, hy = gc[0, 2], gc[0, 3]
        ixc = np.round(gc[:, 0] / hx).astype(np.int64)
        iyc = np.round(gc[:, 1] / hy).astype(np.int64)
        self.nxc, self.nyc = int(ixc.max()) + 1, int(iyc.max()) + 1
        ixf = np.round(gf[:, 0] / gf[0, 2]).astype(np.int64)
        iyf = np.round(gf[:, 1] / gf[0, 3]).astype(np.int64)
        nxf = self.nxc * factor
        # fine cells in [jy][jx] order, coarse cells back from [iy][ix] order
        order_f = np.empty(gf.shape[0], np.int64)
        order_f[iyf * nxf + ixf] = np.arange(gf.shape[0])
        self.perm_f = torch.tensor(order_f, device=device)
        self.pos_c = torch.tensor(iyc * self.nxc + ixc, device=device)
        self.n_fine, self.n_coarse = gf.shape[0], gc.shape[0]

    def __call__(self, fine_cm, ncomp):
        n1, f = self.n1, self.f
        U = fine_cm.view(self.n_fine, ncomp, n1, n1)[self.perm_f]
        U = U.view(self.nyc, self.fy, self.nxc, f, ncomp, n1, n1)
        C = __import__("torch").einsum("sad,tbe,ysxtcde->yxcab", self.Rmy, self.Rm, U)
        C = C.reshape(self.nyc * self.nxc, ncomp, n1, n1)[self.pos_c]
        return C.reshape(-1).contiguous()


class SteadyProfile:
    """Steady subcritical solution of the 1-D channel (q = Q everywhere): the energy
    E = Q^2 / (2 g h^2) + h + b decreases along the flow by the Manning slope,
    dE/dx = -n^2 Q^2 / h^(10/3) (the friction source -g n^2 |v| v / h^(1/3) of the
    model), integrated upstream from the outflow level by classical RK4; at every x the
    depth is the subcritical root of the energy relation for the local bed.  Used as the
    initial state, so that the runs start near the fine model's own steady state."""

    def __init__(self, bed, Q, n, g, surface_out, length, dx=2.5e-3):
        self.bed, self.Q, self.n, self.g = bed, Q, n, g
        m = int(round(length / dx))
        self.x = np.linspace(0.0, length, m + 1)
        xm = 0.5 * (self.x[1:] + self.x[:-1])
        bx, bm = bed(self.x), bed(xm)
        self.hc = (Q * Q / g) ** (1.0 / 3.0)
        E = np.zeros(m + 1)
        h = surface_out - bx[-1]
        E[-1] = self._energy(h, bx[-1])
        for i in range(m, 0, -1):  # x_i -> x_{i-1}: dE = +slope dx going upstream
            e = E[i]
            k1 = self._slope(e, bx[i], h)
            h1 = self._h
            k2 = self._slope(e + 0.5 * dx * k1, bm[i - 1], h1)
            k3 = self._slope(e + 0.5 * dx * k2, bm[i - 1], h1)
            k4 = self._slope(e + dx * k3, bx[i - 1], h1)
            E[i - 1] = e + dx / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
            h = self._h
        self.E = E

    def _energy(self, h, b):
        return self.Q * self.Q / (2.0 * self.g * h * h) + h + b

    def depth(self, E, b, h0):
        """subcritical root of Q^2 / (2 g h^2) + h = E - b (Newton from h0)"""
        h = np.maximum(np.asarray(h0, np.float64), 1.2 * self.hc)
        c = self.Q * self.Q / self.g
        for _ in range(60):
            f = c / (2.0 * h * h) + h - (E - b)
            d = 1.0 - c / (h * h * h)
            step = f / d
            h = np.maximum(h - step, 1.0001 * self.hc)
            if np.all(np.abs(step) < 1e-15 * h):
                break
        return h

    def _slope(self, E, b, h0):
        self._h = float(self.depth(E, b, h0))
        return self.n * self.n * self.Q * self.Q / self._h ** (10.0 / 3.0)

    def h(self, x):
        x = np.asarray(x, np.float64)
        E = np.interp(x, self.x, self.E)
        return self.depth(E, self.bed(x), np.full(x.shape, 1.5))

    def initial(self, xy):
        """free surface, u, v at points (the Problem.initial convention)"""
        x = xy[:, 0]
        h = self.h(x)
        return np.stack([h + self.bed(x), self.Q / h, np.zeros(x.shape)], 1)


# ------------------------------------------------------------ data, training -------------
def l2_norms(cm, ncomp, degree, hx, hy):
    """per-component L2 norms over the mesh of a cell-major DG vector (torch or numpy)"""
    n1 = degree + 1
    _, M1 = restriction_1d(degree, 1)
    M2 = np.kron(M1, M1)
    d = np.asarray(cm.cpu().numpy() if hasattr(cm, "cpu") else cm).reshape(-1, ncomp, n1 * n1)
    return [float(np.sqrt(hx * hy * np.einsum("nk,kl,nl->", d[:, c], M2, d[:, c])))
            for c in range(ncomp)]


def _integrate(op, dt, n_steps, w_h, w_q, chunks=()):
    """n_steps RK4 steps by graph replay; returns after each of the step counts in
    `chunks` (ascending) a copy of the state: [(steps, zeta, q)]"""
    import torch
    from .integrator import DeviceIntegrator
    rk = DeviceIntegrator(op)
    side = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        rk.capture(dt, w_h, w_q)
    side.synchronize()
    out, done = [], 0
    marks = sorted(set([c for c in chunks if 0 <= c <= n_steps] + [n_steps]))
    for mk in marks:
        with torch.cuda.stream(side):
            if mk > done:
                rk.replay(mk - done)
        side.synchronize()
        done = mk
        if mk in chunks:
            out.append((mk, w_h.clone(), w_q.clone()))
    return out


def _cfl_dt(op, w_h, w_q, cfl, T):
    dt = cfl / float("%.3g" % op.transport_speed(w_h, w_q))
    n = int(np.ceil(T / dt))
    return T / n, n


class Sample:
    """fine run of one bathymetry with the closure data along it"""

    def __init__(self, seed, T=2.0, n_snap=8, cfl=0.075, factor=8, factor_y=1, nx=250,
                 ny=10, degree=2):
        import torch
        from .operator import OceanoOperator  # noqa: F401  (device set-up)
        self.S = SubgridChannel(seed, degree=degree, nx=nx, ny=ny, factor=factor,
                                factor_y=factor_y)
        S = self.S
        opf, _, fh, fq = problems.setup_operator(S.fine)
        opc, _, _, _ = problems.setup_operator(S.plain)
        self.active = S.active_cells(S.plain.mesh)
        opc.set_subgrid_stats(S.cell_stats(S.plain.mesh), self.active)
        R = Restriction(S.plain.mesh, S.fine.mesh, degree, factor, factor_y=factor_y)
        dt, n = _cfl_dt(opf, fh, fq, cfl, T)
        marks = [int(round(k * n / n_snap)) for k in range(n_snap + 1)]
        snaps = _integrate(opf, dt, n, fh, fq, chunks=marks)
        Vi, _ = nodal_to_modal(degree)
        Vi_t = torch.tensor(Vi.T, dtype=torch.float64, device="cuda")
        nd = (degree + 1) ** 2
        X, Tg, states = [], [], []
        vh, vq = opf.initialize_vector(0), opf.initialize_vector(1)
        ch, cq = opc.initialize_vector(0), opc.initialize_vector(1)
        for steps, zh, zq in snaps:
            rh = R(opf.unpack(zh, 1), 1)
            rq = R(opf.unpack(zq, 2), 2)
            opf.apply(0.0, [zh, zq], [vh, vq])
            rvq = R(opf.unpack(vq, 2), 2)
            ph, pq = opc.pack(rh, 1), opc.pack(rq, 2)
            opc.apply(0.0, [ph, pq], [ch, cq])
            tau = (rvq - opc.unpack(cq, 2)).view(-1, 2, nd)
            X.append(opc.closure_features(ph, pq))
            Tg.append((tau @ Vi_t).reshape(-1, 2 * nd).float())
            states.append((steps * dt, rh, rq))
        self.X_all, self.T_all = torch.cat(X), torch.cat(Tg)
        keep = torch.tensor(np.tile(self.active, len(X)) != 0, device="cuda")
        self.X, self.T = self.X_all[keep].contiguous(), self.T_all[keep].contiguous()
        self.states = states  # restricted fine states at the snapshot times
        self.dt_fine, self.n_fine = dt, n
        self.time = T
        self.cfl = cfl
        del opf


def fit(X, T, steps=20000, batch=4096, lr=2e-3, seed=0, noise=0.05, log=None):
    """train a closure network on features X and targets T (device float32 tensors);
    normalisation from the data, Adam with a cosine learning-rate decay.  Each batch's
    inputs get Gaussian noise of `noise` times each feature's spread, so that the learned
    source varies slowly with the state (the coupled coarse model stays stable); the
    output range seen in training (widened by 10%) bounds the correction at inference."""
    import torch
    from .mlp import ClosureNetwork
    n_in, n_out = X.shape[1], T.shape[1]
    xm, xs = X.double().mean(0), X.double().std(0)
    xs = torch.where(xs > 1e-12 * (1 + xm.abs()), xs, torch.ones_like(xs))
    tm, ts = T.double().mean(0), T.double().std(0)
    ts = torch.clamp(ts, min=1e-6 * float(ts.max()))
    net = ClosureNetwork(n_in, n_out, seed=seed)
    net.set_normalization(xm.cpu().numpy(), xs.cpu().numpy(), tm.cpu().numpy(),
                          ts.cpu().numpy())
    tn = (T.double() - tm) / ts
    lo, hi = tn.min(0).values, tn.max(0).values
    pad = 0.1 * (hi - lo)
    net.set_output_range((lo - pad).cpu().numpy(), (hi + pad).cpu().numpy())
    xs_f = (noise * xs).float()
    g = torch.Generator(device="cpu").manual_seed(seed)
    gn = torch.Generator(device=X.device).manual_seed(seed + 1)
    n = X.shape[0]
    hist = []
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(X.device)
        rate = lr * 0.5 * (1 + np.cos(np.pi * it / steps))
        want = log is not None and (it % max(1, steps // 20) == 0 or it == steps - 1)
        xb = X[idx]
        if noise > 0:
            xb = xb + xs_f * torch.randn(xb.shape, generator=gn, device=X.device)
        loss = net.train_step(xb.contiguous(), T[idx].contiguous(), rate, want_loss=want)
        if want:
            hist.append((it, loss))
            log("  step %6d  lr %.2e  loss %.4e" % (it, rate, loss))
    return net, hist


def _frozen_source(op, data, mesh, tau_nodal):
    """a fixed momentum source given by nodal values per cell ([n_cells, 2, nd]), through
    the wind-stress slot of the source (values at the cell Gauss points; the Gauss rule
    integrates it against the basis exactly, so M^-1 of its integral is tau itself)"""
    n1 = int(round(np.sqrt(tau_nodal.shape[2])))
    xy = op.quadrature_points("cell_0")
    npq = xy.shape[0] // op.n_cells
    d0 = data["cell_0"].copy()
    g = mesh.geom
    nodes = gauss_lobatto_nodes(n1)
    cell = np.repeat(np.arange(op.n_cells), npq)
    tx = (xy[:, 0] - g[cell, 0]) / g[cell, 2]
    ty = (xy[:, 1] - g[cell, 1]) / g[cell, 3]
    phi = (lagrange(nodes, ty)[:, :, None] * lagrange(nodes, tx)[:, None, :]).reshape(
        -1, n1 * n1)
    d0[:, 2] = np.einsum("pk,pk->p", phi, tau_nodal[cell, 0])
    d0[:, 3] = np.einsum("pk,pk->p", phi, tau_nodal[cell, 1])
    op.set_static_data("cell_0", d0)


def evaluate(sample, net=None, cfl=None, exact=True, gains=(), baselines=True):
    """coarse runs of one held-out sample from the restricted fine initial state to the
    final time: plain sub-grid model, plain + network (at the network's own output gain,
    and as "learned@g" at each gain g of `gains`), the pointwise-sampled bed (the
    reference's own sampling of the bed, OceanDG.h:498-569) and (the ceiling of any
    momentum-source correction) plain + the exact correction of the initial state frozen in
    time; returns the L2 errors (zeta, hu, hv) of each against the restricted fine
    solution"""
    S = sample.S
    t0, rh0, rq0 = sample.states[0]
    tT, rhT, rqT = sample.states[-1]
    cfl = cfl or sample.cfl
    out = {}
    cases = [("plain", S.plain, False, None), ("learned", S.plain, True, None)]
    cases += [("learned@%g" % g, S.plain, True, g) for g in gains]
    if baselines:
        cases += [("sampled", problems.channel(degree=S.degree, nx=S.nx, ny=S.ny, seed=S.seed,
                                               manning=S.manning), False, None),
                  ("exact", S.plain, False, None)]
    hx, hy = S.length / S.nx, 1.0 / S.ny
    nd = (S.degree + 1) ** 2
    own_gain = net.output_gain if net is not None else 1.0
    for name, prob, with_net, gain in cases:
        if (with_net and net is None) or (name == "exact" and not exact):
            continue
        if with_net:
            net.set_output_gain(own_gain if gain is None else gain)
        op, data, _, _ = problems.setup_operator(prob)
        if name == "exact":
            _, V = nodal_to_modal(S.degree)
            Tm = sample.T_all[:op.n_cells].double().cpu().numpy()
            _frozen_source(op, data, prob.mesh,
                           np.stack([Tm[:, :nd] @ V.T, Tm[:, nd:] @ V.T], 1))
        if with_net:
            op.set_subgrid_stats(S.cell_stats(prob.mesh), sample.active)
            op.set_closure(net)
        h, q = op.pack(rh0, 1), op.pack(rq0, 2)
        dt, n = _cfl_dt(op, h, q, cfl, tT - t0)
        _integrate(op, dt, n, h, q)
        eh = l2_norms(op.unpack(h, 1) - rhT, 1, S.degree, hx, hy)
        eq = l2_norms(op.unpack(q, 2) - rqT, 2, S.degree, hx, hy)
        out[name] = [eh[0], eq[0], eq[1]]
        if with_net:
            op.set_closure(None)
    if net is not None:
        net.set_output_gain(own_gain)
    return out


GAINS = (0.0, 0.03125, 0.0625, 0.125, 0.25, 0.5, 1.0)


def calibrate_gain(net, seeds, gains=GAINS, T=2.0, n_snap=2, log=None):
    """A-posteriori calibration of the correction on training bathymetries: coupled coarse
    runs with the network's output scaled by each gain, scored by the sum over the runs of
    the zeta and hu errors relative to the plain model's; the best gain is kept (0, the
    plain model, when every nonzero gain degrades the coupled runs).  Returns (gain,
    {gain: score})."""
    score = {g: 0.0 for g in gains}
    for seed in seeds:
        s = Sample(seed, T=T, n_snap=n_snap)
        r = evaluate(s, net, gains=gains, baselines=False)
        pz, pq = r["plain"][0], r["plain"][1]
        for g in gains:
            e = r["learned@%g" % g]
            score[g] += e[0] / pz + e[1] / pq
        if log:
            log("calibration seed %d: " % seed +
                " ".join("g=%g z %.2e hu %.2e" % (g, r["learned@%g" % g][0],
                                                 r["learned@%g" % g][1]) for g in gains))
        del s
    best = min(gains, key=lambda g: (score[g], g))
    return best, score


def run_experiment(n_train=256, n_eval=32, steps=20000, noise=0.02, T=2.0, n_snap=4,
                   train_seed0=1000, eval_seed0=2000, net_seed=0, n_cal=8, log=print):
    """The closure experiment: fine runs of n_train seeded bathymetries give the training
    data, the network is trained on the GPU, its output gain is calibrated on coupled
    coarse runs of n_cal training bathymetries (calibrate_gain), and n_eval held-out
    bathymetries are run with the plain sub-grid model, plain + network, the sampled bed and
    the exact-correction ceiling.  Returns a report dict (errors are root-sum-squares over
    the held-out runs)."""
    import time

    import torch
    t0 = time.time()
    X, Tg = [], []
    for i in range(n_train):
        s = Sample(train_seed0 + i, T=T, n_snap=n_snap)
        X.append(s.X)
        Tg.append(s.T)
        if log and (i % 16 == 0 or i == n_train - 1):
            log("train sample %d/%d (seed %d): fine %d steps, %.0f s" %
                (i + 1, n_train, s.S.seed, s.n_fine, time.time() - t0))
        del s
    X, Tg = torch.cat(X), torch.cat(Tg)
    t1 = time.time()
    net, hist = fit(X, Tg, steps=steps, noise=noise, seed=net_seed)
    torch.cuda.synchronize()
    t_fit = time.time() - t1
    if log:
        log("trained on %d samples in %.1f s (%d Adam steps), final loss %.4e" %
            (X.shape[0], t_fit, steps, hist[-1][1] if hist else float("nan")))
    gain, score = calibrate_gain(net, range(train_seed0, train_seed0 + n_cal), T=T, log=log)
    net.set_output_gain(gain)
    rep = {
        "n_train": n_train, "n_eval": n_eval, "adam_steps": steps, "input_noise": noise,
        "time_s": T, "snapshots": n_snap + 1, "train_samples": int(X.shape[0]),
        "train_seconds": t_fit, "calibration_seeds": n_cal, "gain": gain,
        "calibration_scores": {str(k): v for k, v in score.items()}}
    rep.update(evaluate_heldout(net, n_eval, T, n_snap, eval_seed0, log=log))
    rep["total_seconds"] = time.time() - t0
    return rep, net


def evaluate_heldout(net, n_eval=32, T=2.0, n_snap=4, eval_seed0=2000, log=None):
    """root-sum-square errors over n_eval held-out bathymetries: plain, plain + network at
    its calibrated gain ("learned") and at gain 1 ("learned@1"), the sampled bed and the
    exact-correction ceiling; with the gains against the plain model and the sampled bed"""
    import torch
    scale = torch.tensor(net.normalization[3], device="cuda", dtype=torch.float32)
    tot, apriori = {}, []
    for i in range(n_eval):
        ev = Sample(eval_seed0 + i, T=T, n_snap=n_snap)
        g = net.output_gain
        net.set_output_gain(1.0)
        loss, _ = net.gradients(ev.X, ev.T)
        net.set_output_gain(g)
        zero = float(((ev.T / scale) ** 2).mean())  # loss of no correction
        apriori.append((loss, zero))
        r = evaluate(ev, net, gains=(1.0,) if g != 1.0 else ())
        for k, v in r.items():
            tot.setdefault(k, np.zeros(3))
            tot[k] += np.array(v) ** 2
        if log:
            log("eval seed %d: a-priori loss %.3e (no correction %.3e); " %
                (ev.S.seed, loss, zero) +
                " ".join("%s z %.2e hu %.2e" % (k, v[0], v[1]) for k, v in r.items()))
        del ev
    err = {k: [float(x) for x in np.sqrt(v)] for k, v in tot.items()}
    return {
        "apriori_loss_network": float(np.mean([a for a, _ in apriori])),
        "apriori_loss_no_correction": float(np.mean([b for _, b in apriori])),
        "errors_zeta_hu_hv": err,
        "gain_vs_plain": {k: [err["plain"][j] / v[j] for j in range(2)]
                          for k, v in err.items()},
        "gain_vs_sampled_bed": {k: [err["sampled"][j] / v[j] for j in range(2)]
                                for k, v in err.items()} if "sampled" in err else None}


def frozen_diagnosis(net, seeds, T=2.0, n_snap=4, log=None):
    """Which part of the learned correction harms the coupled run: coarse runs with frozen
    momentum sources (each correction evaluated once at the initial state and held fixed
    through the wind-stress slot, _frozen_source) built from the exact and the learned
    correction (at gain 1) and mixtures of their modes.  The cell-mean modes are the first
    mode of each momentum component.  Returns {seed: {"errors": {case: (zeta, hu)},
    "modes": statistics of the cell-mean modes on the active cells}}."""
    import torch
    g0 = net.output_gain
    net.set_output_gain(1.0)
    out = {}
    for seed in seeds:
        s = Sample(seed, T=T, n_snap=n_snap)
        S = s.S
        n = S.plain.mesh.n_cells
        nd = (S.degree + 1) ** 2
        act = np.asarray(s.active, bool)
        Te = s.T_all[:n].double().cpu().numpy()
        Tl = net.forward(s.X_all[:n].contiguous()).double().cpu().numpy()
        Tl[~act] = 0.0
        means = [0, nd]
        unb = Tl.copy()
        for j in means:
            unb[act, j] -= Tl[act, j].mean()
        cases = {"plain": None, "exact": Te, "learned": Tl}
        mix = Tl.copy()
        mix[:, means] = Te[:, means]
        cases["learned, exact cell means"] = mix
        mix = Te.copy()
        mix[:, means] = Tl[:, means]
        cases["exact, learned cell means"] = mix
        cases["learned, net force removed"] = unb
        cases["learned x 0.25, net force removed"] = 0.25 * unb
        t0, rh0, rq0 = s.states[0]
        tT, rhT, rqT = s.states[-1]
        hx, hy = S.length / S.nx, 1.0 / S.ny
        _, V = nodal_to_modal(S.degree)
        err = {}
        for name, tau in cases.items():
            op, data, _, _ = problems.setup_operator(S.plain)
            if tau is not None:
                _frozen_source(op, data, S.plain.mesh,
                               np.stack([tau[:, :nd] @ V.T, tau[:, nd:] @ V.T], 1))
            h, q = op.pack(rh0, 1), op.pack(rq0, 2)
            dt, nst = _cfl_dt(op, h, q, s.cfl, tT - t0)
            _integrate(op, dt, nst, h, q)
            eh = l2_norms(op.unpack(h, 1) - rhT, 1, S.degree, hx, hy)
            eq = l2_norms(op.unpack(q, 2) - rqT, 2, S.degree, hx, hy)
            err[name] = (float(eh[0]), float(eq[0]))
        d = (Tl - Te)[act]
        modes = {"exact_rms": float(np.sqrt((Te[act][:, means] ** 2).mean())),
                 "learned_error_rms": float(np.sqrt((d[:, means] ** 2).mean())),
                 "learned_minus_exact_hu_mean_average": float(d[:, 0].mean()),
                 "exact_hu_mean_average": float(Te[act][:, 0].mean())}
        out[seed] = {"errors": err, "modes": modes}
        if log:
            pz, pq = err["plain"]
            log("seed %d: " % seed + "; ".join("%s gain %.2f %.2f" % (k, pz / v[0], pq / v[1])
                                               for k, v in err.items()) + "; %s" % modes)
        del s
        torch.cuda.synchronize()
    net.set_output_gain(g0)
    return out


def main(argv=None):
    import argparse
    import json
    ap = argparse.ArgumentParser(description="learned sub-grid closure experiment")
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--eval", type=int, default=32)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--noise", type=float, default=0.02)
    ap.add_argument("--time", type=float, default=2.0)
    ap.add_argument("--calibrate", default=None,
                    help="a trained network (.npz): calibrate its output gain on training "
                         "bathymetries and evaluate it on the held-out ones (no training)")
    ap.add_argument("--cal-seeds", type=int, default=8)
    ap.add_argument("--diagnose", type=int, nargs="*", default=None,
                    help="held-out seeds: frozen-source diagnosis of the shipped network "
                         "(frozen_diagnosis)")
    ap.add_argument("--out", default="closure_run")
    a = ap.parse_args(argv)
    import os
    os.makedirs(a.out, exist_ok=True)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.diagnose is not None:
        from .mlp import ClosureNetwork
        net = ClosureNetwork.load(os.path.join(os.path.dirname(__file__), "data",
                                               "closure_channel_p2.npz"))
        rep = frozen_diagnosis(net, a.diagnose or [2000, 2001, 2002, 2003], T=a.time, log=log)
        with open(os.path.join(a.out, "diagnosis.json"), "w") as f:
            json.dump({str(k): v for k, v in rep.items()}, f, indent=1)
        print(json.dumps({str(k): v for k, v in rep.items()}))
        return
    if a.calibrate:
        import time
        from .mlp import ClosureNetwork
        t0 = time.time()
        net = ClosureNetwork.load(a.calibrate)
        net.set_output_gain(1.0)
        gain, score = calibrate_gain(net, range(1000, 1000 + a.cal_seeds), T=a.time, log=log)
        net.set_output_gain(gain)
        log("calibrated output gain %g (scores %s)" % (gain, score))
        rep = {"calibration_seeds": a.cal_seeds, "gain": gain,
               "calibration_scores": {str(k): v for k, v in score.items()}}
        rep.update(evaluate_heldout(net, a.eval, a.time, log=log))
        rep["total_seconds"] = time.time() - t0
        net.save(os.path.join(a.out, "closure_channel_p2.npz"))
        with open(os.path.join(a.out, "report.json"), "w") as f:
            json.dump(rep, f, indent=1)
        print(json.dumps(rep))
        return
    rep, net = run_experiment(a.train, a.eval, a.steps, a.noise, a.time, log=log)
    net.save(os.path.join(a.out, "closure_channel_p2.npz"))
    with open(os.path.join(a.out, "report.json"), "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
