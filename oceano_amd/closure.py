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
                 discharge=4.42, h_out=2.0, length=25.0, n_quad=48):
        self.seed, self.degree, self.nx, self.ny, self.factor = seed, degree, nx, ny, factor
        self.manning, self.discharge, self.h_out, self.length = manning, discharge, h_out, length
        self.dx = length / nx
        self.fine = problems.channel(degree=degree, nx=nx * factor, ny=ny * factor, seed=seed,
                                     manning=manning, discharge=discharge, h_out=h_out)
        self._fine_data = self.fine.problem_data
        # sub-grid statistics and the Q_p projection of the bed, per coarse column
        xq, wq = gauss(n_quad)
        x = (np.arange(nx)[:, None] + xq[None, :]) * self.dx
        b = self.bed(x.ravel()).reshape(nx, n_quad)
        mean = b @ wq
        dev = b - mean[:, None]
        var = (dev ** 2) @ wq
        third = (dev ** 3) @ wq
        skew = np.zeros(nx)
        big = var > 1e-14
        skew[big] = third[big] / var[big] ** 1.5
        self.stats = np.stack([mean, var, skew], 1)  # [nx, 3]
        P = legendre_modes(degree + 1, xq)  # [q, m]
        self.bed_modes = (b * wq[None, :]) @ P  # [nx, n1]
        self.plain = self._plain_problem()

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

    def cell_stats(self, mesh):
        """[n_cells, 3] sub-grid (mean, var, skew) of the bed for the cells of the coarse
        mesh (the bed depends on x only)."""
        ix = np.clip(np.round(mesh.geom[:, 0] / self.dx).astype(np.int64), 0, self.nx - 1)
        return self.stats[ix]


class Restriction:
    """Exact L2 projection from the fine to the coarse channel mesh (torch, on the device,
    cell-major vectors as produced by OceanoOperator.unpack)."""

    def __init__(self, coarse_mesh, fine_mesh, degree, factor, device="cuda"):
        import torch
        self.n1 = degree + 1
        self.f = factor
        self.Rm = torch.tensor(restriction_1d(degree, factor)[0], dtype=torch.float64,
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
        U = U.view(self.nyc, f, self.nxc, f, ncomp, n1, n1)
        C = __import__("torch").einsum("sad,tbe,ysxtcde->yxcab", self.Rm, self.Rm, U)
        C = C.reshape(self.nyc * self.nxc, ncomp, n1, n1)[self.pos_c]
        return C.reshape(-1).contiguous()
