"""Pins of the CPU oracle (oracle/oracle.cpp) against mathematics and invariants.

The reference ships no golden vectors or known-answer tests for this path (its tests/
holds two run decks), and deal.II -- which holds the quadrature, basis, sum-factorisation
and inverse-mass arithmetic -- is not available, so every piece is pinned here by closed
forms, library routines (numpy) and invariants chosen so that a dropped term, a wrong
sign/index or a transposed operand fails one of them.
"""
import numpy as np
import pytest

import pyoracle as po
from oceano_amd import problems
from oceano_amd.integrator import butcher_rows
from oceano_amd.mesh import Mesh


# ----------------------------------------------------------------- quadrature -------
@pytest.mark.parametrize("n", range(1, 9))
def test_gauss_matches_numpy(n):
    x, w = po.rule("gauss", n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, (xn + 1) / 2, atol=1e-15, rtol=0)
    assert np.allclose(w, wn / 2, atol=1e-15, rtol=0)


@pytest.mark.parametrize("n", range(2, 9))
def test_gauss_lobatto_nodes_weights_exactness(n):
    x, w = po.rule("lobatto", n)
    assert x[0] == 0.0 and x[-1] == 1.0
    # interior nodes are the roots of P'_{n-1}
    c = np.zeros(n)
    c[-1] = 1
    roots = np.sort((np.polynomial.legendre.legroots(np.polynomial.legendre.legder(c)) + 1) / 2)
    assert np.allclose(x[1:-1], roots, atol=1e-14)
    # exact for x^k, k <= 2n-3, not for 2n-2
    for k in range(2 * n - 2):
        assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) < 1e-14
    k = 2 * n - 2
    assert abs(np.dot(w, x ** k) - 1.0 / (k + 1)) > 1e-10


# ----------------------------------------------------------------- basis -------------
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_lagrange_identities(p):
    nodes, _ = po.rule("lobatto", p + 1)
    V, D = po.lagrange(nodes, nodes)
    assert np.allclose(V, np.eye(p + 1), atol=1e-15)
    pts = np.linspace(0.03, 0.97, 11)
    V, D = po.lagrange(nodes, pts)
    assert np.allclose(V.sum(1), 1.0, atol=1e-14)       # partition of unity
    assert np.allclose(D.sum(1), 0.0, atol=1e-12)
    for k in range(p + 1):                               # exact for degree <= p
        assert np.allclose(V @ nodes ** k, pts ** k, atol=1e-13)
        assert np.allclose(D @ nodes ** k, k * pts ** max(k - 1, 0) * (k > 0), atol=1e-12)


# ------------------------------------------------------ single-cell brute force -------
def _single_cell_oracle(p, hx=0.37, hy=0.21):
    geom = np.array([[0.1, 0.2, hx, hy]])
    inner = np.array([[0, 1, -1, 0, 0, -1], [0, 3, -1, 0, 2, -1]], np.int32)  # periodic
    return po.Oracle(geom, inner, np.zeros((0, 3), np.int32), p)


def _dense_mass(p, hx, hy, rule_n):
    nodes, _ = po.rule("lobatto", p + 1)
    x, w = rule_n
    V, _ = po.lagrange(nodes, x)
    M1 = (V * w[:, None]).T @ V
    return np.kron(M1 * hy, M1 * hx)  # index i + n1*j: kron(y, x)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_inverse_mass_is_dense_inverse(p):
    """CellwiseInverseMassMatrix == inverse of the exactly integrated mass matrix."""
    hx, hy = 0.37, 0.21
    o = _single_cell_oracle(p, hx, hy)
    M = _dense_mass(p, hx, hy, po.rule("gauss", p + 4))  # exact for degree 2p
    rng = np.random.default_rng(p)
    x = rng.standard_normal((p + 1) ** 2)
    y = o.inverse_mass_tensor(M @ x, 1)
    assert np.allclose(y, x, rtol=0, atol=1e-12 * np.abs(x).max())


@pytest.mark.parametrize("p", [1, 2, 3])
def test_depth_mass_is_dense_mass(p):
    """cell_mass_height with z_b = 0 and a wet state is M_GL zeta (GL(p+2) is exact)."""
    hx, hy = 0.37, 0.21
    o = _single_cell_oracle(p, hx, hy)
    o.initialize_data_at_quadrature(lambda xy: np.zeros((len(xy), 5)))
    M = _dense_mass(p, hx, hy, po.rule("lobatto", p + 2))
    Mex = _dense_mass(p, hx, hy, po.rule("gauss", p + 4))
    assert np.allclose(M, Mex, atol=1e-15)
    z = 1.0 + 0.1 * np.random.default_rng(0).standard_normal((p + 1) ** 2)
    out = np.zeros_like(z)
    o.cell_mass_height(z, out)
    assert np.allclose(out, M @ z, atol=1e-15)


# ------------------------------------------------------------ depth-mass Jacobi -------
@pytest.mark.parametrize("p", [1, 2, 3])
def test_jacobi_fixed_point(p):
    """front = M h(zeta): the first sweep changes nothing and the loop stops after it."""
    P = problems.lake_at_rest(degree=p)
    o = po.Oracle.from_mesh(P.mesh, p, bcs=P.bcs)
    o.initialize_data_at_quadrature(P.problem_data)
    z = 1.0 + 0.01 * np.random.default_rng(1).standard_normal(o.N * o.nd)
    front = o.zeros(1)
    o.cell_mass_height(z, front)
    zz = z.copy()
    o.inverse_mass_height(zz, front)
    assert np.abs(zz - z).max() < 1e-13
    assert o.last_sweeps() == o.N  # one sweep per cell


# spectral radius of I - D^-1 M for the lumped GL(p+2) mass (documented for the reference)
RHO = {1: 0.889, 2: 0.840, 3: 0.816, 4: 0.803}


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_jacobi_matches_dense_iteration(p):
    """15 capped sweeps equal the dense linear Jacobi iteration on a wet cell; its
    iteration matrix has the spectral radius quoted for the reference's scheme."""
    hx, hy = 0.37, 0.21
    o = _single_cell_oracle(p, hx, hy)
    o.initialize_data_at_quadrature(lambda xy: np.zeros((len(xy), 5)))
    M = _dense_mass(p, hx, hy, po.rule("lobatto", p + 2))
    D = M.sum(1)  # lumped: row sums (integral of phi_i)
    Jm = np.eye(len(D)) - M / D[:, None]
    rho = max(abs(np.linalg.eigvals(Jm)))
    assert abs(rho - RHO[p]) < 1e-3
    rng = np.random.default_rng(p)
    z0 = 1.0 + 0.01 * rng.standard_normal(len(D))
    target = z0 + 1e-3 * rng.standard_normal(len(D))
    front = M @ target
    z = z0.copy()
    for _ in range(15):
        z = z + (front - M @ z) / D
    zz = z0.copy()
    o.inverse_mass_height(zz, front)
    assert o.last_sweeps() == 15
    assert np.allclose(zz, z, atol=1e-14)
    # and it is deliberately not converged: the remaining error follows rho^15
    assert np.abs(zz - target).max() > 1e-3 * 1e-3


def test_dry_lift():
    """An all-dry cell restarts from -max(z_b) + 1e-9 (OceanDG.h:1399-1422); with the
    front of that lifted state the first sweep is a fixed point."""
    o = _single_cell_oracle(1)
    o.set_static("cell_1", np.full(9, -0.2))
    lift = -(-0.2) + 1e-09
    front = np.zeros(4)
    o.cell_mass_height(np.full(4, lift), front)
    zz = np.full(4, 0.1)          # below the bed everywhere: h = 0.1 - 0.2 < 0
    o.inverse_mass_height(zz, front)
    assert np.array_equal(zz, np.full(4, lift))
    assert o.last_sweeps() == 1


def test_partially_dry_degenerate_lumped_mass():
    """Reference behaviour kept as is: when every quadrature point supporting a basis
    function is dry in a partially wet cell, the masked lumped mass D_i is 0 and the
    update 1/D_i * rhs_i is not finite (operators.h:294-297 has no guard)."""
    o = _single_cell_oracle(1)
    zb = np.array([-0.3, -0.25, -0.2, -0.31, -0.26, -0.21, -0.32, -0.27, -0.22])
    o.set_static("cell_1", zb)
    zz = np.full(4, 0.1)
    o.inverse_mass_height(zz, np.zeros(4))
    assert not np.isfinite(zz).all()


# ------------------------------------------------------------ fluxes ------------------
def _rhs_two_cells(state_m, state_p, p=1, zb=(0.0, 0.0)):
    """Two cells side by side, periodic in y, walls nowhere: returns L of cell 0 / 1."""
    geom = np.array([[0.0, 0.0, 1.0, 1.0], [1.0, 0.0, 1.0, 1.0]])
    inner = np.array([[0, 1, -1, 1, 0, -1], [1, 1, -1, 0, 0, -1],
                      [0, 3, -1, 0, 2, -1], [1, 3, -1, 1, 2, -1]], np.int32)
    o = po.Oracle(geom, inner, np.zeros((0, 3), np.int32), p)
    o.initialize_data_at_quadrature(
        lambda xy: np.stack([np.where(xy[:, 0] < 1.0, zb[0], zb[1])] + [0 * xy[:, 0]] * 4, 1))
    nd = (p + 1) ** 2
    h = np.concatenate([np.full(nd, state_m[0]), np.full(nd, state_p[0])])
    q = np.concatenate([np.full(nd, state_m[1]), np.full(nd, state_m[2]),
                        np.full(nd, state_p[1]), np.full(nd, state_p[2])])
    return o.rhs(h, q)


def test_flux_conservation_and_upwinding():
    """Piecewise-constant states: cell integrals vanish, so L is the face flux.  Mass is
    conserved exactly, and the LLF flux reduces to the physical flux for equal states."""
    g = 9.81
    Lh, Lq = _rhs_two_cells((1.2, 0.3, 0.1), (0.9, -0.2, 0.05))
    assert abs(Lh.sum()) < 1e-14
    # equal states on both sides: no net face contribution anywhere
    Lh, Lq = _rhs_two_cells((1.0, 0.3, 0.1), (1.0, 0.3, 0.1))
    assert np.abs(Lh).max() < 1e-14 and np.abs(Lq).max() < 1e-14
    # closed form for one face: zeta jump only, q = 0 -> F = 0.5*lambda*(z_m - z_p)
    Lh, Lq = _rhs_two_cells((1.2, 0.0, 0.0), (1.0, 0.0, 0.0))
    lam = max(np.sqrt(g * 1.2), np.sqrt(g * 1.0))
    F = 0.5 * lam * (1.2 - 1.0)
    # p = 1, unit cells: the face x=1 carries nodes 1, 3 of cell 0 (minus side, gets -F)
    # and nodes 0, 2 of cell 1 (plus side, gets +F); each node's face weight is 1/2
    nodes = Lh.reshape(2, 4)
    assert abs((nodes[0, 1] + nodes[0, 3]) + F) < 1e-13
    assert abs((nodes[1, 0] + nodes[1, 2]) - F) < 1e-13
    # pressure: centred strong jump P = g/4 (h_p + h_m)(zeta_p - zeta_m) n, submitted with
    # the same sign on both sides (-P on the minus side, -P on the plus side)
    P = g * 0.25 * (1.0 + 1.2) * (1.0 - 1.2)
    qx = Lq.reshape(2, 2, 4)[:, 0]
    assert abs((qx[0, 1] + qx[0, 3]) + P) < 1e-13
    assert abs((qx[1, 0] + qx[1, 2]) + P) < 1e-13
    assert np.abs(Lq.reshape(2, 2, 4)[:, 1]).max() < 1e-14  # no y-momentum


# ------------------------------------------------------- whole-operator invariants -----
@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("frac", [None, 0.3])
def test_free_stream_and_mass(p, frac):
    m = Mesh.rectangle(0, 2, 0, 1, 8, 4, periodic=(True, True)).refine_global(1)
    if frac:
        m.refine_fraction(frac, 5)
    o = po.Oracle.from_mesh(m, p)
    o.initialize_data_at_quadrature(lambda xy: np.zeros((len(xy), 5)))
    h = o.project(np.ones(o.N * o.nd), 1)
    qv = np.zeros((o.N * o.nd, 2))
    qv[:, 0], qv[:, 1] = 1.0, 0.5
    Lh, Lq = o.rhs(h, o.project(qv, 2))
    assert np.abs(Lh).max() < 1e-14 and np.abs(Lq).max() < 1e-14
    rng = np.random.default_rng(1)
    Lh, _ = o.rhs(1 + 0.1 * rng.standard_normal(o.N * o.nd),
                  0.1 * rng.standard_normal(2 * o.N * o.nd))
    assert abs(Lh.sum()) < 1e-13


@pytest.mark.parametrize("p", [1, 2, 3])
def test_lake_at_rest(p):
    P = problems.lake_at_rest(degree=p)
    o = po.Oracle.from_mesh(P.mesh, p, bcs=P.bcs)
    d = o.initialize_data_at_quadrature(P.problem_data)
    h, q = o.project_hydro(P.initial, d["cell_2"][:, 0])
    Lh, Lq = o.rhs(h, q)
    assert np.abs(Lh).max() < 1e-14 and np.abs(Lq).max() < 1e-13


def test_projection_error_p1_is_roundoff():
    """At p = 1 the projection's Gauss(p+1) points are the error quadrature's points, so the
    t = 0 error is round-off (README.md:71 reports 6.291e-16 / 4.93e-15)."""
    P = problems.vortex(degree=1, refine=1)
    o = po.Oracle.from_mesh(P.mesh, 1)
    d = o.initialize_data_at_quadrature(P.problem_data)
    h, q = o.project_hydro(P.initial, d["cell_2"][:, 0])
    err = o.errors(h, q, P.exact, 0.0)
    assert err[0] < 1e-14 and err[1] < 1e-13


def test_vortex_convergence():
    """The stage reproduces the travelling vortex with the DG order of accuracy."""
    errs = []
    for refine in (1, 2):
        P = problems.vortex(degree=2, refine=refine, u_inf=(1.0, 0.0))
        o = po.Oracle.from_mesh(P.mesh, 2)
        d = o.initialize_data_at_quadrature(P.problem_data)
        h, q = o.project_hydro(P.initial, d["cell_2"][:, 0])
        dt = 2.5e-4 / refine
        n = int(round(0.01 / dt))
        rows, _ = butcher_rows("stage_4_order_4", dt)
        ki_h = [o.zeros(1) for _ in range(5)]
        ki_q = [o.zeros(2) for _ in range(5)]
        r_h, r_q = o.zeros(1), o.zeros(2)
        for _ in range(n):
            o.time_step(rows, h, q, ki_h, ki_q, r_h, r_q)
        errs.append(o.errors(h, q, P.exact, n * dt))
    order_h = np.log2(errs[0][0] / errs[1][0])
    order_q = np.log2(errs[0][1] / errs[1][1])
    assert order_h > 2.3 and order_q > 2.3, (errs, order_h, order_q)


# --------------------------------------------- time-loop diagnostics (N1 / N3) ---------
@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_velocity_field_reproduces_polynomial_velocity(p):
    """evaluate_velocity_field: for a wet constant depth H (h^4 >= cu^4, so the
    desingularised inverse depth is 1/H to round-off) and q = H u with u a polynomial of
    degree p per direction, the collocation back to the nodal basis gives u's nodal
    values exactly."""
    m = Mesh.rectangle(0, 2, 0, 1, 4, 2, periodic=(True, True))
    o = po.Oracle.from_mesh(m, p)
    o.initialize_data_at_quadrature(lambda xy: np.zeros((len(xy), 5)))
    H = 1.7
    ux = lambda xy: 0.3 + xy[:, 0] ** p - 0.2 * xy[:, 1] * xy[:, 0]
    uy = lambda xy: -0.1 + 0.5 * xy[:, 1] ** p
    xy = o.quadrature_points("cell_2")
    h = o.project(np.full(len(xy), H), 1)
    q = o.project(np.ascontiguousarray(np.stack([H * ux(xy), H * uy(xy)], 1)), 2)
    vel = o.velocity_field(h, q)
    ref = o.project(np.ascontiguousarray(np.stack([ux(xy), uy(xy)], 1)), 2)
    assert np.abs(vel - ref).max() < 1e-13


def test_transport_speed_closed_form():
    """compute_cell_transport_speed on a uniform state: the largest cell gives
    lambda_max (1/h) sqrt(g H) + max(|U|/hx, |V|/hy); for square cells lambda_max = 1/h,
    for an hx < hy cell the five power iterations give sqrt((1 + r^2 hx^2/hy^2)/(1 + r^2))/hx
    with r = (hx/hy)^10."""
    g, H, U, V = 9.81, 2.0, 0.7, -0.4
    for nx, ny in ((8, 4), (8, 8)):
        m = Mesh.rectangle(0, 2, 0, 1, nx, ny, periodic=(True, True))
        o = po.Oracle.from_mesh(m, 2, g=g)
        o.initialize_data_at_quadrature(lambda xy: np.zeros((len(xy), 5)))
        n = o.N * o.nd
        h = o.project(np.full(n, H), 1)
        q = o.project(np.ascontiguousarray(np.tile([H * U, H * V], (n, 1))), 2)
        hx, hy = 2.0 / nx, 1.0 / ny
        r = (min(hx, hy) / max(hx, hy)) ** 10
        hs, hl = min(hx, hy), max(hx, hy)
        lam = np.sqrt((1 + r * r * hs * hs / (hl * hl)) / (1 + r * r)) / hs
        # inverse_depth = sqrt(2) H / sqrt(2 H^4) = 1/H up to rounding
        expect = lam * np.sqrt(g * H) + max(abs(U) / hx, abs(V) / hy)
        assert abs(o.transport_speed(h, q) - expect) < 1e-12 * expect


def test_volume_is_integral_of_depth():
    """check_mass cell integral: zeta + z_b with a polynomial bed integrates exactly by
    GL(p+2) (degree <= 2p+1); depth floor applies where the bed is above zeta."""
    p = 2
    m = Mesh.rectangle(0, 2, 0, 1, 6, 3, periodic=(True, True))
    o = po.Oracle.from_mesh(m, p)
    zb = lambda xy: 0.5 + 0.1 * xy[:, 0] - 0.2 * xy[:, 1] ** 2
    o.initialize_data_at_quadrature(
        lambda xy: np.concatenate([zb(xy)[:, None], np.zeros((len(xy), 4))], 1))
    xy = o.quadrature_points("cell_2")
    h = o.project(1.0 + 0.3 * xy[:, 0], 1)
    # integral over [0,2]x[0,1] of 1 + 0.3x + 0.5 + 0.1x - 0.2y^2
    expect = 2.0 * 1.5 + 0.4 * 2.0 - 0.2 * 2.0 / 3.0
    assert abs(o.volume(h) - expect) < 1e-13


# ---------------------------------------------------- IMEX ARK stage (N4) -------------
def _ark_step(o, h, q, dt, v=None):
    S = 3
    w_h, w_q = h.copy(), q.copy()
    ki_h = [o.zeros(1) for _ in range(S + 1)]
    ki_q = [o.zeros(2) for _ in range(2 * S)]
    o.time_step_ark(dt, w_h, w_q, ki_h, ki_q, o.zeros(1), o.zeros(2), v=v)
    return w_h, w_q


def test_ark_without_friction_is_the_explicit_scheme():
    """With zero friction the stiff terms vanish and the modified mass is the mass, so the
    IMEX step is the explicit stage_3_order_2 step (the momentum update is assembled as
    M^-1 (M q + sum a_j L_j) instead of q + sum a_j M^-1 L_j: round-off apart)."""
    P = problems.vortex(degree=2, refine=1, refine_fraction=0.3, seed=2)
    o = po.Oracle.from_mesh(P.mesh, 2)
    d = o.initialize_data_at_quadrature(P.problem_data)
    h, q = o.project_hydro(P.initial, d["cell_2"][:, 0])
    dt = 4e-4
    ah, aq = _ark_step(o, h, q, dt)
    rows, _ = butcher_rows("stage_3_order_2", dt)
    S = len(rows)
    eh, eq = h.copy(), q.copy()
    o.time_step(rows, eh, eq, [o.zeros(1) for _ in range(S + 1)],
                [o.zeros(2) for _ in range(S + 1)], o.zeros(1), o.zeros(2))
    assert np.linalg.norm(ah - eh) < 1e-13 * np.linalg.norm(eh)
    assert np.linalg.norm(aq - eq) < 1e-13 * np.linalg.norm(eq)
    assert np.linalg.norm(aq - q) > 1e-6 * np.linalg.norm(q)  # the step did something


def test_ark_uniform_flow_friction_closed_form():
    """Uniform depth H and discharge Q over a flat periodic bed: every flux term vanishes
    and the momentum obeys dq/dt = -J(q) q, J = g n^2 |q/H| / H^(4/3) (Manning).  The IMEX
    step then reduces to the scalar recursion of the scheme: r_1 = (q + at_00 F0) /
    (1 + d_0 J0), r_2 = (q + at_10 F0 + at_11 F1) / (1 + d_1 J1), q' = q + sum b_j F_j
    (F_j = -J(r_j) r_j, dt-scaled coefficients), at every node."""
    from oceano_amd.integrator import ark_tableau
    p, H, Q, n, g = 2, 1.3, np.array([0.9, -0.4]), 0.03, 9.81
    m = Mesh.rectangle(0, 2, 0, 1, 4, 2, periodic=(True, True))
    o = po.Oracle.from_mesh(m, p, g=g)
    o.initialize_data_at_quadrature(
        lambda xy: np.stack([np.zeros(len(xy)), np.full(len(xy), n), np.zeros(len(xy)),
                             np.zeros(len(xy)), np.zeros(len(xy))], 1))
    N = o.N * o.nd
    h = o.project(np.full(N, H), 1)
    q = o.project(np.ascontiguousarray(np.tile(Q, (N, 1))), 2)
    dt = 0.05
    wh, wq = _ark_step(o, h, q, dt)

    def J(r):
        u = r / H
        return g * n * n * np.linalg.norm(u) / np.exp(np.log(H) * 1.33333333333)
    a, b, c, at, dtl = ark_tableau()
    F = lambda r: -J(r) * r
    r0 = Q
    r1 = (Q + dt * at[0][0] * F(r0)) / (1 + dt * dtl[0] * J(r0))
    r2 = (Q + dt * at[1][0] * F(r0) + dt * at[1][1] * F(r1)) / (1 + dt * dtl[1] * J(r1))
    q_new = Q + dt * (b[0] * F(r0) + b[1] * F(r1) + b[2] * F(r2))
    wq = wq.reshape(o.N, 2, o.nd)
    assert np.abs(wq[:, 0] - q_new[0]).max() < 1e-13
    assert np.abs(wq[:, 1] - q_new[1]).max() < 1e-13
    assert np.abs(wh - H).max() < 1e-13
    assert abs(q_new[0] - Q[0]) > 1e-4  # friction acted


def test_threaded_oracle_matches_sequential():
    """the OpenMP loops of the CPU baseline give the sequential stage to rounding (face
    contributions accumulate in another order)"""
    from oceano_amd.integrator import butcher_rows
    P = problems.vortex(degree=2, refine=1, refine_fraction=0.3, seed=3)
    out = []
    for threads in (1, 4):
        o = po.Oracle.from_mesh(P.mesh, 2)
        d = o.initialize_data_at_quadrature(P.problem_data)
        h, q = o.project_hydro(P.initial, d["cell_2"][:, 0])
        o.set_threads(threads)
        rows, _ = butcher_rows("stage_4_order_4", 2e-4)
        S = len(rows)
        o.time_step(rows, h, q, [o.zeros(1) for _ in range(S + 1)],
                    [o.zeros(2) for _ in range(S + 1)], o.zeros(1), o.zeros(2))
        out.append((h, q))
    assert np.linalg.norm(out[1][0] - out[0][0]) < 1e-14 * np.linalg.norm(out[0][0])
    assert np.linalg.norm(out[1][1] - out[0][1]) < 1e-14 * np.linalg.norm(out[0][1])
