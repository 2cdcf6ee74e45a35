"""CPU checks of the learned-closure pieces: the network oracle's backward pinned by finite
differences, BF16 rounding, the modal basis, the fine-to-coarse restriction and the plain
sub-grid bed (no reference counterpart: oceano has no closure)."""
import numpy as np
import pytest

import mlp_oracle as mo
from oceano_amd import closure
from oceano_amd.mesh import Mesh


def _params(rng, n_in, n_out, scale=0.3):
    p = np.zeros(mo.NPARAM)
    for (W, b), (k, n) in zip(mo.unpack(p), [(n_in, mo.HID), (mo.HID, mo.HID),
                                              (mo.HID, mo.HID), (mo.HID, n_out)]):
        W[:k, :n] = rng.uniform(-scale, scale, (k, n))
        b[:n] = rng.uniform(-0.1, 0.1, n)
    return p


def test_bf16_round_to_nearest_even():
    one = np.float32(1.0)
    ulp = 2.0 ** -7  # BF16 spacing at 1
    vals = np.array([1.0, 1.0 + ulp / 2, 1.0 + 1.5 * ulp, 1.0 + ulp / 4, -3.0 - ulp,
                     1.0 + ulp / 2 + 2.0 ** -20], np.float32)
    got = mo.to_bf16(vals)
    # ties go to the even mantissa: 1 + ulp/2 -> 1, 1 + 1.5 ulp -> 1 + 2 ulp
    assert got[0] == one and got[1] == one and got[2] == np.float32(1.0 + 2 * ulp)
    assert got[3] == one and got[5] == np.float32(1.0 + ulp)
    # -3 - 2^-7 sits halfway between -3 and -3 - 2^-6 (spacing 2^-6 at 3): even is -3
    assert got[4] == np.float32(-3.0)


def test_gelu_derivative():
    z = np.linspace(-4, 4, 41)
    h = 1e-6
    fd = (mo.gelu(z + h) - mo.gelu(z - h)) / (2 * h)
    assert np.abs(fd - mo.gelu_prime(z)).max() < 1e-8
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15  # Phi(1)


def test_backward_matches_finite_differences():
    """FP64 mode: every parameter block's gradient against central differences of the
    loss (a transposed product, a wrong activation index or a missing bias term fails)"""
    rng = np.random.default_rng(3)
    n_in, n_out, n = 7, 5, 6
    p = _params(rng, n_in, n_out)
    x = rng.standard_normal((n, n_in))
    t = rng.standard_normal((n, n_out))
    norm = (rng.standard_normal(n_in) * 0.1, 1 + rng.random(n_in),
            rng.standard_normal(n_out) * 0.1, 1 + rng.random(n_out))
    net = mo.Network(p, n_in, n_out, norm=norm, quantize=False)
    loss, g = net.loss_and_gradient(x, t)
    blocks = [(mo.W1, n_in, mo.HID), (mo.B1, 1, mo.HID), (mo.W2, mo.HID, mo.HID),
              (mo.B3, 1, mo.HID), (mo.W4, mo.HID, n_out), (mo.B4, 1, n_out)]
    h = 1e-6
    for off, k, ncol in blocks:
        width = mo.OUT if off == mo.W4 else mo.HID
        for _ in range(6):
            i, j = rng.integers(k), rng.integers(ncol)
            idx = off + i * width + j if k > 1 else off + j
            pp, pm = p.copy(), p.copy()
            pp[idx] += h
            pm[idx] -= h
            lp = mo.Network(pp, n_in, n_out, norm=norm, quantize=False).loss_and_gradient(x, t)[0]
            lm = mo.Network(pm, n_in, n_out, norm=norm, quantize=False).loss_and_gradient(x, t)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[idx]) < 1e-7 * max(1.0, abs(fd)), (off, i, j, fd, g[idx])
    # padded inputs / outputs get no gradient
    W1 = g[mo.W1:mo.B1].reshape(mo.IN, mo.HID)
    assert np.all(W1[n_in:] == 0)
    W4 = g[mo.W4:mo.B4].reshape(mo.HID, mo.OUT)
    assert np.all(W4[:, n_out:] == 0) and np.all(g[mo.B4 + n_out:] == 0)


def test_quantized_forward_close_to_fp64():
    rng = np.random.default_rng(4)
    p = _params(rng, 30, 18)
    x = rng.standard_normal((40, 30))
    y64 = mo.Network(p, 30, 18, quantize=False).predict(x)
    y16 = mo.Network(p, 30, 18, quantize=True).predict(x)
    assert np.linalg.norm(y16 - y64) < 2e-2 * np.linalg.norm(y64)


def test_modal_basis():
    for p in (1, 2):
        Vi, V = closure.nodal_to_modal(p)
        n1 = p + 1
        nodes = closure.gauss_lobatto_nodes(n1)
        xq, wq = closure.gauss(n1 + 1)
        # orthonormality of the modes in the L2(0,1)^2 product
        L = closure.legendre_modes(n1, xq)
        M1 = np.einsum("q,qa,qb->ab", wq, L, L)
        assert np.abs(M1 - np.eye(n1)).max() < 1e-14
        # nodal values of mode (a, b) are V[:, a + n1 b]; a bilinear function's modes
        f = lambda x, y: 1.0 + 2.0 * x - 3.0 * y + 0.5 * x * y
        nod = np.array([f(nodes[i], nodes[j]) for j in range(n1) for i in range(n1)])
        m = Vi @ nod
        # mean = coefficient of mode (0,0) (psi_0 = 1)
        assert abs(m[0] - (1.0 + 1.0 - 1.5 + 0.125)) < 1e-14
        assert np.abs(V @ m - nod).max() < 1e-14


def test_restriction_exact_on_coarse_polynomials():
    p, f = 2, 4
    cm = Mesh.rectangle(0, 25, 0, 1, 10, 2)
    fm = Mesh.rectangle(0, 25, 0, 1, 40, 8)
    torch = pytest.importorskip("torch")
    R = closure.Restriction(cm, fm, p, f, device="cpu")
    nodes = closure.gauss_lobatto_nodes(p + 1)

    def nodal(m, fun):
        g = m.geom
        out = np.zeros((g.shape[0], (p + 1) ** 2))
        for c in range(g.shape[0]):
            X = g[c, 0] + g[c, 2] * nodes
            Y = g[c, 1] + g[c, 3] * nodes
            out[c] = [fun(X[i], Y[j]) for j in range(p + 1) for i in range(p + 1)]
        return out
    fun = lambda x, y: 1 + 0.3 * x - 0.01 * x * x + 2 * y * y - 0.5 * x * y
    uc = R(torch.tensor(nodal(fm, fun).ravel()), 1).numpy().reshape(-1, (p + 1) ** 2)
    assert np.abs(uc - nodal(cm, fun)).max() < 1e-12
    # a function discontinuous across fine cells: its projection keeps the cell mean
    rng = np.random.default_rng(1)
    uf = rng.standard_normal((fm.n_cells, (p + 1) ** 2))
    uc = R(torch.tensor(uf.ravel()), 1).numpy().reshape(-1, (p + 1) ** 2)
    Vi, _ = closure.nodal_to_modal(p)
    mean_c = uc @ Vi[0]
    mean_f = (uf @ Vi[0]).reshape(8, 40)
    # coarse cell (ix, iy) holds fine cells ix*4.., iy*4..
    want = mean_f.reshape(2, 4, 10, 4).mean(axis=(1, 3))
    ixc = np.round(cm.geom[:, 0] / 2.5).astype(int)
    iyc = np.round(cm.geom[:, 1] / 0.5).astype(int)
    assert np.abs(mean_c - want[iyc, ixc]).max() < 1e-13


def test_plain_bed_is_the_cellwise_projection():
    S = closure.SubgridChannel(1000, nx=250, factor=8)
    xq, wq = closure.gauss(48)
    for ix in (0, 85, 100, 117, 249):
        x = (ix + xq) * S.dx
        b, bp = S.bed(x), S.plain_bed(x)
        # same mean and first moments up to degree p (orthogonality of the residual)
        P = closure.legendre_modes(3, xq)
        assert np.abs(((b - bp) * wq) @ P).max() < 1e-12
        assert abs((b * wq).sum() - S.stats[ix, 0]) < 1e-14
    assert S.stats[:, 1].max() > 1e-5  # the ridges are sub-grid: nonzero variance
