"""Mesh builder geometry checks and the C-ABI library surface (CPU only)."""
import ctypes
import os
import re

import numpy as np
import pytest

from oceano_amd import _capi
from oceano_amd.mesh import Mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = set()
    for h in ("oceano_dg.h", "oceano_mlp.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(oc_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding declares every one of them
    assert set(names) == set(_capi.SIGNATURES), set(names) ^ set(_capi.SIGNATURES)


def _face_points(geom, c, f, t0, t1):
    x0, y0, hx, hy = geom[c]
    if f == 0:
        return (x0, y0 + t0 * hy), (x0, y0 + t1 * hy)
    if f == 1:
        return (x0 + hx, y0 + t0 * hy), (x0 + hx, y0 + t1 * hy)
    if f == 2:
        return (x0 + t0 * hx, y0), (x0 + t1 * hx, y0)
    return (x0 + t0 * hx, y0 + hy), (x0 + t1 * hx, y0 + hy)


def _check_faces(m, period=(0.0, 0.0)):
    g = m.geom
    for cm, fm, sm, cp, fp, sp in m.inner:
        assert fp == fm - 1 and fm in (1, 3)
        tm = (0.0, 1.0) if sm < 0 else (0.5 * sm, 0.5 * sm + 0.5)
        tp = (0.0, 1.0) if sp < 0 else (0.5 * sp, 0.5 * sp + 0.5)
        a = np.array(_face_points(g, cm, fm, *tm))
        b = np.array(_face_points(g, cp, fp, *tp))
        d = a - b
        for k in range(2):
            if period[k]:
                d[:, k] -= period[k] * np.round(d[:, k] / period[k])
        assert np.abs(d).max() < 1e-12
        assert abs(m.level[cm] - m.level[cp]) <= 1
        assert (sm >= 0) == (m.level[cm] < m.level[cp])
        assert (sp >= 0) == (m.level[cp] < m.level[cm])


def test_c2_mesh():
    """20x10 refined 4x plus 30% seeded refinement: 97,280 cells (2:1 holds)."""
    m = Mesh.rectangle(0, 2, 0, 1, 20, 10, periodic=(True, True)).refine_global(4)
    assert m.n_cells == 51200
    m.refine_fraction(0.3, 20261202)
    assert m.n_cells == 51200 + 3 * 15360
    assert m.boundary.shape[0] == 0
    areas = m.geom[:, 2] * m.geom[:, 3]
    assert abs(areas.sum() - 2.0) < 1e-12
    _check_faces(m, period=(2.0, 1.0))


def test_faces_cover_every_cell_side():
    m = Mesh.rectangle(0, 2, 0, 1, 6, 3, boundary_ids=(1, 2, 0, 0)).refine_global(1)
    m.refine_fraction(0.4, 3)
    _check_faces(m)
    cover = np.zeros((m.n_cells, 4))
    for cm, fm, sm, cp, fp, sp in m.inner:
        cover[cm, fm] += 0.5 if sm >= 0 else 1.0
        cover[cp, fp] += 0.5 if sp >= 0 else 1.0
    for c, f, b in m.boundary:
        cover[c, f] += 1.0
        x0, y0, hx, hy = m.geom[c]
        expect = {0: 1, 1: 2, 2: 0, 3: 0}[f]
        assert b == expect
    assert np.all(cover == 1.0)


def test_refinement_is_reproducible():
    a = Mesh.rectangle(0, 2, 0, 1, 10, 5).refine_global(1).refine_fraction(0.3, 7)
    b = Mesh.rectangle(0, 2, 0, 1, 10, 5).refine_global(1).refine_fraction(0.3, 7)
    c = Mesh.rectangle(0, 2, 0, 1, 10, 5).refine_global(1).refine_fraction(0.3, 8)
    assert np.array_equal(a.geom, b.geom) and not np.array_equal(a.geom, c.geom)


def test_unbalanced_refinement_is_rejected():
    m = Mesh.rectangle(0, 1, 0, 1, 2, 2).refine_fraction(1.0, 1)
    # refining one fine cell once more next to a coarse cell violates 2:1 only if levels
    # differ by 2; refining a single cell of a uniform mesh twice does that
    m2 = Mesh.rectangle(0, 1, 0, 1, 4, 4)
    m2.refine_fraction(1.0 / 16, 1)
    with pytest.raises(_capi.OceanoError) as e:
        # refine the four children of the refined cell's first child only: pick a seed
        # whose draws fall in the fine block until it fails
        for seed in range(200):
            m3 = Mesh.rectangle(0, 1, 0, 1, 4, 4)
            m3.refine_fraction(1.0 / 16, 1)
            m3.refine_fraction(1.0 / 19, seed)
    assert e.value.code == _capi.OC_ERR_MESH


def test_vortex_and_channel_functions():
    from oceano_amd import problems
    # vortex: depth h0 outside the radius, h0 - amplitude at the centre; cyclostrophic
    # balance g dh/dr = u_theta^2 / r (the vortex is a steady solution in the moving frame)
    prm = [1.0, 1.0, 0.0, 0.1, 0.35, 0.5, 0.5, 9.81, 0.0, 0.0]
    out = problems.vortex_exact(np.array([[0.5, 0.5], [0.5 + 0.36, 0.5]]), 0.0, prm)
    # the formula's decimal constants (0.091145833 for 35/384, ...) leave 4e-9
    assert abs(out[0, 0] - 0.9) < 1e-8 and abs(out[1, 0] - 1.0) < 1e-15
    r = np.linspace(0.02, 0.33, 25)
    eps = 1e-6
    pts = np.stack([0.5 + r, 0.5 + 0 * r], 1)
    h = problems.vortex_exact(pts, 0.0, prm)
    hp = problems.vortex_exact(pts + [eps, 0], 0.0, prm)
    hm = problems.vortex_exact(pts - [eps, 0], 0.0, prm)
    dhdr = (hp[:, 0] - hm[:, 0]) / (2 * eps)
    utheta = h[:, 2]  # v = (x - x0) omega at y = y0
    assert np.allclose(9.81 * dhdr, utheta ** 2 / r, rtol=1e-5, atol=1e-8)
    # translation: at time t the pattern moved by u_inf t
    a = problems.vortex_exact(pts, 0.0, prm)
    b = problems.vortex_exact(pts + [0.1, 0], 0.1, prm)
    assert np.allclose(a, b, atol=1e-14)
    # channel bed: parabolic bump of height 0.2 at x = 10 plus bumps within [0.02, 0.05]
    d = problems.channel_data(np.array([[10.0, 0.5], [5.0, 0.5]]), [0.02, 0, 0, 0, 1, 1, 0, 1], 7)
    assert abs(d[0, 0] + 0.2) < 1e-15 and d[1, 0] == 0.0 and d[0, 1] == 0.02


# ------------------------------------------------------ Gmsh decks of the reference ----
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_read_vortex_deck():
    """tests/shallowWaterVortex/shallowWaterVortex_20x10cells.msh (a fixture of the
    reference, copied as data): the 20 x 10 grid on [0,2]x[0,1] (node coordinates within
    ~1e-12 of the grid lines), boundary id 1 at x = 0 (curve 7), 2 at x = 2 (curve 5) and 0
    on the y sides (curves 4, 6) -- the ids the vortex case's boundary conditions use."""
    m = Mesh.read_msh(os.path.join(GOLDEN, "shallowWaterVortex_20x10cells.msh"))
    r = Mesh.rectangle(0, 2, 0, 1, 20, 10, boundary_ids=(1, 2, 0, 0))
    assert m.n_cells == 200
    assert np.array_equal(m.geom, r.geom)
    assert np.array_equal(m.inner, r.inner) and np.array_equal(m.boundary, r.boundary)
    ids = {(int(f), int(b)) for _, f, b in m.boundary}
    assert ids == {(0, 1), (1, 2), (2, 0), (3, 0)}
    # the C1 configuration: the deck refined three times, periodic
    c1 = Mesh.read_msh(os.path.join(GOLDEN, "shallowWaterVortex_20x10cells.msh"),
                       periodic=(True, True)).refine_global(3)
    assert c1.n_cells == 12800 and c1.boundary.shape[0] == 0


def test_read_channel_deck():
    """tests/channelFlow/channelFlow_10x2cells.msh: 10 x 2 cells on [0,100]x[0,1]."""
    m = Mesh.read_msh(os.path.join(GOLDEN, "channelFlow_10x2cells.msh"))
    assert m.n_cells == 20
    assert np.allclose(m.geom[:, 2], 10.0) and np.allclose(m.geom[:, 3], 0.5)
    sides = {}
    for _, f, b in m.boundary:
        sides.setdefault(int(f), set()).add(int(b))
    assert all(len(v) == 1 for v in sides.values()) and set(sides) == {0, 1, 2, 3}


def test_read_msh_errors(tmp_path):
    bad = tmp_path / "tri.msh"
    bad.write_text("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n$Nodes\n1 3 1 3\n2 1 0 3\n1\n2\n3\n"
                   "0 0 0\n1 0 0\n0 1 0\n$EndNodes\n$Elements\n1 1 1 1\n2 1 2 1\n1 1 2 3\n"
                   "$EndElements\n")
    with pytest.raises(RuntimeError, match="element type 2"):
        Mesh.read_msh(str(bad))
    with pytest.raises(RuntimeError, match="cannot open"):
        Mesh.read_msh(str(tmp_path / "missing.msh"))
