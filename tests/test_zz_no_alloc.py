"""oc_dg_perform_stage_hydro_noalloc (the stage for callers that cannot allocate) against
oc_dg_perform_stage_hydro through the C ABI: one RK4 step from the same state gives
bitwise the same state and k vectors, on a larger operator and then a smaller one (the
second call reuses the shared scratch the first left behind); the FP32 path at p = 4 uses
the scratch's deferred-solve masks.  An operator with more batches than the scratch holds
is refused."""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import make_pair, to_dev
from oceano_amd import _capi, problems
from oceano_amd.integrator import butcher_rows
from oceano_amd.mesh import Mesh
from oceano_amd.operator import OceanoOperator

pytestmark = pytest.mark.gpu


def _step(op, entry, h, q, rows):
    S = len(rows)
    w_h, w_q = to_dev(op, h, 1), to_dev(op, q, 2)
    ki_h = [op.initialize_vector(0) for _ in range(S + 1)]
    ki_q = [op.initialize_vector(1) for _ in range(S + 1)]
    r_h = [op.initialize_vector(0) for _ in range(2)]
    r_q = [op.initialize_vector(1) for _ in range(2)]
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    kh = (C.c_void_p * (S + 1))(*[k.data_ptr() for k in ki_h])
    kq = (C.c_void_p * (S + 1))(*[k.data_ptr() for k in ki_q])
    for s in range(S):
        ri = (w_h, w_q) if s == 0 else (r_h[s % 2], r_q[s % 2])
        fac = np.ascontiguousarray(rows[s][:s + 1], np.float64)
        rc = entry(op._h, s, 0.0, _capi.dptr(fac), s + 1, C.c_void_p(ri[0].data_ptr()),
                   C.c_void_p(ri[1].data_ptr()), None, kh, kq, S + 1,
                   C.c_void_p(w_h.data_ptr()), C.c_void_p(w_q.data_ptr()),
                   C.c_void_p(r_h[(s + 1) % 2].data_ptr()),
                   C.c_void_p(r_q[(s + 1) % 2].data_ptr()), 0, stream)
        _capi.check(rc, "stage")
    torch.cuda.synchronize()
    return [w_h, w_q] + ki_h[1:] + ki_q[1:]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_stage_matches_stage(cuda, precision):
    lib = _capi.lib()
    for degree, refine in ((4, 2), (3, 1)):   # the larger operator first
        P = problems.vortex(degree=degree, refine=refine, refine_fraction=0.3, seed=9)
        print("[noalloc] %s p=%d cells=%d" % (precision, degree, P.mesh.n_cells), flush=True)
        op, o, h, q = make_pair(P, precision=precision)
        hmin = float(np.min(P.mesh.geom[:, 2:]))
        rows, _ = butcher_rows("stage_4_order_4", 0.1 * hmin / (8.0 * (2 * degree + 1)))
        ref = _step(op, lib.oc_dg_perform_stage_hydro, h, q, rows)
        out = _step(op, lib.oc_dg_perform_stage_hydro_noalloc, h, q, rows)
        for a, b in zip(ref, out):
            assert torch.equal(a, b)
        assert float((ref[0] - to_dev(op, h, 1)).abs().max()) > 0  # the step moved


def test_noalloc_refuses_too_many_batches(cuda):
    n = _capi.OC_NOALLOC_MAX_BATCHES * 64 + 1
    m = Mesh.rectangle(0.0, 2.0, 0.0, 1.0, 2 * 725, 725, periodic=(True, True))
    assert m.n_cells >= n
    op = OceanoOperator(m, 1)
    S = 4
    v1, v2 = op.initialize_vector(0), op.initialize_vector(1)
    kh = (C.c_void_p * (S + 1))(*[v1.data_ptr()] * (S + 1))
    kq = (C.c_void_p * (S + 1))(*[v2.data_ptr()] * (S + 1))
    fac = np.zeros(1)
    rc = _capi.lib().oc_dg_perform_stage_hydro_noalloc(
        op._h, 0, 0.0, _capi.dptr(fac), 1, C.c_void_p(v1.data_ptr()),
        C.c_void_p(v2.data_ptr()), None, kh, kq, S + 1, C.c_void_p(v1.data_ptr()),
        C.c_void_p(v2.data_ptr()), C.c_void_p(v1.data_ptr()), C.c_void_p(v2.data_ptr()), 0,
        None)
    assert rc == _capi.OC_ERR_ARGUMENT
    assert b"exceed the scratch" in _capi.lib().oc_last_error()
