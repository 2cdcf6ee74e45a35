"""Work-layout options (cell-slot order, batch-to-XCD mapping, work items per workgroup)
change only the speed: an RK4 step under every option set the autotuner tries is bitwise the
same, in both precisions, among the candidates with the same paired_items setting (pairing
sums the face integrals of hanging faces in another order: tests/test_gpu_paired.py), and
the autotuner returns one of its candidates with a timing for each.

The slot orders also change how the FP32 stage deals the second subfaces of coarse faces
over the lanes of a wave (one item per lane, several passes when a wave holds more than 64
such items): bitwise equality across the orders checks that dealing, and the test confirms
on the host that one of the orders gives some wave more than 64 items."""
import numpy as np
import pytest
import torch

from helpers import to_host
from oceano_amd import problems, tune
from oceano_amd.integrator import DeviceIntegrator

pytestmark = pytest.mark.gpu


def _max_second_subfaces_per_wave(op, mesh):
    """largest number of (cell, coarse face) pairs in one 64-cell batch of the operator's
    slot order (the cell of each slot read back through pack)"""
    N, nd = mesh.n_cells, op.dofs_per_cell
    ids = op.pack(np.repeat(np.arange(N, dtype=np.float64), nd), 1).cpu().numpy()
    slots = ids.reshape(-1, nd, 64)[:, 0, :].ravel()[:N].astype(np.int64)
    ncoarse = np.zeros(N, np.int64)
    for F in mesh.inner:
        for c, f, sub in ((F[0], F[1], F[2]), (F[3], F[4], F[5])):
            if sub >= 0:
                ncoarse[c] += 1
    # each coarse face appears once per subface in the face list: count faces, not subfaces
    ncoarse //= 2
    return max(int(ncoarse[slots[b:b + 64]].sum()) for b in range(0, N, 64))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_options_do_not_change_results(cuda, precision):
    P = problems.vortex(degree=3, refine=1, refine_fraction=0.3, seed=12)
    print("[tune] %s, %d cells" % (precision, P.mesh.n_cells), flush=True)
    ref = {}
    cands = tune.DEFAULT_CANDIDATES + [
        {"slot_chunk": 4, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 0},
        {"slot_chunk": 4, "xcd_swizzle": 1, "waves_per_block": 1, "paired_items": 1}]
    most = 0
    try:
        for opts in cands:
            tune.apply_options(opts)
            op, _, w_h, w_q = problems.setup_operator(P, precision=precision)
            most = max(most, _max_second_subfaces_per_wave(op, P.mesh))
            rk = DeviceIntegrator(op)
            for _ in range(2):
                rk.step(1e-4, w_h, w_q)
            torch.cuda.synchronize()
            out = (to_host(op, w_h, 1), to_host(op, w_q, 2))
            r = ref.setdefault(opts["paired_items"], out)
            assert (out[0] == r[0]).all() and (out[1] == r[1]).all(), opts
    finally:
        tune.apply_options(tune.LIBRARY_DEFAULTS)
    print("  most second subfaces in one wave: %d" % most)
    assert most > 64


def test_autotune_picks_a_candidate(cuda):
    P = problems.vortex(degree=2, refine=2)
    print("[tune] autotune p=2, %d cells" % P.mesh.n_cells, flush=True)
    try:
        chosen, table = tune.autotune(P, "mixed", dt=1e-4, steps=4, repeats=1)
    finally:
        tune.apply_options(tune.LIBRARY_DEFAULTS)
    assert chosen in tune.DEFAULT_CANDIDATES
    assert len(table) == len(tune.DEFAULT_CANDIDATES)
    assert all(t > 0 for _, t in table)
    assert min(table, key=lambda e: e[1])[0] == chosen
