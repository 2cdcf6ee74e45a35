"""Work-layout options (cell-slot order, batch-to-XCD mapping) change only the speed: an RK4
step under every option set the autotuner tries is bitwise the same, in both precisions,
and the autotuner returns one of its candidates with a timing for each."""
import pytest
import torch

from helpers import to_host
from oceano_amd import problems, tune
from oceano_amd.integrator import DeviceIntegrator

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_options_do_not_change_results(cuda, precision):
    P = problems.vortex(degree=3, refine=1, refine_fraction=0.3, seed=12)
    print("[tune] %s, %d cells" % (precision, P.mesh.n_cells), flush=True)
    ref = None
    cands = tune.DEFAULT_CANDIDATES + [{"slot_chunk": 4, "xcd_swizzle": 1}]
    try:
        for opts in cands:
            tune.apply_options(opts)
            op, _, w_h, w_q = problems.setup_operator(P, precision=precision)
            rk = DeviceIntegrator(op)
            for _ in range(2):
                rk.step(1e-4, w_h, w_q)
            torch.cuda.synchronize()
            out = (to_host(op, w_h, 1), to_host(op, w_q, 2))
            if ref is None:
                ref = out
            else:
                assert (out[0] == ref[0]).all() and (out[1] == ref[1]).all(), opts
    finally:
        tune.apply_options(tune.LIBRARY_DEFAULTS)


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
