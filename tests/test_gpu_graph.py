"""hcg_graph_create / hcg_graph_launch: the device-resident launch captured once as a HIP graph
and replayed.  Every replay returns exactly what hcg_track_batch_device returns for the same
inputs (records, endpoints, last points, valuations bit for bit, and the launch counters), with
the program's own eager launches and synchronizations between the replays; after the caller
rewrites the target parameters in place, the next replay equals an eager launch on the new
targets."""
from pathlib import Path

import numpy as np
import pytest
import torch

from hcgpu import learned
from hcgpu.configs import cfg2_problem
from hcgpu.device import DeviceBatch
from hcgpu.homotopies import ParameterHomotopy
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu
STARTS = Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy"


def _bits(out):
    rec, sol, last, val = out
    return [np.ascontiguousarray(a).view(np.uint8).copy() for a in (rec, sol, last, val)]


def _equal(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("layout,use_learned,ntargets", [("lane", 0, 4), ("lane", 1, 4), ("group", 0, 1)])
def test_graph_replays_equal_eager_launches(gpu_ctx, layout, use_learned, ntargets):
    F, p0, Q = cfg2_problem(2 * ntargets)
    S = np.load(STARTS)
    H = ParameterHomotopy(F, p0, Q[:ntargets], lift_affine=True)
    prog = gpu_ctx.program(H, layout=layout)
    if use_learned:
        learned.load(gpu_ctx, learned.saved_weights("param10_step_mlp"))
    opts = make_options(learned_step=use_learned)
    starts = np.tile(S, (ntargets, 1))
    groups = np.repeat(np.arange(ntargets, dtype=np.int32), len(S))
    print(f"graph test: layout {layout}, learned {use_learned}, {len(starts)} paths", flush=True)

    eager = DeviceBatch(gpu_ctx, prog, starts, H.p, H.q, groups)
    eager.launch(opts)
    ref = _bits(eager.results())
    ref_counters = gpu_ctx.last_counters()

    replayed = DeviceBatch(gpu_ctx, prog, starts, H.p, H.q, groups)
    g = replayed.graph(opts)
    try:
        for r in range(3):
            ms, counters = g.launch()
            torch.cuda.synchronize()
            assert _equal(_bits(replayed.results()), ref), f"replay {r} differs from the eager launch"
            assert counters == ref_counters
            assert ms > 0
            eager.launch(opts)  # eager work on the same context between replays
            torch.cuda.synchronize()
            assert _equal(_bits(eager.results()), ref)
        # new targets written into the captured buffers
        Q2 = H.target_matrix(Q[ntargets:2 * ntargets])  # in the tapes' (lifted) parameter space
        replayed.set_targets(Q2)
        torch.cuda.synchronize()
        g.launch()
        fresh = DeviceBatch(gpu_ctx, prog, starts, H.p, Q2, groups)
        fresh.launch(opts)
        got, want = _bits(replayed.results()), _bits(fresh.results())
        assert _equal(got, want), "replay on rewritten targets differs from an eager launch on them"
        assert not np.array_equal(got[1], ref[1])  # the new targets moved the endpoints
        rec = replayed.results()[0]
        assert (rec["code"] == 1).all()
    finally:
        g.close()
