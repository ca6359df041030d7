"""Edge cases of the batch entry points on the GPU: empty batches, batches smaller than a wave
or a lane group, ragged multi-system batches, and out-of-range parameter groups in a
device-resident batch (rejected per path on the device, as the header documents).  A path's
result never depends on what else is in its batch, so every subset equals the corresponding
rows of the full batch bit for bit."""
from pathlib import Path

import numpy as np
import pytest
import torch  # noqa: F401  (before libhcgpu.so: one HIP runtime)

from hcgpu.abi import HcgpuError, RETURN_CODES
from hcgpu.configs import cfg2_problem, dense_system
from hcgpu.device import DeviceBatch
from hcgpu.gpu import track_batch, track_multi
from hcgpu.homotopies import ParameterHomotopy, total_degree
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu
STARTS = Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy"


def _rows_equal(sub, full, idx):
    assert np.array_equal(sub.code, full.code[idx])
    assert np.array_equal(sub.accepted_steps, full.accepted_steps[idx])
    assert np.array_equal(sub.solution.view(np.uint8), full.solution[idx].view(np.uint8))


@pytest.mark.parametrize("layout", ["lane", "group"])
def test_empty_and_ragged_batches(gpu_ctx, layout):
    H, S = total_degree(dense_system(4, 3, seed=100), seed=101)
    prog = gpu_ctx.program(H, layout=layout)
    print(f"edges: layout {layout}", flush=True)
    empty = track_batch(prog, np.zeros((0, H.n), dtype=complex), make_options())
    assert len(empty.code) == 0 and empty.solution.shape == (0, H.n)
    full = track_batch(prog, S, make_options())
    for idx in ([0], list(range(15)), list(range(17)), list(range(63)), list(range(65)), [80, 3, 41]):
        _rows_equal(track_batch(prog, S[idx], make_options()), full, idx)


def test_multi_system_batch_with_an_empty_system(gpu_ctx):
    Ha, Sa = total_degree(dense_system(4, 3, seed=100), seed=101)
    Hb, Sb = total_degree(dense_system(3, 2, seed=7), seed=8)
    multi = gpu_ctx.program_multi([Ha, Hb])
    fa = track_batch(gpu_ctx.program(Ha), Sa, make_options())
    outs = track_multi(multi, [Sa, np.zeros((0, Hb.n), dtype=complex)], make_options())
    assert len(outs[1].code) == 0
    _rows_equal(outs[0], fa, list(range(len(Sa))))
    outs = track_multi(multi, [Sa[:5], Sb], make_options())
    _rows_equal(outs[0], fa, list(range(5)))
    fb = track_batch(gpu_ctx.program(Hb), Sb, make_options())
    _rows_equal(outs[1], fb, list(range(len(Sb))))


def test_out_of_range_groups(gpu_ctx):
    """The host entry refuses out-of-range groups; the device entry (which cannot read its
    inputs on the host) marks those paths terminated_unknown and tracks the others."""
    F, p0, Q = cfg2_problem(2)
    S = np.load(STARTS)[:64]
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    prog = gpu_ctx.program(H)
    groups = np.zeros(len(S), dtype=np.int32)
    groups[[3, 40]] = [2, -1]
    with pytest.raises(HcgpuError, match="out of range"):
        track_batch(prog, S, make_options(), group_of_path=groups)
    d = DeviceBatch(gpu_ctx, prog, S, H.p, H.q, groups)
    d.launch(make_options())
    rec = d.results()[0]
    unknown = RETURN_CODES.index("terminated_unknown")
    assert rec["code"][3] == unknown and rec["code"][40] == unknown
    ok = np.setdiff1d(np.arange(len(S)), [3, 40])
    ref = track_batch(prog, S[ok], make_options(), group_of_path=np.zeros(len(ok), dtype=np.int32))
    assert np.array_equal(rec["code"][ok], ref.code)


def test_graph_refuses_an_empty_batch(gpu_ctx):
    H, S = total_degree(dense_system(4, 3, seed=100), seed=101)
    d = DeviceBatch(gpu_ctx, gpu_ctx.program(H), S[:1])
    d.b.npaths = 0
    with pytest.raises(HcgpuError, match="at least one path"):
        d.graph(make_options())
