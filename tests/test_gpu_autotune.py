"""The autotuner times build/launch candidates on a sample and caches the fastest; every
candidate gives the same results (tuning never changes the arithmetic)."""
import numpy as np
import pytest

import systems
from hcgpu import autotune
from hcgpu.gpu import track_batch
from hcgpu.homotopies import total_degree
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu


def test_autotune_picks_a_candidate_and_results_do_not_change(gpu_ctx, tmp_path, monkeypatch):
    monkeypatch.setattr(autotune, "_cache_path", lambda: tmp_path / "autotune.json")
    H, S = total_degree(systems.winding_family(4), seed=3)
    S = np.tile(S, (64, 1))
    cands = ((1, 4), (2, 8), (4, 16))
    best = autotune.tune(gpu_ctx, H, S, candidates=cands)
    assert (best["waves_per_eu"], best["blocks_per_cu"]) in cands
    assert len(best["table"]) == 3
    assert autotune.tuned_choice(gpu_ctx, H)["waves_per_eu"] == best["waves_per_eu"]
    outs = []
    for w, bpc in cands:
        o = make_options(blocks_per_cu=bpc)
        outs.append(track_batch(gpu_ctx.program(H, waves_per_eu=w), S, o))
    for g in outs[1:]:
        assert (g.code == outs[0].code).all()
        np.testing.assert_array_equal(g.solution, outs[0].solution)


def test_default_program_uses_the_tuned_choice(gpu_ctx, tmp_path, monkeypatch):
    """After tuning (layout among the candidates), GpuContext.program() builds the tuned
    layout and register budget and track_batch launches the tuned workgroups per CU; the
    results equal the default build's bit for bit."""
    monkeypatch.setattr(autotune, "_cache_path", lambda: tmp_path / "autotune.json")
    H, S = total_degree(systems.winding_family(4), seed=5)
    S = np.tile(S, (64, 1))
    cands = autotune.candidates_for(H.n, lane=((1, 4), (2, 8)), group=(2,))
    best = autotune.tune(gpu_ctx, H, S, candidates=cands)
    print("tuned", {k: best[k] for k in ("layout", "waves_per_eu", "blocks_per_cu")}, best["table"])
    assert {r["layout"] for r in best["table"]} == {"lane", "group"}
    prog = gpu_ctx.program(H)
    assert (prog.layout, prog.blocks_per_cu) == (best["layout"], best["blocks_per_cu"])
    if best["layout"] == "lane":
        assert prog.waves_per_eu == best["waves_per_eu"]
    out = track_batch(prog, S)
    ref = track_batch(gpu_ctx.program(H, waves_per_eu=2, layout="lane"), S)
    np.testing.assert_array_equal(out.code, ref.code)
    np.testing.assert_array_equal(out.solution, ref.solution)
