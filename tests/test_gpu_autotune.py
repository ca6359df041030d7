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
