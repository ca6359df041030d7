"""Per-configuration parity of the GPU tracker, each benchmark configuration at its own size:

* dense n = 4 cubics, 81 total-degree paths;
* dense n = 8 cubics, 6,561 total-degree paths in ONE launch;
* configs[2]+[3]: n = 10 sparse quadrics, 1,024 starts x 256 targets = 262,144 paths in one
  launch, plain and learned-step;
* shared-factor systems n = 4, 6, 8, 12 in one multi-system launch, learned-step.

For each: return codes and singular flags equal the oracle's on every path the oracle runs
(all of them for configs[0], [1]; all of n = 4, 6, 8 and the first 8,192 of n = 12 for [4];
2 of the 256 targets for [2]), counts of nonsingular / singular / at-infinity endpoints equal,
and endpoints of nonsingular successes within 1e-10 relative of (a) the oracle's and (b) their
200-bit Newton refinement on the expanded polynomial system, on a sample of 256 endpoints
(tests/refine.py).  Each test prints one CONFIG_TABLE line with the figures.
"""
import numpy as np
import pytest

import oracle
from hcgpu.configs import CFG4_SEEDS, cfg2_problem, cfg4_problem, dense_system, shared_factor_system
from hcgpu.gpu import track_batch, track_multi
from hcgpu.homotopies import ParameterHomotopy, total_degree
from hcgpu.solve import make_options
from refine import refine, rel_error

pytestmark = pytest.mark.gpu
TOL = 1e-10
SAMPLE = 256


def _counts(code, singular):
    code, singular = np.asarray(code), np.asarray(singular, dtype=bool)
    return dict(nonsingular=int(((code == 1) & ~singular).sum()), singular=int(((code == 1) & singular).sum()),
                at_infinity=int(np.isin(code, (2, 3)).sum()), other=int((~np.isin(code, (1, 2, 3))).sum()))


def _same_as_oracle(g_code, g_sing, g_sol, o):
    assert list(g_code) == list(o.code)
    assert list(g_sing) == list(o.singular)
    ok = (o.code == 1) & (~o.singular)
    if not ok.any():
        return 0.0
    rel = rel_error(g_sol[ok], o.solution[ok])
    assert rel.max() <= TOL, rel.max()
    return float(rel.max())


def _refined_sample(F, X, rng, size=SAMPLE):
    idx = rng.choice(len(X), size=size, replace=False) if len(X) > size else np.arange(len(X))
    R, last = refine(F, X[idx])
    assert (last < 1e-30).all(), last.max()
    rel = rel_error(X[idx], R)
    assert rel.max() <= TOL, rel.max()
    return float(rel.max()), len(idx)


@pytest.mark.parametrize("cfg,n,seeds", [(0, 4, (100, 101)), (1, 8, (1, 2))])
def test_dense_total_degree_one_launch(gpu_ctx, cfg, n, seeds):
    F = dense_system(n, 3, seed=seeds[0])
    H, S = total_degree(F, seed=seeds[1])
    g = track_batch(gpu_ctx.program(H), S, make_options())
    ms = gpu_ctx.last_kernel_ms()
    o = oracle.track_batch(H, S, make_options())
    rel_o = _same_as_oracle(g.code, g.singular, g.solution, o)
    assert _counts(g.code, g.singular) == _counts(o.code, o.singular)
    fin = (g.code == 1) & (~g.singular)
    rel_r, m = _refined_sample(F, g.solution[fin], np.random.default_rng(cfg))
    print(f"CONFIG_TABLE cfg{cfg}: {len(S)} paths in one launch ({ms:.1f} ms); GPU {_counts(g.code, g.singular)}; "
          f"oracle identical codes on all {len(S)}; max rel vs oracle {rel_o:.2e}; "
          f"max rel vs 200-bit refined ({m} endpoints) {rel_r:.2e}")


def test_cfg2_full_batch_plain_and_learned(gpu_ctx):
    from hcgpu import learned
    F, p0, Q = cfg2_problem(256)
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    from pathlib import Path
    S = np.load(Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy")
    starts = np.tile(S, (len(Q), 1))
    groups = np.repeat(np.arange(len(Q), dtype=np.int32), len(S))
    prog = gpu_ctx.program(H)
    gp = track_batch(prog, starts, make_options(), group_of_path=groups)
    learned.load(gpu_ctx, learned.saved_weights("param10_step_mlp"))
    gl = track_batch(prog, starts, make_options(learned_step=1), group_of_path=groups)
    assert list(gl.code) == list(gp.code) and list(gl.singular) == list(gp.singular)
    # no path jumps: every learned endpoint is the plain tracker's solution (all 262,144 paths)
    ok = (gp.code == 1) & (gl.code == 1)
    jump = np.abs(gl.solution[ok] - gp.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(gp.solution[ok]).max(axis=1))
    assert jump.max() <= TOL, f"a learned path ended at another solution (rel {jump.max():.2e})"
    # the oracle on 2 of the 256 targets (2,048 paths): every path, every field that decides counts
    m = 2 * len(S)
    o = oracle.track_batch(H, starts[:m], make_options(), group_of_path=groups[:m])
    rel_o = _same_as_oracle(gp.code[:m], gp.singular[:m], gp.solution[:m], o)
    rel_ol = _same_as_oracle(gl.code[:m], gl.singular[:m], gl.solution[:m], o)
    # 256 sampled endpoints of the learned run, refined at their own targets
    fin = np.nonzero((gl.code == 1) & (~gl.singular))[0]
    rng = np.random.default_rng(2)
    pick = rng.choice(fin, size=SAMPLE, replace=False)
    worst = 0.0
    for gi in np.unique(groups[pick]):
        sel = pick[groups[pick] == gi]
        R, last = refine(F.fix_parameters(Q[gi]), gl.solution[sel])
        assert (last < 1e-30).all()
        rel = rel_error(gl.solution[sel], R)
        assert rel.max() <= TOL, rel.max()
        worst = max(worst, float(rel.max()))
    print(f"CONFIG_TABLE cfg2/3: 262144 paths, plain {_counts(gp.code, gp.singular)}, learned identical codes; "
          f"oracle identical codes on {m}; max rel vs oracle plain {rel_o:.2e} learned {rel_ol:.2e}; "
          f"max rel vs 200-bit refined ({SAMPLE} learned endpoints) {worst:.2e}")


def test_cfg4_mixed_learned_vs_oracle(gpu_ctx):
    from hcgpu import learned
    probs = cfg4_problem()
    Hs = [H for H, _ in probs]
    starts = [S for _, S in probs]
    multi = gpu_ctx.program_multi(Hs)
    learned.load(gpu_ctx, learned.saved_weights("mixed_step_mlp"))
    outs = track_multi(multi, starts, make_options(learned_step=1))
    plain = track_multi(multi, starts, make_options())
    rng = np.random.default_rng(4)
    lines = []
    for H, S, g, gp in zip(Hs, starts, outs, plain):
        # every path: the plain tracker's code, and no jump to another solution
        assert list(g.code) == list(gp.code) and list(g.singular) == list(gp.singular)
        ok = (g.code == 1) & (gp.code == 1)
        if ok.any():
            jump = np.abs(g.solution[ok] - gp.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(gp.solution[ok]).max(axis=1))
            assert jump.max() <= TOL, f"n={H.n}: a learned path ended at another solution (rel {jump.max():.2e})"
        m = len(S) if H.n < 12 else 8192
        o = oracle.track_batch(H, S[:m], make_options())
        rel_o = _same_as_oracle(g.code[:m], g.singular[:m], g.solution[:m], o)
        assert list(g.winding_number[:m]) == list(o.winding_number)
        assert _counts(g.code[:m], g.singular[:m]) == _counts(o.code, o.singular)
        F = shared_factor_system(H.n, seed=CFG4_SEEDS["system"] + H.n)
        fin = (g.code == 1) & (~g.singular)
        rel_r, k = _refined_sample(F, g.solution[fin], rng, SAMPLE // 4) if fin.any() else (0.0, 0)
        lines.append(f"n={H.n}: {len(S)} paths {_counts(g.code, g.singular)}, oracle identical on {m}, "
                     f"rel vs oracle {rel_o:.2e}, vs refined ({k}) {rel_r:.2e}")
    print("CONFIG_TABLE cfg4 (learned, one launch): " + "; ".join(lines))
