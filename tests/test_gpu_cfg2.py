"""Batched parameter homotopy on the GPU: n = 10 sparse quadrics whose coefficients are
affine in 16 parameters, 1,024 start solutions at p0 (tests/golden/cfg2_starts.npy),
tracked to many target parameter vectors in one launch.

* parity with the oracle on 4 targets (4,096 paths): identical codes, endpoints within
  1e-10 relative (normally bit for bit);
* full size, 256 targets (262,144 paths): size-independent properties -- every path
  ends at a nonsingular solution of F(x; q_g), the residual is at rounding level, and
  each target's 1,024 endpoints are pairwise distinct (Bézout's bound: a complete set).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from hcgpu.configs import cfg2_problem
from hcgpu.homotopies import ParameterHomotopy
from hcgpu.solve import EndgameTracker, make_options

pytestmark = pytest.mark.gpu

STARTS = Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy"


def _batch(ntargets):
    F, p0, Q = cfg2_problem(ntargets)
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    S = np.load(STARTS)
    starts = np.tile(S, (len(Q), 1))
    groups = np.repeat(np.arange(len(Q), dtype=np.int32), len(S))
    return F, Q, H, starts, groups


def test_cfg2_parity_with_oracle(gpu_ctx):
    F, Q, H, starts, groups = _batch(4)
    T = EndgameTracker(H, ctx=gpu_ctx)
    g = T.track_batch(starts, group_of_path=groups)
    o = oracle.track_batch(H, starts, make_options(), group_of_path=groups)
    print("kernel ms", gpu_ctx.last_kernel_ms(), "counters", gpu_ctx.last_counters())
    assert list(g.code) == list(o.code)
    assert list(g.singular) == list(o.singular)
    assert (g.code == 1).all() and not g.singular.any()
    rel = np.abs(g.solution - o.solution).max(axis=1) / np.maximum(1.0, np.abs(o.solution).max(axis=1))
    assert rel.max() <= 1e-10, rel.max()
    assert (g.accepted_steps == o.accepted_steps).all()


def _residuals(F, Q, sol, groups):
    """max_i |f_i(x; q)| / Σ_α |c_iα| |x^α| for every path, from the expanded 16-parameter system."""
    out = np.empty(len(sol))
    for gi in range(len(Q)):
        Fq = F.fix_parameters(Q[gi])
        idx = np.nonzero(groups == gi)[0]
        X = sol[idx]
        num = np.zeros((len(idx), len(Fq.polys)), dtype=complex)
        den = np.zeros((len(idx), len(Fq.polys)))
        for i, f in enumerate(Fq.polys):
            for ex, c in f.terms.items():
                m = np.prod(X ** np.asarray(ex), axis=1)
                num[:, i] += c * m
                den[:, i] += abs(c) * np.abs(m)
        out[idx] = (np.abs(num) / den).max(axis=1)
    return out


def test_cfg2_full_size_properties(gpu_ctx):
    F, Q, H, starts, groups = _batch(256)
    T = EndgameTracker(H, ctx=gpu_ctx)
    g = T.track_batch(starts, group_of_path=groups)
    ms = gpu_ctx.last_kernel_ms()
    print(f"262,144 paths in {ms:.1f} ms: {len(starts) / ms * 1e3:.0f} paths/s", gpu_ctx.last_counters())
    assert (g.code == 1).all(), {c: int((g.code == c).sum()) for c in np.unique(g.code)}
    assert not g.singular.any()
    assert (g.accuracy < 1e-10).all()
    res = _residuals(F, Q, g.solution, groups)
    assert res.max() < 1e-12, res.max()
    for gi in range(len(Q)):
        X = g.solution[groups == gi]
        key = np.round(X[:, :3], 6)
        assert len({tuple(r) for r in key}) == len(X), f"target {gi}: repeated endpoints"
