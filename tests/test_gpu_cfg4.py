"""Mixed sizes in one launch: total-degree homotopies of shared-factor systems
(f_i = l(x) q_i(x) + r_i(x), the cubic parts sharing the linear factor l) with
n = 4, 6, 8, 12 tracked by ONE persistent kernel launch (hcg_track_multi).

* every system's results in the multi launch equal its single-system launch (same
  kernel body per system, so bit for bit);
* n = 4, 6, 8 in full and the first 1,024 paths of n = 12 agree with the oracle:
  identical return codes, singular flags and winding numbers, endpoints within 1e-10;
* full size (538,812 paths): every path ends as success or at infinity, finite
  nonsingular endpoints have rounding-level residuals and are pairwise distinct.
"""
import numpy as np
import pytest

import oracle
from hcgpu.configs import cfg4_problem
from hcgpu.gpu import track_batch, track_multi
from hcgpu.solve import make_options

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def probs():
    return cfg4_problem()


def _assert_same(g, o, label):
    assert list(g.code) == list(o.code), label
    assert list(g.singular) == list(o.singular), label
    assert list(g.winding_number) == list(o.winding_number), label
    ok = (o.code == 1) & (~o.singular)
    if ok.any():
        rel = np.abs(g.solution[ok] - o.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(o.solution[ok]).max(axis=1))
        assert rel.max() <= 1e-10, (label, rel.max())


def test_multi_launch_equals_single_launches(gpu_ctx, probs):
    Hs = [H for H, _ in probs]
    starts = [S if H.n < 12 else S[:2048] for H, S in probs]
    multi = gpu_ctx.program_multi(Hs)
    outs = track_multi(multi, starts, make_options())
    print("multi launch ms", gpu_ctx.last_kernel_ms(), gpu_ctx.last_counters())
    for H, st, g in zip(Hs, starts, outs):
        s = track_batch(gpu_ctx.program(H), st, make_options())
        assert list(g.code) == list(s.code)
        np.testing.assert_array_equal(g.solution, s.solution)
        assert (g.accepted_steps == s.accepted_steps).all()


def _residuals(F, X):
    num = np.zeros((len(X), len(F.polys)), dtype=complex)
    den = np.zeros((len(X), len(F.polys)))
    for i, f in enumerate(F.polys):
        for ex, c in f.terms.items():
            m = np.prod(X ** np.asarray(ex), axis=1)
            num[:, i] += c * m
            den[:, i] += abs(c) * np.abs(m)
    return (np.abs(num) / den).max(axis=1)


def test_full_size_one_launch(gpu_ctx, probs):
    Hs = [H for H, _ in probs]
    starts = [S for _, S in probs]
    multi = gpu_ctx.program_multi(Hs)
    outs = track_multi(multi, starts, make_options())
    ms = gpu_ctx.last_kernel_ms()
    total = sum(len(S) for S in starts)
    print(f"{total} paths in {ms:.1f} ms: {total / ms * 1e3:.0f} paths/s", gpu_ctx.last_counters())
    for H, S, g in zip(Hs, starts, outs):
        codes = dict(zip(*np.unique(g.codes(), return_counts=True)))
        print(f"n={H.n}: {codes}, singular {int(g.singular.sum())}")
        m = len(S) if H.n < 12 else 1024
        o = oracle.track_batch(H, S[:m], make_options())
        sub = type(g)(g.records[:m], g.solution[:m], g.last_point[:m], g.valuation[:m])
        _assert_same(sub, o, f"n={H.n}")
        assert set(np.unique(g.code)) <= {1, 2, 3}, codes
        fin = (g.code == 1) & (~g.singular)
        X = g.solution[fin]
        if len(X):
            assert _residuals(H.F.system, X).max() < 1e-11
            key = np.round(X[:, :3], 6)
            assert len({tuple(r) for r in key}) == len(X), f"n={H.n}: repeated finite endpoints"


def test_learned_mixed_same_results(gpu_ctx, probs):
    """The mixed workload with the learned proposer (network trained on shared-factor
    systems disjoint from these): every path ends with the plain tracker's return code, finite
    endpoints agree within 1e-10, and the batch needs fewer steps."""
    from hcgpu import learned
    w = learned.saved_weights("mixed_step_mlp")
    assert w is not None, "hcgpu/weights/mixed_step_mlp.npy is missing (train_mixed.py writes it)"
    Hs = [H for H, _ in probs]
    starts = [S if H.n < 12 else S[:32768] for H, S in probs]
    multi = gpu_ctx.program_multi(Hs)
    plain = track_multi(multi, starts, make_options())
    learned.load(gpu_ctx, w)
    outs = track_multi(multi, starts, make_options(learned_step=1))
    sp = sl = 0
    for H, g, p in zip(Hs, outs, plain):
        assert list(g.code) == list(p.code), f"n={H.n}"
        ok = (p.code == 1) & (~p.singular)
        if ok.any():
            rel = np.abs(g.solution[ok] - p.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(p.solution[ok]).max(axis=1))
            assert rel.max() <= 1e-10, (H.n, rel.max())
        sp += int((p.accepted_steps + p.rejected_steps).sum())
        sl += int((g.accepted_steps + g.rejected_steps).sum())
    print(f"steps plain {sp} learned {sl}")
    assert sl < 0.85 * sp
