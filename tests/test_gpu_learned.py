"""Learned step-size proposer on the GPU.

* the HIP training kernels' loss and gradients equal the CPU reference (oracle/mlp_ref.py)
  on the same batch, within FP32 accumulation error;
* the plain GPU tracker's step log (the training data) equals the oracle's, record for
  record;
* GPU training lowers the loss;
* with the trained network the tracker reaches the same endpoints as the oracle's plain
  tracker (identical return codes, endpoints within 1e-10) in fewer steps per path.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
import systems
from hcgpu import learned
from hcgpu.configs import cfg2_problem, parameter_draws
from hcgpu.gpu import track_batch
from hcgpu.homotopies import ParameterHomotopy, total_degree
from hcgpu.solve import make_options
from oracle import mlp_ref

pytestmark = pytest.mark.gpu

STARTS = Path(__file__).resolve().parent / "golden" / "cfg2_starts.npy"


def test_mlp_gradients_match_cpu_reference(gpu_ctx):
    rng = np.random.default_rng(0)
    n = 1000
    X = rng.standard_normal((n, 8)).astype(np.float32)
    Y = np.stack([rng.standard_normal(n), (rng.random(n) < 0.3)], axis=1).astype(np.float32)
    W = np.stack([(rng.random(n) < 0.8), np.ones(n)], axis=1).astype(np.float32)
    p = learned.init_params(1)
    pred, loss, grad = learned.forward_backward(gpu_ctx, p, X, Y, W)
    rl, rg, ro = mlp_ref.loss_and_grad(p, X, Y, W)
    np.testing.assert_allclose(pred, ro, rtol=1e-4, atol=1e-5)
    assert loss == pytest.approx(rl, rel=1e-5)
    err = np.abs(grad - rg).max() / np.abs(rg).max()
    assert err < 1e-4, err


def test_step_log_matches_oracle(gpu_ctx):
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    prog = gpu_ctx.program(H)
    g = track_batch(prog, S, make_options(), log_capacity=4096)
    o = oracle.track_batch(H, S, make_options(), log_capacity=4096)

    def srt(L):
        return L[np.lexsort((L[:, 1], L[:, 0]))]
    assert len(g.log) == len(o.log) > 0
    np.testing.assert_array_equal(srt(g.log), srt(o.log))


def _cfg2_batch(Q):
    F, p0, _ = cfg2_problem(1)
    H = ParameterHomotopy(F, p0, Q, lift_affine=True)
    S = np.load(STARTS)
    return H, np.tile(S, (len(Q), 1)), np.repeat(np.arange(len(Q), dtype=np.int32), len(S))


@pytest.fixture(scope="module")
def trained(gpu_ctx):
    H, starts, groups = _cfg2_batch(parameter_draws(16, 4, seed=30))
    out = track_batch(gpu_ctx.program(H), starts, make_options(), group_of_path=groups, log_capacity=600_000)
    X, Y, W = learned.training_set(out.log, out.code)
    mu, sd = learned.feature_scaling(X)
    Xn = ((X - mu) / sd).astype(np.float32)
    p0 = learned.init_params(0)
    _, l0, _ = learned.forward_backward(gpu_ctx, p0, Xn, Y, W)
    p, l1 = learned.train(gpu_ctx, p0, Xn, Y, W, epochs=6)
    return learned.fold_scaling(p, mu, sd), l0, l1


def test_training_lowers_the_loss(trained):
    _, l0, l1 = trained
    print("loss", l0, "->", l1)
    assert l1 < 0.5 * l0


def test_learned_tracking_same_endpoints_fewer_steps(gpu_ctx, trained):
    params = learned.saved_weights("param10_step_mlp")
    if params is None:
        params = trained[0]
    H, starts, groups = _cfg2_batch(cfg2_problem(4)[2])
    prog = gpu_ctx.program(H)
    plain = track_batch(prog, starts, make_options(), group_of_path=groups)
    learned.load(gpu_ctx, params)
    g = track_batch(prog, starts, make_options(learned_step=1), group_of_path=groups)
    o = oracle.track_batch(H, starts, make_options(), group_of_path=groups)
    assert list(plain.code) == list(o.code)
    assert list(g.code) == list(o.code)
    assert list(g.singular) == list(o.singular)
    ok = o.code == 1
    rel = np.abs(g.solution[ok] - o.solution[ok]).max(axis=1) / np.maximum(1.0, np.abs(o.solution[ok]).max(axis=1))
    assert rel.max() <= 1e-10, rel.max()
    sp = (plain.accepted_steps + plain.rejected_steps).mean()
    sl = (g.accepted_steps + g.rejected_steps).mean()
    rate = g.learned_accepted.sum() / max(1, g.learned_proposed.sum())
    print(f"steps/path plain {sp:.2f} learned {sl:.2f}; learned proposals accepted {rate:.3f}")
    assert sl < 0.8 * sp
    assert rate > 0.8
