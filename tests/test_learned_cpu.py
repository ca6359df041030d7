"""CPU checks of the learned step-size proposer's host pieces and of its CPU reference.

* the numpy reference gradient (oracle/mlp_ref.py) against central finite differences;
* BF16 rounding: round to nearest, ties to even;
* folding the feature scaling into layer 1 leaves the network's function unchanged;
* training targets built from a hand-made step log;
* the oracle's step log: feature columns are what they claim to be.
"""
import numpy as np
import pytest

import oracle
import systems
from hcgpu import learned
from hcgpu.homotopies import total_degree
from hcgpu.solve import make_options
from oracle import mlp_ref


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 8))
    Y = np.stack([rng.standard_normal(n), (rng.random(n) < 0.3).astype(float)], axis=1)
    W = np.stack([(rng.random(n) < 0.8).astype(float), np.ones(n)], axis=1)
    return X, Y, W


def test_reference_gradient_matches_finite_differences():
    X, Y, W = _data(37, 1)
    p = learned.init_params(3).astype(np.float64) + np.random.default_rng(4).normal(0, 0.05, learned.NPARAMS)
    p = p.astype(np.float32).astype(np.float64)
    _, g, _ = mlp_ref.loss_and_grad(p, X, Y, W, bf16=False)
    rng = np.random.default_rng(5)
    idx = np.concatenate([rng.choice(np.arange(mlp_ref.OFF_W1, mlp_ref.OFF_B1), 8),
                          rng.choice(np.arange(mlp_ref.OFF_W2, mlp_ref.OFF_B2), 8),
                          np.arange(mlp_ref.OFF_W3, mlp_ref.NPARAMS)[::9], [mlp_ref.OFF_B1 + 3, mlp_ref.OFF_B2 + 5]])

    def f(q):
        # float64 evaluation: bypass the FP32 cast of the weights
        W1, b1, W2, b2, W3, b3 = mlp_ref.unpack(q)
        a1 = X @ W1.T + b1
        a2 = mlp_ref.gelu(a1) @ W2.T + b2
        o = mlp_ref.gelu(a2) @ W3.T + b3
        e = o[:, 0] - Y[:, 0]
        z = o[:, 1]
        bce = np.maximum(z, 0) - z * Y[:, 1] + np.log1p(np.exp(-np.abs(z)))
        return (W[:, 0] * 0.5 * e * e + W[:, 1] * bce).sum() / len(X)

    for i in idx:
        h = 1e-6
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd = (f(q1) - f(q2)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_round_to_nearest_even():
    r = mlp_ref.bf16_round(np.array([1.0, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, -2.5, 1 + 2.0 ** -7 + 2.0 ** -8],
                                    dtype=np.float32))
    np.testing.assert_array_equal(r, np.array([1.0, 1.0, 1 + 2.0 ** -6, -2.5, 1 + 2.0 ** -6], dtype=np.float32))


def test_fold_scaling_preserves_the_function():
    X, _, _ = _data(50, 2)
    X = X * np.array([3, 0.1, 10, 1, 2, 5, 0.5, 4]) + np.arange(8)
    mu, sd = learned.feature_scaling(X.astype(np.float32))
    p = learned.init_params(7)
    o1, _ = mlp_ref.forward(p, (X - mu) / sd, bf16=False)
    o2, _ = mlp_ref.forward(learned.fold_scaling(p, mu, sd), X, bf16=False)
    np.testing.assert_allclose(o1, o2, rtol=1e-4, atol=1e-4)


def test_output_offset_shifts_only_the_log_step_output():
    """with_output_offset(p, κ) lowers the log-step output by exactly κ and leaves the
    at-infinity output and every other parameter unchanged."""
    X, _, _ = _data(40, 3)
    p = learned.init_params(3)
    q = learned.with_output_offset(p, -0.25)
    o1, _ = mlp_ref.forward(p, X, bf16=False)
    o2, _ = mlp_ref.forward(q, X, bf16=False)
    np.testing.assert_allclose(o2[:, 0] - o1[:, 0], 0.25, rtol=0, atol=1e-6)
    np.testing.assert_array_equal(o2[:, 1], o1[:, 1])
    diff = np.flatnonzero(p != q)
    assert list(diff) == [learned.OFF_B3]


@pytest.mark.parametrize("name", ["param10_step_mlp", "mixed_step_mlp"])
def test_committed_offset_follows_the_selection_rule(name):
    """The offset recorded beside each committed network is the one select_output_offset's rule
    picks from the recorded table: kept (training codes and endpoints equal the plain
    tracker's), within 5% of the fastest kept candidate, and closest to 0 among those."""
    import json
    from pathlib import Path
    meta = json.loads((Path(learned.__file__).parent / "weights" / f"{name}.json").read_text())
    kept = [r for r in meta["table"] if r["kept"]]
    assert all(r["identical_codes"] and r["max_rel_endpoint_diff"] <= 1e-8 for r in kept)
    fastest = min(r["learned_ms"] for r in kept)
    pick = min((r for r in kept if r["learned_ms"] <= fastest * 1.05), key=lambda r: abs(r["kappa"]))
    assert pick["kappa"] == meta["kappa"]
    assert learned.saved_weights(name).shape == (learned.NPARAMS,)


def test_training_targets_from_a_step_log():
    """Path 0: steps of size 0.1 (acc), 0.2 (rejected), 0.15 (acc), 0.3 (acc), 0.3 (acc).
    Target of step 0 = log((0.15 + 0.3) / plain_0); step 2 = log((0.3 + 0.3) / plain_2)."""
    rows = []
    sizes = [0.1, 0.2, 0.15, 0.3, 0.3]
    accs = [1, 0, 1, 1, 1]
    plain = [0.2, 0.15, 0.3, 0.3, 0.3]
    for k, (s, a, pl) in enumerate(zip(sizes, accs, plain)):
        r = np.zeros(16, dtype=np.float32)
        r[0], r[1], r[2], r[10], r[15] = 0, k, np.log(s), a, np.log(pl)
        rows.append(r)
    r = np.zeros(16, dtype=np.float32)  # path 1: one record, at infinity
    r[0], r[1], r[2], r[10], r[15] = 1, 0, np.log(0.5), 1, np.log(0.5)
    rows.append(r)
    log = np.stack(rows)[::-1]  # order must not matter
    X, Y, W = learned.training_set(log, np.array([1, 2]))
    # rows come back sorted by (path, step)
    np.testing.assert_allclose(np.exp(X[:5, 0]), sizes, rtol=1e-6)
    assert list(W[:, 0]) == [1, 0, 1, 0, 0, 0]
    assert Y[0, 0] == pytest.approx(np.log(0.45 / 0.2), rel=1e-5)
    assert Y[2, 0] == pytest.approx(np.log(0.6 / 0.3), rel=1e-5)
    assert list(Y[:, 1]) == [0, 0, 0, 0, 0, 1]


def test_oracle_step_log_features():
    H, S = total_degree(systems.beyond_polyhedral(), gamma=1.3j + 0.4)
    o = oracle.track_batch(H, S, make_options(), log_capacity=10000)
    log = o.log
    assert len(log) == int((o.accepted_steps + o.rejected_steps).sum())
    for p in range(len(S)):
        R = log[log[:, 0] == p]
        R = R[np.argsort(R[:, 1])]
        assert list(R[:, 1]) == list(range(len(R)))
        acc = R[:, 10] > 0.5
        assert acc.sum() == o.accepted_steps[p]
        # t after each step decreases along the path (tracking 1 -> 0), column 5 is t
        assert (np.diff(R[acc, 5]) <= 0).all()
        # the next step's size is the proposal logged by the previous step (col 12)
        np.testing.assert_allclose(R[1:, 2], R[:-1, 12], rtol=1e-6, atol=1e-6)
