"""Learned step-size proposer: training data from the plain tracker's step log, training on
the GPU with libhcgpu's HIP kernels, and the weights the tracker loads.

Network: 8 features -> 64 -> 64 (GELU) -> 2, FP32 master weights, BF16 in the forward
pass, FP32 accumulation.  Layout of the flat parameter vector (HCG_MLP_NPARAMS floats):
W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2].

Step record (16 floats, written by the tracker after every step):
  0 path, 1 step index, 2-9 features (log Δs, θ, log(e Δs^4), t, log10 cond, log ω,
  log μ, log τ), 10 accepted, 11 Newton iterations, 12 log next proposed Δs,
  13 flags (1 extended precision, 2 singular endgame, 4 learned proposal),
  14 log10 accuracy, 15 log of the plain rule's next proposal.

Targets:
  output 0 -- log of the distance the plain tracker covered with its next two accepted
              steps, relative to its own next proposal: log((Δs_{k1} + Δs_{k2}) / Δs_plain_k).
              Only accepted steps outside extended precision and the singular endgame,
              followed by two accepted steps, carry weight.
  output 1 -- at-infinity logit: 1 if the path ended at infinity (or at zero).
"""
from __future__ import annotations

import ctypes
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

from . import abi

IN, HID, OUT = 8, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + OUT * HID
NPARAMS = OFF_B3 + OUT
WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, dtype=np.float32)

    def glorot(fan_out, fan_in):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        return rng.uniform(-lim, lim, size=(fan_out, fan_in)).astype(np.float32)

    p[OFF_W1:OFF_B1] = glorot(HID, IN).ravel()
    p[OFF_W2:OFF_B2] = glorot(HID, HID).ravel()
    p[OFF_W3:OFF_B3] = glorot(OUT, HID).ravel()
    return p


def training_set(log: np.ndarray, path_codes: np.ndarray):
    """(X, Y, W) from a step log (records x 16) and the per-path return codes."""
    log = np.asarray(log, dtype=np.float32)
    order = np.lexsort((log[:, 1], log[:, 0]))
    R = log[order]
    n = len(R)
    X = np.ascontiguousarray(R[:, 2:10], dtype=np.float32)
    Y = np.zeros((n, OUT), dtype=np.float32)
    W = np.zeros((n, OUT), dtype=np.float32)
    path = R[:, 0].astype(np.int64)
    acc = R[:, 10] > 0.5
    clean = (R[:, 13].astype(np.int64) & 3) == 0
    # index of the next accepted record of the same path, for every record
    nxt = np.full(n, -1, dtype=np.int64)
    last = -1
    for i in range(n - 1, -1, -1):
        if i + 1 < n and path[i + 1] != path[i]:
            last = -1
        nxt[i] = last
        if acc[i]:
            last = i
    k1 = nxt
    k2 = np.where(k1 >= 0, nxt[np.maximum(k1, 0)], -1)
    ok = acc & clean & (k1 >= 0) & (k2 >= 0)
    ok[ok] &= clean[k1[ok]] & clean[k2[ok]]
    D = np.exp(R[np.maximum(k1, 0), 2].astype(np.float64)) + np.exp(R[np.maximum(k2, 0), 2].astype(np.float64))
    y0 = np.log(D) - R[:, 15].astype(np.float64)
    ok &= np.isfinite(y0)
    Y[ok, 0] = y0[ok]
    W[ok, 0] = 1.0
    codes = np.asarray(path_codes)
    inf = np.isin(codes, (2, 3))
    Y[:, 1] = inf[np.clip(path, 0, len(codes) - 1)].astype(np.float32)
    W[:, 1] = 1.0
    return X, Y, W


def feature_scaling(X: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    mu = X.mean(axis=0).astype(np.float64)
    sd = X.std(axis=0).astype(np.float64)
    sd[sd < 1e-6] = 1.0
    return mu, sd


def fold_scaling(params: np.ndarray, mu: np.ndarray, sd: np.ndarray) -> np.ndarray:
    """Network trained on (x - mu) / sd  ->  the same network on raw features."""
    p = np.array(params, dtype=np.float64)
    W1 = p[OFF_W1:OFF_B1].reshape(HID, IN)
    b1 = p[OFF_B1:OFF_W2]
    W1n = W1 / sd[None, :]
    b1n = b1 - W1n @ mu
    p[OFF_W1:OFF_B1] = W1n.ravel()
    p[OFF_B1:OFF_W2] = b1n
    return p.astype(np.float32)


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def forward_backward(ctx, params, X, Y, W):
    """One forward + backward on the GPU: (predictions, mean loss, gradient)."""
    L = abi.lib()
    L.hcg_mlp_forward_backward.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                           ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    n = len(X)
    pred = np.zeros((n, OUT), dtype=np.float32)
    grad = np.zeros(NPARAMS, dtype=np.float32)
    loss = ctypes.c_float()
    abi.check(L.hcg_mlp_forward_backward(ctx.handle, _fp(p), n, _fp(X), _fp(Y), _fp(W), _fp(pred),
                                         ctypes.byref(loss), _fp(grad)))
    return pred, loss.value, grad


def train(ctx, params, X, Y, W, epochs: int = 8, batch: int = 4096, lr: float = 2e-3, seed: int = 0):
    """Adam on the GPU over (X, Y, W); returns (params, final mean loss)."""
    L = abi.lib()
    L.hcg_mlp_train.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int32,
                                ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                ctypes.c_uint64, ctypes.POINTER(ctypes.c_float)]
    p = np.array(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    W = np.ascontiguousarray(W, dtype=np.float32)
    fl = ctypes.c_float()
    abi.check(L.hcg_mlp_train(ctx.handle, _fp(p), len(X), _fp(X), _fp(Y), _fp(W), epochs, batch, lr, seed,
                              ctypes.byref(fl)))
    return p, fl.value


def load(ctx, params) -> None:
    """Make `params` the network the tracker uses when options.learned_step = 1."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.shape != (NPARAMS,):
        raise ValueError(f"expected {NPARAMS} parameters")
    L = abi.lib()
    L.hcg_mlp_load.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]
    abi.check(L.hcg_mlp_load(ctx.handle, _fp(p)))


def saved_weights(name: str) -> Optional[np.ndarray]:
    f = WEIGHTS_DIR / f"{name}.npy"
    return np.load(f) if f.exists() else None


def fit_from_log(ctx, log, path_codes, epochs: int = 8, batch: int = 4096, lr: float = 2e-3, seed: int = 0):
    """Full pipeline: targets, feature scaling, GPU training, scaling folded into layer 1."""
    X, Y, W = training_set(log, path_codes)
    mu, sd = feature_scaling(X)
    Xn = ((X - mu) / sd).astype(np.float32)
    p0 = init_params(seed)
    p, loss = train(ctx, p0, Xn, Y, W, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return fold_scaling(p, mu, sd), loss, dict(records=len(X), weighted=int(W[:, 0].sum()),
                                                 mean_target=float(Y[W[:, 0] > 0, 0].mean()) if W[:, 0].any() else 0.0)


def combine_logs(parts):
    """Concatenate (step log, path codes) pairs of several launches into one training log:
    path indices of later parts are offset so that every path stays distinct."""
    logs, codes = [], []
    off = 0
    for log, c in parts:
        L = np.array(log, dtype=np.float32, copy=True)
        L[:, 0] += off
        logs.append(L)
        codes.append(np.asarray(c))
        off += len(c)
    return np.concatenate(logs), np.concatenate(codes)


OFFSET_CANDIDATES = (0.0, -0.1, -0.2, -0.3, -0.4)


def with_output_offset(params, kappa: float) -> np.ndarray:
    """The network with κ subtracted from its log-step output (κ < 0: longer proposals)."""
    p = np.array(params, dtype=np.float32, copy=True)
    p[OFF_B3] -= np.float32(kappa)
    return p


def select_output_offset(ctx, params, runs, kappas=OFFSET_CANDIDATES, endpoint_rtol: float = 1e-8,
                         slack: float = 0.05, repeats: int = 3):
    """Model selection on the training systems: `runs` is a list of callables, each tracking
    one training batch with the options it is given and returning (return codes, endpoints,
    kernel ms).  Every κ is tried on every run.  A candidate is kept only if, on every training
    path, its return code equals the plain tracker's and, for paths that end in success, its
    endpoint equals the plain tracker's within `endpoint_rtol` (relative to max(1, |x|∞)):
    a longer step must not let a path jump to another solution.  Of the kept candidates within
    `slack` of the fastest (median over `repeats` passes of the total kernel time) the one
    closest to κ = 0 wins: rare jumps that a training set of this size cannot show grow with the
    step length.  Returns (weights with that offset, κ, table)."""
    from .solve import make_options

    def timed(opts):
        """Results of every run and the median over `repeats` passes of their total kernel time."""
        for run in runs:
            run(opts)  # warm
        passes = [[run(opts) for run in runs] for _ in range(repeats)]
        return passes[0], float(np.median([sum(r[2] for r in p) for p in passes]))

    plain, plain_ms = timed(make_options())
    table = []
    for k in kappas:
        load(ctx, with_output_offset(params, k))
        got, ms = timed(make_options(learned_step=1))
        same = all(np.array_equal(g[0], pl[0]) for g, pl in zip(got, plain))
        rel = 0.0
        for g, pl in zip(got, plain):
            ok = (pl[0] == 1) & (g[0] == 1)
            if ok.any():
                a, b = np.asarray(g[1])[ok], np.asarray(pl[1])[ok]
                r = np.abs(a - b).max(axis=1) / np.maximum(1.0, np.abs(b).max(axis=1))
                rel = max(rel, float(np.nan_to_num(r, nan=np.inf).max()))
        keep = same and rel <= endpoint_rtol
        table.append({"kappa": k, "learned_ms": round(ms, 3), "plain_ms": round(plain_ms, 3),
                      "ratio": round(plain_ms / ms, 4), "identical_codes": bool(same),
                      "max_rel_endpoint_diff": rel, "kept": bool(keep)})
    kept = [r for r in table if r["kept"]]
    if not kept:
        raise RuntimeError("no output offset keeps the plain tracker's return codes and endpoints on the "
                           "training systems")
    fastest = min(r["learned_ms"] for r in kept)
    best = min((r for r in kept if r["learned_ms"] <= fastest * (1 + slack)), key=lambda r: abs(r["kappa"]))["kappa"]
    return with_output_offset(params, best), best, table
