"""CPU reference of the learned step-size proposer's network (test infrastructure only).

The MLP has no counterpart in the reference; this numpy restatement is what the GPU
training kernels (homotopycontinuation.jl_amd/csrc/mlp.hip) are checked against:
8 -> 64 -> 64 (GELU, erf form) -> 2, weights rounded to BF16 (round to nearest even)
in the forward pass, straight-through gradients to the FP32 master weights, loss per
record  w0 (o0 - y0)^2 / 2 + w1 BCEWithLogits(o1, y1), mean over the batch.
Computed in float64 so that it is the exact value the FP32 kernels approximate.
"""
import math

import numpy as np

IN, HID, OUT = 8, 64, 2
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + OUT * HID
NPARAMS = OFF_B3 + OUT

_erf = np.vectorize(math.erf)


def bf16_round(a):
    """FP32 -> BF16 (round to nearest, ties to even) -> FP32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    r = np.where(special, u.astype(np.uint32), r)
    return r.view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(HID, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3].reshape(OUT, HID), p[OFF_B3:NPARAMS])


def gelu(a):
    return 0.5 * a * (1.0 + _erf(a / math.sqrt(2.0)))


def gelu_d(a):
    return 0.5 * (1.0 + _erf(a / math.sqrt(2.0))) + a * np.exp(-0.5 * a * a) / math.sqrt(2.0 * math.pi)


def forward(params, X, bf16=True):
    p = bf16_round(params) if bf16 else np.asarray(params, dtype=np.float32)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = np.asarray(X, dtype=np.float64)
    a1 = X @ W1.T + b1
    a2 = gelu(a1) @ W2.T + b2
    o = gelu(a2) @ W3.T + b3
    return o, (a1, a2)


def loss_and_grad(params, X, Y, W, bf16=True):
    """(mean loss, gradient w.r.t. the master weights, outputs)."""
    o, (a1, a2) = forward(params, X, bf16)
    p = bf16_round(params) if bf16 else np.asarray(params, dtype=np.float32)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Y = np.asarray(Y, dtype=np.float64)
    Wt = np.asarray(W, dtype=np.float64)
    n = len(X)
    e = o[:, 0] - Y[:, 0]
    z = o[:, 1]
    bce = np.maximum(z, 0) - z * Y[:, 1] + np.log1p(np.exp(-np.abs(z)))
    loss = float((Wt[:, 0] * 0.5 * e * e + Wt[:, 1] * bce).sum() / n)
    dO = np.stack([Wt[:, 0] * e, Wt[:, 1] * (1.0 / (1.0 + np.exp(-z)) - Y[:, 1])], axis=1) / n
    h2 = gelu(a2)
    h1 = gelu(a1)
    gW3 = dO.T @ h2
    gb3 = dO.sum(axis=0)
    d2 = (dO @ W3) * gelu_d(a2)
    gW2 = d2.T @ h1
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ W2) * gelu_d(a1)
    gW1 = d1.T @ np.asarray(X, dtype=np.float64)
    gb1 = d1.sum(axis=0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, g, o
