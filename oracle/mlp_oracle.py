"""CPU FP64 forward / backward of the surrogate MLP -- TEST INFRASTRUCTURE ONLY.

The checker of the GPU training kernels (massivemds_amd/csrc/mds_mlp.hip):
same network (input [ (x_i - xc_i) / sigma_x, embedding ] padded to 32, two
tanh layers of 256, linear output padded to 16), same loss (mean squared
error of the P outputs against (g_i - gc_i) / sigma_g over N P values), with
per-location centres xc, gc (N x P) and per-dimension scales (P).  With
``bf16=True`` the values the GPU keeps in BF16 (inputs, weights, activations
and the backward signals) are rounded the same way, and the input
normalisation is done in FP32 as on the GPU, so that the comparison isolates
the kernels' arithmetic from BF16 rounding.  Its own gradient is pinned by
central finite differences in tests/test_mlp_oracle.py.  The network has no
counterpart in the reference.
"""
from __future__ import annotations

import numpy as np

HIDDEN, EMBED, IN, OUT = 256, 16, 32, 16


def to_bf16(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def _inputs(x, xc, sx, p, bf16):
    n = x.shape[0]
    if bf16:  # FP32 as on the GPU: (x - xc) * (1 / sigma_x)
        inv = (np.float32(1.0) / np.asarray(sx[:p], dtype=np.float32)).astype(np.float32)
        d = (np.asarray(x, dtype=np.float32) - np.asarray(xc, dtype=np.float32)).astype(np.float32)
        return (d * inv).astype(np.float32).astype(np.float64)
    return (np.asarray(x) - np.asarray(xc)) / np.asarray(sx[:p])


def forward_backward(params, x, g, xc, gc, sx, sg, p, bf16=False):
    """Returns (loss, gradients in the flat parameter order)."""
    q = to_bf16 if bf16 else _id
    W1, b1, W2, b2, W3, b3, E = params
    n = x.shape[0]
    a0 = np.zeros((n, IN))
    a0[:, :p] = _inputs(x, xc, sx, p, bf16)
    a0[:, p:p + EMBED] = E
    a0 = q(a0)
    z1 = a0 @ q(W1) + b1
    h1 = q(np.tanh(z1))
    z2 = h1 @ q(W2) + b2
    h2 = q(np.tanh(z2))
    out = h2 @ q(W3) + b3
    t = np.zeros((n, OUT))
    t[:, :p] = (np.asarray(g) - np.asarray(gc)) / np.asarray(sg[:p])
    err = np.zeros((n, OUT))
    err[:, :p] = out[:, :p] - t[:, :p]
    loss = float(np.sum(err ** 2) / (n * p))
    dout = q(2.0 * err / (n * p))
    dW3 = h2.T @ dout
    db3 = dout.sum(axis=0)
    dh2 = q((dout @ q(W3).T) * (1.0 - h2 ** 2))
    dW2 = h1.T @ dh2
    db2 = dh2.sum(axis=0)
    dh1 = q((dh2 @ q(W2).T) * (1.0 - h1 ** 2))
    dW1 = a0.T @ dh1
    db1 = dh1.sum(axis=0)
    da0 = dh1 @ q(W1).T
    dE = da0[:, p:p + EMBED]
    return loss, [dW1, db1, dW2, db2, dW3, db3, dE]


def predict(params, x, xc, gc, sx, sg, p, bf16=False):
    """Predicted gradient rows gc_i + sigma_g * network output."""
    q = to_bf16 if bf16 else _id
    W1, b1, W2, b2, W3, b3, E = params
    n = x.shape[0]
    a0 = np.zeros((n, IN))
    a0[:, :p] = _inputs(x, xc, sx, p, bf16)
    a0[:, p:p + EMBED] = E
    h1 = q(np.tanh(q(a0) @ q(W1) + b1))
    h2 = q(np.tanh(h1 @ q(W2) + b2))
    out = h2 @ q(W3) + b3
    return np.asarray(gc) + out[:, :p] * np.asarray(sg[:p])
