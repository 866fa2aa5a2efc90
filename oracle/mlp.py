"""Oracle of the learned-CF coordinate network (test infrastructure only).

FP64 numpy restatement of the network the library trains and evaluates on the GPU: dense
layers [5, H x layers, 2] with SiLU (z * sigmoid(z)) after every hidden layer, loss
sum((y - t)^2) / (2n), and its gradient by the chain rule.  The parameter vector has, per layer,
the weights (inputs x outputs, row-major) and then the biases.

The network has no counterpart in the reference (its CFs are FFT tables), so there is no
reference routine to follow.  Pins: the analytic gradient against central finite differences
of the loss in FP64 (tests/test_oracle_pins.py), and SiLU' against its closed form.
"""
from __future__ import annotations

import numpy as np


def dims(hidden: int, layers: int):
    return [5] + [hidden] * layers + [2]


def unpack(p: np.ndarray, hidden: int, layers: int):
    d = dims(hidden, layers)
    out, o = [], 0
    for i in range(len(d) - 1):
        W = p[o:o + d[i] * d[i + 1]].reshape(d[i], d[i + 1])
        o += d[i] * d[i + 1]
        b = p[o:o + d[i + 1]]
        o += d[i + 1]
        out.append((W, b))
    assert o == p.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, x, hidden, layers):
    layers_ = unpack(np.asarray(p, np.float64), hidden, layers)
    a = np.asarray(x, np.float64)
    for W, b in layers_[:-1]:
        a = silu(a @ W + b)
    W, b = layers_[-1]
    return a @ W + b


def forward_f16(p, x, hidden, layers, rnd=None):
    """The rounding of the fused FP16 inference: weights and inputs rounded to FP16, every hidden
    activation rounded to FP16 after its SiLU, biases as stored (FP32); products and sums here in
    FP64 (the GPU sums in FP32).  rnd: the rounding (default: to numpy float16 and back)."""
    if rnd is None:
        rnd = lambda a: np.asarray(a).astype(np.float16).astype(np.float64)  # noqa: E731
    layers_ = unpack(np.asarray(p, np.float64), hidden, layers)
    a = rnd(np.asarray(x, np.float64))
    for W, b in layers_[:-1]:
        a = rnd(silu(a @ rnd(W) + np.float32(0) + b))
    W, b = layers_[-1]
    return a @ rnd(W) + b


def loss_grad(p, x, t, hidden, layers):
    """(loss, gradient) in FP64."""
    p = np.asarray(p, np.float64)
    ls = unpack(p, hidden, layers)
    x = np.asarray(x, np.float64)
    t = np.asarray(t, np.float64)
    n = x.shape[0]
    acts, zs = [x], []
    a = x
    for W, b in ls[:-1]:
        z = a @ W + b
        zs.append(z)
        a = silu(z)
        acts.append(a)
    W, b = ls[-1]
    y = a @ W + b
    loss = float(np.sum((y - t) ** 2) / (2.0 * n))
    dz = (y - t) / n
    grads = []
    for li in range(len(ls) - 1, -1, -1):
        W, _ = ls[li]
        grads.append((acts[li].T @ dz, dz.sum(axis=0)))
        if li == 0:
            break
        dz = (dz @ W.T) * silu_grad(zs[li - 1])
    grads.reverse()
    g = np.concatenate([np.concatenate([gw.ravel(), gb]) for gw, gb in grads])
    return loss, g
