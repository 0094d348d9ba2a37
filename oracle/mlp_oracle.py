"""FP64 restatement of the surrogate MLP (forward, backward, MSE loss, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere in the product.

The surrogate has no counterpart in the reference (it is new in this
library), so this oracle is pinned by mathematics alone: the backward pass
against central finite differences of the forward pass, the forward pass
against brute-force loops on tiny inputs, and Adam against torch.optim.Adam
(tests/test_oracle_mlp.py).

Layout follows the library: weights W_l are [out_l, in_l], a layer is
Z = H W^T + b, hidden layers apply ReLU, the last layer is the identity; the
loss is scale * sum(mask * (Y - T)^2) / 2 so that dL/dY = scale * mask * (Y - T)
(mask is 0/1).

With `bf16=True` the operands are rounded to bfloat16 at the points where the
GPU path stores them in bf16 (input, weights, hidden activations, output
gradient and hidden gradients fed to the matrix products); everything else is
exact FP64, so the oracle differs from the GPU only by FP32 accumulation order.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    out[nan] = np.nan
    return out


def forward(params, X, bf16=False):
    """params: list of (W [out,in], b [out]).  Returns (Y, hs) where hs[l] is
    the input of layer l (hs[0] = X)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    h = rnd(X)
    hs = [h]
    L = len(params)
    for l, (W, b) in enumerate(params):
        z = h @ rnd(W).T + np.asarray(b, np.float64)
        if l < L - 1:
            h = np.maximum(z, 0.0)
            hs.append(rnd(h))
            h = hs[-1]
        else:
            return z, hs
    raise ValueError("no layers")


def loss_and_grad(Y, T, mask=None, scale=1.0):
    d = (Y - T) * (1.0 if mask is None else mask)
    return 0.5 * scale * float((d * d).sum()), scale * d


def backward(params, hs, gY, bf16=False):
    """Gradients (dW_l, db_l) of a loss whose output gradient is gY."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    grads = [None] * len(params)
    g = gY
    for l in range(len(params) - 1, -1, -1):
        W, _ = params[l]
        gb = rnd(g)
        grads[l] = (gb.T @ hs[l], g.sum(axis=0))
        if l > 0:
            dh = gb @ rnd(W)
            g = np.where(hs[l] > 0, dh, 0.0)
    return grads


def adam(w, g, m, v, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba, 2015, Algorithm 1); returns new (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
