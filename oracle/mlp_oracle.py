"""FP64 restatement of the surrogate MLP (forward, backward, MSE loss, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ and nowhere in the product.

The surrogate has no counterpart in the reference (it is new in this
library), so this oracle is pinned by mathematics alone: the backward pass
against central finite differences of the forward pass, the forward pass
against brute-force loops on tiny inputs, Adam against torch.optim.Adam, and
the GRU forward and backward against torch.nn.GRU + autograd in FP64 and
central differences (tests/test_oracle_mlp.py).

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


# ---------------------------------------------------------------------------
# GRU surrogate: gates (r, z, n) as torch.nn.GRU, linear readout per step.
#   gi = x W_ih^T + b_ih, gh = h W_hh^T + b_hh
#   r = s(gi_r + gh_r), z = s(gi_z + gh_z), n = tanh(gi_n + r gh_n)
#   h' = (1 - z) n + z h,   y = h' W_o^T + b_o
# params: dict W_ih [3H, I], b_ih, W_hh [3H, H], b_hh, W_o [O, H], b_o.

def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def gru_forward(p, X, h0=None, bf16=False):
    """X [T, B, I] -> (Y [T, B, O], cache)."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    T, B, _ = X.shape
    H = p["W_hh"].shape[1]
    h = np.zeros((B, H)) if h0 is None else np.asarray(h0, np.float64)
    Wih, Whh, Wo = rnd(p["W_ih"]), rnd(p["W_hh"]), rnd(p["W_o"])
    Xr = rnd(X)
    hs, gis, ghs, ys = [h], [], [], []
    for t in range(T):
        gi = Xr[t] @ Wih.T + p["b_ih"]
        gh = rnd(h) @ Whh.T + p["b_hh"]
        r = _sig(gi[:, :H] + gh[:, :H])
        z = _sig(gi[:, H:2 * H] + gh[:, H:2 * H])
        n = np.tanh(gi[:, 2 * H:] + r * gh[:, 2 * H:])
        h = (1 - z) * n + z * h
        gis.append(gi)
        ghs.append(gh)
        hs.append(h)
        ys.append(rnd(h) @ Wo.T + p["b_o"])
    return np.stack(ys), dict(hs=hs, gis=gis, ghs=ghs, Xr=Xr)


def gru_backward(p, cache, gY, bf16=False):
    """Gradients of every parameter for output gradient gY [T, B, O]."""
    rnd = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    hs, gis, ghs, Xr = cache["hs"], cache["gis"], cache["ghs"], cache["Xr"]
    T = len(gis)
    H = p["W_hh"].shape[1]
    Wih, Whh, Wo = rnd(p["W_ih"]), rnd(p["W_hh"]), rnd(p["W_o"])
    g = {k: np.zeros_like(np.asarray(v, np.float64)) for k, v in p.items()}
    dnext = np.zeros_like(hs[0])
    for t in range(T - 1, -1, -1):
        gy = rnd(gY[t])
        g["W_o"] += gy.T @ rnd(hs[t + 1])
        g["b_o"] += gY[t].sum(0)
        d = gy @ Wo + dnext
        gi, gh, h = gis[t], ghs[t], hs[t]
        r = _sig(gi[:, :H] + gh[:, :H])
        z = _sig(gi[:, H:2 * H] + gh[:, H:2 * H])
        ghn = gh[:, 2 * H:]
        n = np.tanh(gi[:, 2 * H:] + r * ghn)
        dn = d * (1 - z) * (1 - n * n)
        dz = d * (h - n) * z * (1 - z)
        dr = dn * ghn * r * (1 - r)
        dgi = np.concatenate([dr, dz, dn], axis=1)
        dgh = np.concatenate([dr, dz, dn * r], axis=1)
        g["W_ih"] += rnd(dgi).T @ Xr[t]
        g["W_hh"] += rnd(dgh).T @ rnd(h)
        g["b_ih"] += dgi.sum(0)
        g["b_hh"] += dgh.sum(0)
        dnext = d * z + rnd(dgh) @ Whh
    return g
