"""Learned surrogate of a data-assimilation window (no counterpart in the
reference: new in this library).

`MLPSurrogate` is a ReLU multilayer perceptron trained and evaluated with the
library's own kernels (bf16 matrix-core GEMM with FP32 accumulation, fused
bias + ReLU, ReLU backward, bias gradient, masked MSE gradient, Adam on FP32
master weights, normalise/pad).  PyTorch provides only device memory and the
stream.

Layer l computes Z = H W_l^T + b_l with W_l stored [out, in]; widths are
padded to multiples of 32 and the batch to a multiple of 32 (padded rows
carry zero loss weight).  Backward per layer, with G the bf16 gradient of Z:
dW = G^T H (GEMM of the two transposes), db = column sums of G (FP32),
dH = G W (GEMM against the transposed weights), masked by H > 0.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


def _pad(n, m=32):
    return (int(n) + m - 1) // m * m


class MLPSurrogate:
    def __init__(self, n_in, n_out, hidden=(512, 512, 512), max_batch=256, device=0, seed=0, lr=1e-3,
                 betas=(0.9, 0.999), eps=1e-8):
        nat.require_gpu()
        self.lib = nat.lib()
        self.dev = torch.device("cuda", device)
        self.n_in, self.n_out = int(n_in), int(n_out)
        self.sizes = [_pad(n_in)] + [_pad(h) for h in hidden] + [_pad(n_out)]
        self.max_batch = _pad(max_batch)
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.t = 0
        g = torch.Generator().manual_seed(seed)
        f32, bf = dict(dtype=torch.float32, device=self.dev), dict(dtype=torch.bfloat16, device=self.dev)
        self.W, self.b, self.Wb, self.WbT = [], [], [], []
        self.mW, self.vW, self.mb, self.vb, self.dW, self.db = [], [], [], [], [], []
        for l, (i, o) in enumerate(zip(self.sizes[:-1], self.sizes[1:])):
            w = torch.randn((o, i), generator=g) * np.sqrt(2.0 / i)  # He initialisation
            if l == 0:
                w[:, self.n_in:] = 0.0
            self.W.append(w.to(self.dev))
            self.b.append(torch.zeros(o, **f32))
            self.Wb.append(torch.empty((o, i), **bf))
            self.WbT.append(torch.empty((i, o), **bf))
            for lst in (self.mW, self.vW, self.dW):
                lst.append(torch.zeros((o, i), **f32))
            for lst in (self.mb, self.vb, self.db):
                lst.append(torch.zeros(o, **f32))
        B = self.max_batch
        self.H = [torch.empty((B, s), **f32) for s in self.sizes]       # layer inputs (H[0] = X) and output
        self.Hb = [torch.empty((B, s), **bf) for s in self.sizes[:-1]]
        self.HbT = [torch.empty((s, B), **bf) for s in self.sizes[:-1]]
        self.G = [torch.empty((B, s), **f32) for s in self.sizes[1:]]    # gradient of each layer's Z
        self.Gb = [torch.empty((B, s), **bf) for s in self.sizes[1:]]
        self.GbT = [torch.empty((s, B), **bf) for s in self.sizes[1:]]
        self.dH = torch.empty((B, max(self.sizes)), **f32)
        self.T = torch.empty((B, self.sizes[-1]), **f32)
        self.M = torch.empty((B, self.sizes[-1]), **f32)
        self.lsum = torch.empty(((B * self.sizes[-1] + 255) // 256,), **f32)
        self.x_scale = torch.ones(self.n_in, **f32)
        self.x_shift = torch.zeros(self.n_in, **f32)
        self.y_scale = torch.ones(self.n_out, **f32)
        self.y_shift = torch.zeros(self.n_out, **f32)
        self._y_std = torch.ones(self.n_out, **f32)
        self._y_mean = torch.zeros(self.n_out, **f32)
        self._refresh_bf16()

    # -- helpers -----------------------------------------------------------
    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _refresh_bf16(self):
        s = self._stream()
        for l, w in enumerate(self.W):
            o, i = w.shape
            nat.check(self.lib.dtb_mlp_f32_to_bf16(o * i, w.data_ptr(), self.Wb[l].data_ptr(), s))
            nat.check(self.lib.dtb_mlp_transpose_bf16(o, i, self.Wb[l].data_ptr(), self.WbT[l].data_ptr(), s))

    def set_normalization(self, X, Y):
        """Per-feature standardisation from training data (host or device arrays)."""
        X = torch.as_tensor(X, dtype=torch.float64)
        Y = torch.as_tensor(Y, dtype=torch.float64)
        xs = X.std(0).clamp_min(1e-12)
        ys = Y.std(0).clamp_min(1e-12)
        self.x_scale = (1.0 / xs).float().to(self.dev)
        self.x_shift = (-X.mean(0) / xs).float().to(self.dev)
        self.y_scale = (1.0 / ys).float().to(self.dev)
        self.y_shift = (-Y.mean(0) / ys).float().to(self.dev)
        self._y_std = ys.float().to(self.dev)
        self._y_mean = Y.mean(0).float().to(self.dev)

    def _load_x(self, X, rows):
        X = X.to(self.dev, torch.float32).contiguous()
        nat.check(self.lib.dtb_mlp_affine(rows, self.n_in, self.sizes[0], X.data_ptr(), X.stride(0),
                                          self.x_scale.data_ptr(), self.x_shift.data_ptr(), self.H[0].data_ptr(),
                                          self.Hb[0].data_ptr(), self.sizes[0], self._stream()))

    def _forward(self, B):
        s = self._stream()
        L = len(self.W)
        for l in range(L):
            o, i = self.W[l].shape
            last = l == L - 1
            nat.check(self.lib.dtb_mlp_gemm_bf16_nt(B, o, i, self.Hb[l].data_ptr(), i, self.Wb[l].data_ptr(), i,
                                                    self.b[l].data_ptr(), 0 if last else 1, self.H[l + 1].data_ptr(),
                                                    o, None if last else self.Hb[l + 1].data_ptr(), o, s))

    # -- public API --------------------------------------------------------
    def predict(self, X):
        """X [rows, n_in] -> [rows, n_out] FP32 on the device (denormalised)."""
        X = torch.as_tensor(X)
        n = X.shape[0]
        out = torch.empty((n, self.n_out), dtype=torch.float32, device=self.dev)
        for r0 in range(0, n, self.max_batch):
            r = min(self.max_batch, n - r0)
            B = _pad(r)
            self._load_x(X[r0:r0 + r], r)
            if B > r:
                self.Hb[0][r:B].zero_()
            self._forward(B)
            nat.check(self.lib.dtb_mlp_affine(r, self.n_out, self.n_out, self.H[-1].data_ptr(), self.sizes[-1],
                                              self._y_std.data_ptr(), self._y_mean.data_ptr(),
                                              out[r0:r0 + r].data_ptr(), None, self.n_out, self._stream()))
        return out

    def gradients(self, X, Y, mask=None):
        """Forward + backward on one batch (rows <= max_batch); leaves dW, db
        and returns the device loss partial sums.  The loss is
        sum(mask*(Yn - Tn)^2) / (2*rows) on standardised outputs."""
        s = self._stream()
        X = torch.as_tensor(X)
        r = X.shape[0]
        assert r <= self.max_batch
        B = _pad(r)
        no = self.sizes[-1]
        self._load_x(X, r)
        Yd = torch.as_tensor(Y).to(self.dev, torch.float32).contiguous()
        nat.check(self.lib.dtb_mlp_affine(r, self.n_out, no, Yd.data_ptr(), Yd.stride(0), self.y_scale.data_ptr(),
                                          self.y_shift.data_ptr(), self.T.data_ptr(), None, no, s))
        self.M[:B].zero_()
        if mask is None:
            self.M[:r, :self.n_out].fill_(1.0)
        else:
            self.M[:r, :self.n_out].copy_(torch.as_tensor(mask).to(self.dev, torch.float32))
        if B > r:
            self.Hb[0][r:B].zero_()
            self.T[r:B].zero_()
        self._forward(B)
        L = len(self.W)
        nat.check(self.lib.dtb_mlp_mse_grad(B * no, self.H[L].data_ptr(), self.T.data_ptr(), self.M.data_ptr(),
                                            1.0 / r, self.G[L - 1].data_ptr(), self.Gb[L - 1].data_ptr(),
                                            self.lsum.data_ptr(), s))
        for l in range(L - 1, -1, -1):
            o, i = self.W[l].shape
            nat.check(self.lib.dtb_mlp_transpose_bf16(B, o, self.Gb[l].data_ptr(), self.GbT[l].data_ptr(), s))
            nat.check(self.lib.dtb_mlp_transpose_bf16(B, i, self.Hb[l].data_ptr(), self.HbT[l].data_ptr(), s))
            # dW[o, i] = sum_b G[b, o] H[b, i]
            nat.check(self.lib.dtb_mlp_gemm_bf16_nt(o, i, B, self.GbT[l].data_ptr(), B, self.HbT[l].data_ptr(), B,
                                                    None, 0, self.dW[l].data_ptr(), i, None, 0, s))
            nat.check(self.lib.dtb_mlp_bias_grad(B, o, self.G[l].data_ptr(), self.db[l].data_ptr(), s))
            if l > 0:
                # dH[b, i] = sum_o G[b, o] W[o, i], then the ReLU mask of layer l-1
                nat.check(self.lib.dtb_mlp_gemm_bf16_nt(B, i, o, self.Gb[l].data_ptr(), o, self.WbT[l].data_ptr(), o,
                                                        None, 0, self.dH.data_ptr(), i, None, 0, s))
                nat.check(self.lib.dtb_mlp_relu_backward(B * i, self.dH.data_ptr(), self.H[l].data_ptr(),
                                                         self.G[l - 1].data_ptr(), self.Gb[l - 1].data_ptr(), s))
        return self.lsum

    def step(self):
        """One Adam step on every parameter from the last `gradients` call."""
        s = self._stream()
        self.t += 1
        for l in range(len(self.W)):
            o, i = self.W[l].shape
            nat.check(self.lib.dtb_mlp_adam(o * i, self.W[l].data_ptr(), self.dW[l].data_ptr(), self.mW[l].data_ptr(),
                                            self.vW[l].data_ptr(), self.lr, self.b1, self.b2, self.eps, self.t,
                                            self.Wb[l].data_ptr(), s))
            nat.check(self.lib.dtb_mlp_adam(o, self.b[l].data_ptr(), self.db[l].data_ptr(), self.mb[l].data_ptr(),
                                            self.vb[l].data_ptr(), self.lr, self.b1, self.b2, self.eps, self.t, None,
                                            s))
            nat.check(self.lib.dtb_mlp_transpose_bf16(o, i, self.Wb[l].data_ptr(), self.WbT[l].data_ptr(), s))

    def fit(self, X, Y, epochs=50, batch=None, seed=0, verbose=False):
        """Mini-batch Adam over (X, Y) kept on the device; returns the mean
        standardised loss per epoch."""
        X = torch.as_tensor(X).to(self.dev, torch.float32)
        Y = torch.as_tensor(Y).to(self.dev, torch.float32)
        n = X.shape[0]
        batch = min(self.max_batch, n) if batch is None else int(batch)
        g = torch.Generator().manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), dtype=torch.float64, device=self.dev)
            for r0 in range(0, n - batch + 1, batch):
                idx = perm[r0:r0 + batch]
                ls = self.gradients(X.index_select(0, idx), Y.index_select(0, idx))
                tot += ls.sum().double() / batch
                self.step()
            hist.append(float(tot) / max(1, n // batch) / self.n_out)
            if verbose:
                print(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist

    def params_numpy(self):
        """[(W, b)] in FP32, unpadded widths dropped only for clarity of tests."""
        return [(w.cpu().double().numpy(), b.cpu().double().numpy()) for w, b in zip(self.W, self.b)]
