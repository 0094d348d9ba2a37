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
        # the partial sums this batch wrote (the buffer is sized for the largest batch)
        return self.lsum[: (B * no + 255) // 256]

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


class GRUSurrogate:
    """Single-layer GRU (gate order r, z, n as torch.nn.GRU) with a linear
    readout per step, trained by backpropagation through time on the
    library's kernels.

    Forward over a sequence X [T, B, n_in]: one GEMM for every step's input
    projection, then per step a recurrent GEMM and the fused gate kernel,
    then one GEMM for every step's readout.  Backward: one GEMM each for the
    readout gradients, per step (reverse order) the fused gate backward and
    the recurrent-gradient GEMM, then one GEMM each for dW_ih and dW_hh over
    all T*B rows."""

    def __init__(self, n_in, n_out, hidden=256, max_batch=64, max_steps=32, device=0, seed=0, lr=1e-3,
                 betas=(0.9, 0.999), eps=1e-8):
        nat.require_gpu()
        self.lib = nat.lib()
        self.dev = torch.device("cuda", device)
        self.n_in, self.n_out = int(n_in), int(n_out)
        self.I, self.H, self.O = _pad(n_in), _pad(hidden), _pad(n_out)
        self.max_batch, self.max_steps = _pad(max_batch), int(max_steps)
        self.lr, self.b1, self.b2, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        self.t = 0
        g = torch.Generator().manual_seed(seed)
        f32, bf = dict(dtype=torch.float32, device=self.dev), dict(dtype=torch.bfloat16, device=self.dev)
        H, I, O = self.H, self.I, self.O
        s = 1.0 / np.sqrt(H)
        w_ih = (torch.rand((3 * H, I), generator=g) * 2 - 1) * s
        w_ih[:, self.n_in:] = 0.0
        init = dict(W_ih=w_ih, b_ih=(torch.rand(3 * H, generator=g) * 2 - 1) * s,
                    W_hh=(torch.rand((3 * H, H), generator=g) * 2 - 1) * s,
                    b_hh=(torch.rand(3 * H, generator=g) * 2 - 1) * s,
                    W_o=torch.randn((O, H), generator=g) * s, b_o=torch.zeros(O))
        self.p = {k: v.float().to(self.dev) for k, v in init.items()}
        self.grad = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.pb = {k: torch.empty(self.p[k].shape, **bf) for k in ("W_ih", "W_hh", "W_o")}
        self.pbT = {k: torch.empty(tuple(reversed(self.p[k].shape)), **bf) for k in ("W_hh", "W_o")}
        B, T = self.max_batch, self.max_steps
        self.X = torch.empty((T * B, I), **f32)
        self.Xb = torch.empty((T * B, I), **bf)
        self.XbT = torch.empty((I, T * B), **bf)
        self.Gi = torch.empty((T * B, 3 * H), **f32)
        self.Gh = torch.empty((T * B, 3 * H), **f32)
        self.Hs = torch.empty(((T + 1) * B, H), **f32)
        self.Hb = torch.empty(((T + 1) * B, H), **bf)
        self.HbT = torch.empty((H, (T + 1) * B), **bf)
        self.Y = torch.empty((T * B, O), **f32)
        self.Tg = torch.empty((T * B, O), **f32)
        self.M = torch.empty((T * B, O), **f32)
        self.Gy = torch.empty((T * B, O), **f32)
        self.Gyb = torch.empty((T * B, O), **bf)
        self.GybT = torch.empty((O, T * B), **bf)
        self.dHo = torch.empty((T * B, H), **f32)
        self.dGi = torch.empty((T * B, 3 * H), **f32)
        self.dGib = torch.empty((T * B, 3 * H), **bf)
        self.dGh = torch.empty((T * B, 3 * H), **f32)
        self.dGhb = torch.empty((T * B, 3 * H), **bf)
        self.dGT = torch.empty((3 * H, T * B), **bf)
        self.dhz = torch.empty((B, H), **f32)
        self.dhg = torch.empty((B, H), **f32)
        self.lsum = torch.empty(((T * B * O + 255) // 256,), **f32)
        self.x_scale = torch.ones(self.n_in, **f32)
        self.x_shift = torch.zeros(self.n_in, **f32)
        self.y_scale = torch.ones(self.n_out, **f32)
        self.y_shift = torch.zeros(self.n_out, **f32)
        self._y_std = torch.ones(self.n_out, **f32)
        self._y_mean = torch.zeros(self.n_out, **f32)
        self._refresh_bf16()

    def _stream(self):
        return torch.cuda.current_stream(self.dev).cuda_stream

    def _refresh_bf16(self):
        s = self._stream()
        for k in ("W_ih", "W_hh", "W_o"):
            r, c = self.p[k].shape
            nat.check(self.lib.dtb_mlp_f32_to_bf16(r * c, self.p[k].data_ptr(), self.pb[k].data_ptr(), s))
            if k in self.pbT:
                nat.check(self.lib.dtb_mlp_transpose_bf16(r, c, self.pb[k].data_ptr(), self.pbT[k].data_ptr(), s))

    def set_normalization(self, X, Y):
        """Per-feature standardisation over every step and sequence."""
        X = torch.as_tensor(X)
        Y = torch.as_tensor(Y)
        MLPSurrogate.set_normalization(self, X.reshape(-1, X.shape[-1]), Y.reshape(-1, Y.shape[-1]))

    def _gemm(self, M, N, K, A, lda, Bm, ldb, bias, C, ldc, Cb=None, ldcb=0):
        nat.check(self.lib.dtb_mlp_gemm_bf16_nt(M, N, K, A.data_ptr(), lda, Bm.data_ptr(), ldb, nat.ptr(bias), 0,
                                                nat.ptr(C), ldc, nat.ptr(Cb), ldcb, self._stream()))

    def _forward(self, X, h0):
        """X [T, r, n_in] (r <= max_batch rows, padded to B); h0 [r, H] or None."""
        s = self._stream()
        T, r = X.shape[0], X.shape[1]
        assert T <= self.max_steps and r <= self.max_batch
        B = _pad(r)
        H, I, O = self.H, self.I, self.O
        Xp = torch.zeros((T, B, self.n_in), dtype=torch.float32, device=self.dev)
        Xp[:, :r] = torch.as_tensor(X).to(self.dev, torch.float32)
        nat.check(self.lib.dtb_mlp_affine(T * B, self.n_in, I, Xp.data_ptr(), self.n_in, self.x_scale.data_ptr(),
                                          self.x_shift.data_ptr(), None, self.Xb.data_ptr(), I, s))
        self._gemm(T * B, 3 * H, I, self.Xb, I, self.pb["W_ih"], I, self.p["b_ih"], self.Gi, 3 * H)
        self.Hs[:B].zero_()
        if h0 is not None:
            self.Hs[:r].copy_(torch.as_tensor(h0).to(self.dev, torch.float32))
        nat.check(self.lib.dtb_mlp_f32_to_bf16(B * H, self.Hs.data_ptr(), self.Hb.data_ptr(), s))
        for t in range(T):
            hb = self.Hb[t * B:(t + 1) * B]
            gh = self.Gh[t * B:(t + 1) * B]
            self._gemm(B, 3 * H, H, hb, H, self.pb["W_hh"], H, self.p["b_hh"], gh, 3 * H)
            nat.check(self.lib.dtb_gru_forward(B, H, self.Gi[t * B:].data_ptr(), gh.data_ptr(),
                                               self.Hs[t * B:].data_ptr(), self.Hs[(t + 1) * B:].data_ptr(),
                                               self.Hb[(t + 1) * B:].data_ptr(), s))
        self._gemm(T * B, O, H, self.Hb[B:], H, self.pb["W_o"], H, self.p["b_o"], self.Y, O)
        return T, r, B

    def predict(self, X, h0=None):
        """X [T, rows, n_in] -> (Y [T, rows, n_out] denormalised, h_T [rows, H])."""
        T, r, B = self._forward(X, h0)
        out = torch.empty((T * B, self.n_out), dtype=torch.float32, device=self.dev)
        nat.check(self.lib.dtb_mlp_affine(T * B, self.n_out, self.n_out, self.Y.data_ptr(), self.O,
                                          self._y_std.data_ptr(), self._y_mean.data_ptr(), out.data_ptr(), None,
                                          self.n_out, self._stream()))
        return out.reshape(T, B, -1)[:, :r], self.Hs[T * B:T * B + r].clone()

    def gradients(self, X, Y, mask=None, h0=None):
        """BPTT over one batch of sequences; loss sum(mask (Yn-Tn)^2) / (2 T rows)."""
        s = self._stream()
        T, r, B = self._forward(X, h0)
        H, I, O = self.H, self.I, self.O
        TB = T * B
        Yp = torch.zeros((T, B, self.n_out), dtype=torch.float32, device=self.dev)
        Yp[:, :r] = torch.as_tensor(Y).to(self.dev, torch.float32)
        nat.check(self.lib.dtb_mlp_affine(TB, self.n_out, O, Yp.data_ptr(), self.n_out, self.y_scale.data_ptr(),
                                          self.y_shift.data_ptr(), self.Tg.data_ptr(), None, O, s))
        M = self.M[:TB].view(T, B, O)
        M.zero_()
        if mask is None:
            M[:, :r, :self.n_out] = 1.0
        else:
            M[:, :r, :self.n_out] = torch.as_tensor(mask).to(self.dev, torch.float32)
        nat.check(self.lib.dtb_mlp_mse_grad(TB * O, self.Y.data_ptr(), self.Tg.data_ptr(), self.M.data_ptr(),
                                            1.0 / (T * r), self.Gy.data_ptr(), self.Gyb.data_ptr(),
                                            self.lsum.data_ptr(), s))
        # readout
        tr = self.lib.dtb_mlp_transpose_bf16
        nat.check(tr(TB, O, self.Gyb.data_ptr(), self.GybT.data_ptr(), s))
        nat.check(tr(TB, H, self.Hb[B:].data_ptr(), self.HbT.data_ptr(), s))
        self._gemm(O, H, TB, self.GybT, TB, self.HbT, TB, None, self.grad["W_o"], H)
        nat.check(self.lib.dtb_mlp_bias_grad(TB, O, self.Gy.data_ptr(), self.grad["b_o"].data_ptr(), s))
        self._gemm(TB, H, O, self.Gyb, O, self.pbT["W_o"], O, None, self.dHo, H)
        # recurrence, last step first
        for t in range(T - 1, -1, -1):
            last = t == T - 1
            nat.check(self.lib.dtb_gru_backward(
                B, H, self.Gi[t * B:].data_ptr(), self.Gh[t * B:].data_ptr(), self.Hs[t * B:].data_ptr(),
                self.dHo[t * B:].data_ptr(), None if last else self.dhz.data_ptr(),
                None if last else self.dhg.data_ptr(), self.dGi[t * B:].data_ptr(), self.dGib[t * B:].data_ptr(),
                self.dGh[t * B:].data_ptr(), self.dGhb[t * B:].data_ptr(), self.dhz.data_ptr(), s))
            if t > 0:
                self._gemm(B, H, 3 * H, self.dGhb[t * B:(t + 1) * B], 3 * H, self.pbT["W_hh"], 3 * H, None, self.dhg,
                           H)
        # weight gradients over all T*B rows
        nat.check(tr(TB, 3 * H, self.dGib.data_ptr(), self.dGT.data_ptr(), s))
        nat.check(tr(TB, I, self.Xb.data_ptr(), self.XbT.data_ptr(), s))
        self._gemm(3 * H, I, TB, self.dGT, TB, self.XbT, TB, None, self.grad["W_ih"], I)
        nat.check(tr(TB, 3 * H, self.dGhb.data_ptr(), self.dGT.data_ptr(), s))
        nat.check(tr(TB, H, self.Hb.data_ptr(), self.HbT.data_ptr(), s))
        self._gemm(3 * H, H, TB, self.dGT, TB, self.HbT, TB, None, self.grad["W_hh"], H)
        nat.check(self.lib.dtb_mlp_bias_grad(TB, 3 * H, self.dGi.data_ptr(), self.grad["b_ih"].data_ptr(), s))
        nat.check(self.lib.dtb_mlp_bias_grad(TB, 3 * H, self.dGh.data_ptr(), self.grad["b_hh"].data_ptr(), s))
        # the partial sums this batch wrote (the buffer is sized for the longest, largest batch)
        return self.lsum[: (TB * O + 255) // 256]

    def step(self):
        s = self._stream()
        self.t += 1
        for k, w in self.p.items():
            nat.check(self.lib.dtb_mlp_adam(w.numel(), w.data_ptr(), self.grad[k].data_ptr(), self.m[k].data_ptr(),
                                            self.v[k].data_ptr(), self.lr, self.b1, self.b2, self.eps, self.t,
                                            nat.ptr(self.pb.get(k)), s))
        for k in self.pbT:
            r, c = self.p[k].shape
            nat.check(self.lib.dtb_mlp_transpose_bf16(r, c, self.pb[k].data_ptr(), self.pbT[k].data_ptr(), s))

    def fit(self, X, Y, epochs=50, batch=None, seed=0):
        """X [T, n, n_in], Y [T, n, n_out] sequences; mini-batches over n."""
        X = torch.as_tensor(X).to(self.dev, torch.float32)
        Y = torch.as_tensor(Y).to(self.dev, torch.float32)
        n = X.shape[1]
        batch = min(self.max_batch, n) if batch is None else int(batch)
        g = torch.Generator().manual_seed(seed)
        hist = []
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), dtype=torch.float64, device=self.dev)
            nb = 0
            for r0 in range(0, n - batch + 1, batch):
                idx = perm[r0:r0 + batch]
                ls = self.gradients(X.index_select(1, idx), Y.index_select(1, idx))
                tot += ls.sum().double() / (batch * X.shape[0])
                self.step()
                nb += 1
            hist.append(float(tot) / max(1, nb) / self.n_out)
        return hist

    def params_numpy(self):
        return {k: v.cpu().double().numpy() for k, v in self.p.items()}


def gemm_kernel_name(M, N, K):
    """The kernel dtb_mlp_gemm_bf16_nt dispatches an M x N x K product to:
    LDS-staged 256x256 tiles when M and N are multiples of 256, K of 64 and
    there are at least 64 such tiles; else 128x128 tiles while they give at
    least 256 workgroups, else 64x64 tiles, else 32x32 tiles with K split
    over the four wavefronts."""
    if M % 256 == 0 and N % 256 == 0 and K % 64 == 0 and (M // 256) * (N // 256) >= 64:
        return "gemm_nt_lds_kernel (256x256 LDS-staged tiles, bf16 MFMA, FP32 accumulate)"
    if ((N + 127) // 128) * ((M + 127) // 128) >= 256:
        return "gemm_nt_kernel<2,2> (128x128 tiles, bf16 MFMA, FP32 accumulate)"
    if ((N + 63) // 64) * ((M + 63) // 64) >= 256:
        return "gemm_nt_kernel<1,1> (64x64 tiles, bf16 MFMA, FP32 accumulate)"
    return "gemm_nt_splitk_kernel (32x32 tiles, K split over 4 wavefronts, bf16 MFMA, FP32 accumulate)"


def gemm_throughput(M, N, K, reps=50, warmup=5, device=0):
    """Time dtb_mlp_gemm_bf16_nt (C = A·Bᵀ, bf16 operands, FP32 out) on
    random operands: HIP events on the stream the kernel is launched on.
    The product is first checked against torch in FP32 (bf16 inputs are
    exact in FP32; only the summation order differs).  Returns a dict with
    the average microseconds per launch and TFLOP/s (2·M·N·K per launch)."""
    nat.require_gpu()
    if M % 32 or N % 32 or K % 16 or min(M, N, K) <= 0:
        raise ValueError("M and N must be positive multiples of 32 and K of 16")
    lib = nat.lib()
    dev = torch.device("cuda", device)
    g = torch.Generator(device="cpu").manual_seed(0)
    A = torch.randn(M, K, generator=g).to(torch.bfloat16).to(dev)
    B = torch.randn(N, K, generator=g).to(torch.bfloat16).to(dev)
    C = torch.empty(M, N, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev)

    def launch():
        nat.check(lib.dtb_mlp_gemm_bf16_nt(M, N, K, A.data_ptr(), K, B.data_ptr(), K, None, 0, C.data_ptr(), N, None,
                                           0, s.cuda_stream))

    launch()
    ref = A.float() @ B.float().T
    err = float((C - ref).abs().max() / ref.abs().max())
    if not err < 1e-5:
        raise RuntimeError(f"GEMM {M}x{N}x{K} differs from torch: relative error {err:.3g}")
    for _ in range(warmup):
        launch()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        launch()
    e1.record(s)
    e1.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / reps
    return {"M": M, "N": N, "K": K, "us": us, "tflops": 2.0 * M * N * K / us * 1e-6, "max_rel_err": err}


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="surrogate GEMM throughput")
    ap.add_argument("--gemm", type=int, nargs=3, action="append", metavar=("M", "N", "K"), required=True)
    ap.add_argument("--reps", type=int, default=50)
    for shape in ap.parse_args().gemm:
        print(json.dumps(gemm_throughput(*shape, reps=ap.parse_args().reps)), flush=True)
