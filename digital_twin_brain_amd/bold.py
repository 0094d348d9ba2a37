"""Drop-in replacement of `models.bold_model_pytorch.BOLD` (models/bold_model_pytorch.py:6-78).

State tensors `s, q, v, f_in` live on the GPU and are advanced in place by
the HIP kernel dtb_bold_step_{f32,f64}; the dtype follows the input `u`
(float64 inputs run the float64 kernel).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


class BOLD:
    def __init__(self, epsilon, tao_s, tao_f, tao_0, alpha, E_0, V_0, delta_t=1e-3, init_f_in=None, init_s=None,
                 init_v=None, init_q=None):
        self.epsilon = epsilon
        self.tao_s = tao_s
        self.tao_f = tao_f
        self.tao_0 = tao_0
        self.E_0 = E_0
        self.V_0 = V_0
        self.delta_t = delta_t
        self.div_alpha = 1 / alpha
        self.alpha = alpha
        self.f_in = init_f_in
        self.s = init_s
        self.v = init_v
        self.q = init_q
        self._par = np.array([epsilon, tao_s, tao_f, tao_0, alpha, E_0, V_0, delta_t], dtype=np.float64)

    def state_update(self, w):
        """models/bold_model_pytorch.py:34-41: w[..., 0:4] = (s, q, v, f_in), q/v/f_in floored at 1e-5."""
        if isinstance(w, np.ndarray):
            w = torch.from_numpy(w.astype(np.float64)).cuda()
        self.s = w[:, :, 0].reshape(-1)
        self.q = torch.clamp_min(w[:, :, 1], 1e-05).reshape(-1)
        self.v = torch.clamp_min(w[:, :, 2], 1e-05).reshape(-1)
        self.f_in = torch.clamp_min(w[:, :, 3], 1e-05).reshape(-1)

    def _prep(self, x, like):
        x = x.to(device=like.device, dtype=like.dtype)
        return x.contiguous()

    def run(self, u):
        """One Euler step of 1 delta_t; returns the BOLD signal, shaped like u."""
        nat.require_gpu()
        if isinstance(u, np.ndarray):
            u = torch.from_numpy(u.astype(np.float64))
        if not u.is_cuda:
            u = u.cuda()
        if u.dtype not in (torch.float32, torch.float64):
            u = u.to(torch.float32)
        u = u.contiguous()
        if self.s is None:
            self.s = torch.zeros_like(u)
        if self.q is None:
            self.q = torch.zeros_like(u)
        if self.v is None:
            self.v = torch.ones_like(u)
        if self.f_in is None:
            self.f_in = torch.ones_like(u)
        self.s, self.q, self.v, self.f_in = (self._prep(x, u) for x in (self.s, self.q, self.v, self.f_in))
        out = torch.empty_like(u)
        stream = torch.cuda.current_stream(u.device).cuda_stream
        fn = nat.lib().dtb_bold_step_f64 if u.dtype == torch.float64 else nat.lib().dtb_bold_step_f32
        with torch.cuda.device(u.device):
            nat.check(fn(u.numel(), self._par.ctypes.data, self.s.data_ptr(), self.q.data_ptr(), self.v.data_ptr(),
                         self.f_in.data_ptr(), u.data_ptr(), out.data_ptr(), stream))
        return out
