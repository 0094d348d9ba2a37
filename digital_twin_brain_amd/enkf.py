"""Drop-in `diffusion_enkf` (data_assimilation/DataAssimilation.py:40-60) on the GPU.

The analysis runs in FP64 in dtb_enkf_analysis_f64.  The observation noise
xi ~ N(0, 1) [E, V] is drawn with torch on the GPU unless given.
"""
from __future__ import annotations

import torch

from . import _native as nat


def diffusion_enkf(w_hat, bold_sigma, bold_t, solo_rate, debug=False, xi=None, generator=None):
    nat.require_gpu()
    ensembles, brain_num, state_num = w_hat.shape
    dev = w_hat.device if w_hat.is_cuda else torch.device("cuda", torch.cuda.current_device())
    w = w_hat.detach().to(device=dev, dtype=torch.float64).contiguous().clone()
    y = torch.as_tensor(bold_t).to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
    if y.numel() != brain_num:
        raise ValueError(f"bold_t has {y.numel()} values for {brain_num} voxels")
    if xi is None:
        xi = torch.normal(0.0, 1.0, size=(ensembles, brain_num), device=dev, dtype=torch.float64,
                          generator=generator)
    xi = torch.as_tensor(xi).to(device=dev, dtype=torch.float64).contiguous()
    w_in = w.clone() if debug else None
    stream = torch.cuda.current_stream(dev).cuda_stream
    with torch.cuda.device(dev):
        nat.check(nat.lib().dtb_enkf_analysis_f64(int(ensembles), int(brain_num), int(state_num), w.data_ptr(),
                                                   float(bold_sigma), y.data_ptr(), float(solo_rate), xi.data_ptr(),
                                                   stream))
    out = w.to(dtype=w_hat.dtype, device=w_hat.device)
    if debug:
        # data_assimilation/DataAssimilation.py:55-57 (diagnostic only)
        wd = w_in - w_in.mean(dim=0, keepdim=True)
        cxx = (wd[:, :, -1] ** 2).sum(0) / (ensembles - 1) + bold_sigma
        temp = wd[:, :, -1] / cxx.reshape(1, -1) / (ensembles - 1)
        yn = y.reshape(1, -1) + bold_sigma ** 0.5 * xi
        w_debug = w_in[:, :10, -1][None, :, :] + (yn - w_in[:, :, -1]).T[:, :, None] * \
            torch.mm(temp.T, wd[:, :10, -1])[:, None, :]
        return out, w_debug.to(dtype=w_hat.dtype, device=w_hat.device)
    return out
