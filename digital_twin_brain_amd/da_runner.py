"""Batched data assimilation over an uneven ensemble (the C4 workload).

One block holds E members of uneven size plus a hidden "truth" member whose
conductance factors are drawn once and kept fixed.  Every window advances the
whole block (all members and the truth in the same step batches), then the
truth's BOLD at the end of the window is the observation of a diffusion EnKF
analysis over the E members (data_assimilation/DataAssimilation.py:402-453).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import networks
from .data_assimilation import DataAssimilation

HP_COLUMNS = (10, 12)  # g_AMPA and g_GABAa of every sub-population


def run_c4(E=64, voxels=90, windows=20, window_ms=800.0, dt=0.1, hp_lo=0.5, hp_hi=2.0, seed=7, log=None,
           **member_kw):
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    nb, props, mem = networks.build_c4(E=E, voxels=voxels, **member_kw)
    t_build = time.perf_counter() - t0
    P = voxels * 2
    sps = int(round(1.0 / dt))  # BOLD once per ms
    da = DataAssimilation(network=nb, column=False, ensemble=E, dt=dt, bold_steps=sps, da_populations=E * P,
                          da_voxels=E * voxels, write_path="/tmp")
    H = len(HP_COLUMNS)
    low = torch.full((P, H), hp_lo)
    high = torch.full((P, H), hp_hi)
    hp = da.hp_random_initialize(low, high, rng=rng)
    da.hp_index2hidden_state()
    da.da_property_initialize(list(HP_COLUMNS), None, hp)
    # hidden truth: fixed factors on the truth member's sub-populations
    truth_pops = np.arange(E * P, (E + 1) * P)
    truth_hp = rng.uniform(hp_lo, hp_hi, size=(P, H)).astype(np.float32)
    nb.mul_props_by_pop(np.repeat(truth_pops, H).astype(np.uint32), np.tile(HP_COLUMNS, P).astype(np.uint32),
                        truth_hp.reshape(-1))
    steps = int(round(window_ms / dt))
    truth_vox = slice(E * voxels, (E + 1) * voxels)
    obs = []

    def observe(t):
        y = torch.from_numpy(nb.bold_state("bold")[truth_vox])
        obs.append(y.numpy())
        return y

    # a first window so the BOLD state leaves its initial transient
    da.evolve(steps, bold_detail=True)
    w_save, w_fix, times = da.da_rest_run(observe, "/tmp", observation_times=windows, steps=steps, save=False)
    bold_da = w_save[:, :, :, -1]  # [T, E, V] forecast BOLD
    obs = np.stack(obs)
    err = bold_da.mean(1) - obs
    out = dict(windows=windows, steps_per_window=steps, window_s_mean=float(times.mean()),
               window_s_median=float(np.median(times)), build_s=t_build, neurons=int(nb.n), synapses=int(nb.nnz),
               members=E, voxels=voxels, ensemble_bold_rmse_first=float(np.sqrt((err[0] ** 2).mean())),
               ensemble_bold_rmse_last=float(np.sqrt((err[-1] ** 2).mean())))
    if log:
        log(f"C4 DA: {out}")
    return out, da, dict(w_save=w_save, w_fix=w_fix, obs=obs, truth_hp=truth_hp)
