"""Batched data assimilation over an uneven ensemble.

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


def _setup(E, voxels, dt, hp_lo, hp_hi, rng, member_kw, tune=True):
    """Block of E members + truth, DataAssimilation over the members, random
    initial hyper-parameters pushed, truth factors drawn and pushed, delivery
    launch shape autotuned on the live network."""
    nb, props, mem = networks.build_ensemble(E=E, voxels=voxels, **member_kw)
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
    if tune:
        from .autotune import tune_delivery

        nb.run(10 * sps, noise_mode=1, graph=False)  # leave the initial transient before measuring
        tune_delivery(nb, dict(noise_mode=1), steps=2 * sps)
    return nb, da, truth_hp, (low, high)


def run_ensemble_da(E=64, voxels=90, windows=20, window_ms=800.0, dt=0.1, hp_lo=0.5, hp_hi=2.0, seed=7, log=None, tune=True,
           **member_kw):
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    nb, da, truth_hp, _ = _setup(E, voxels, dt, hp_lo, hp_hi, rng, member_kw, tune)
    t_build = time.perf_counter() - t0
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
    spk = np.zeros(100, np.uint32)
    ev = np.zeros(100, np.uint64)
    nb.run(100, noise_mode=1, graph=True, spikes=spk, events=ev)
    out = dict(windows=windows, steps_per_window=steps, window_s_mean=float(times.mean()),
               events_per_step=float(ev.mean()), spikes_per_step=float(spk.mean()), launch=nb.get_launch(),
               window_s_median=float(np.median(times)), build_s=t_build, neurons=int(nb.n), synapses=int(nb.nnz),
               members=E, voxels=voxels, ensemble_bold_rmse_first=float(np.sqrt((err[0] ** 2).mean())),
               ensemble_bold_rmse_last=float(np.sqrt((err[-1] ** 2).mean())))
    if log:
        log(f"ensemble DA: {out}")
    return out, da, dict(w_save=w_save, w_fix=w_fix, obs=obs, truth_hp=truth_hp)


def run_surrogate_da(E=64, voxels=90, train_windows=40, windows=20, window_ms=800.0, dt=0.1, hp_lo=0.5, hp_hi=2.0, seed=7,
           every=5, nrmse_max=0.02, hidden=(512, 512, 512), epochs=None, finetune_epochs=None, per_voxel=True,
           log=None, tune=True, adapt=0.0,
           **member_kw):
    """Surrogate-assisted DA against simulator-only DA on the same ensemble.

    1. `train_windows` simulator windows with fresh random hyper-parameters
       give E training samples each; the window surrogate is fitted on them.
    2. Simulator-only DA runs `windows` windows against the truth member and
       records the observations.
    3. From the same initial hyper-parameters, surrogate-assisted DA runs the
       same number of windows against the recorded observations, simulating
       every `every`-th window and after any window whose NRMSE > nrmse_max.
    Returns timings (gain = simulator-only seconds / surrogate-assisted
    seconds), the fallback pattern, NRMSE per simulated window and the
    ensemble-mean BOLD error of both runs."""
    from .surrogate_da import SurrogateDA, WindowSurrogate, collect_windows

    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    nb, da, truth_hp, (low, high) = _setup(E, voxels, dt, hp_lo, hp_hi, rng, member_kw, tune)
    t_build = time.perf_counter() - t0
    steps = int(round(window_ms / dt))
    truth_vox = slice(E * voxels, (E + 1) * voxels)
    say = log or (lambda *a: None)
    da.evolve(steps, bold_detail=True)
    say(f"surrogate DA: block of {nb.n} neurons built and warmed in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    # training windows: random walks of the logits from the assimilation's own starting spread
    # (twice its per-window step), restarted every 8 windows
    Xs, Ys = [], []
    for w0 in range(0, train_windows, 8):
        da.hp_random_initialize(low, high, rng=np.random.default_rng(seed + 100 + w0))
        Xw, Yw = collect_windows(da, min(8, train_windows - w0), steps, rng=rng, walk_sigma=2 * 0.25 ** 0.5)
        Xs.append(Xw)
        Ys.append(Yw)
    X, Y = torch.cat(Xs), torch.cat(Ys)
    t_collect = time.perf_counter() - t0
    say(f"surrogate DA: {train_windows} training windows ({X.shape[0]} samples) in {t_collect:.1f} s")
    ws = WindowSurrogate(voxels, hidden=hidden, seed=seed, per_voxel=per_voxel)
    ws.add(X, Y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = ws.fit(epochs=epochs)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    say(f"surrogate DA: fitted in {t_fit:.1f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")

    def reset_hp():
        hp = da.hp_random_initialize(low, high, rng=np.random.default_rng(seed + 1))
        da._push(hp)
        da.timings = []

    from .region_da import _restore, _snapshot

    # both runs start from this state (block state, properties, hyper-parameters, RNG streams)
    reset_hp()
    da._hidden_state = None
    snap = _snapshot(da)
    obs = []

    def observe(t):
        y = nb.bold_state("bold")[truth_vox].copy()
        obs.append(y)
        return torch.from_numpy(y)

    w_sim, _, t_sim = da.da_rest_run(observe, "/tmp", observation_times=windows, steps=steps, save=False)
    obs = np.stack(obs)
    say(f"surrogate DA: {windows} simulator-only windows, {t_sim.mean():.3f} s each")
    _restore(da, snap)
    del snap
    sda = SurrogateDA(da, ws, every=every, nrmse_max=nrmse_max, finetune_epochs=finetune_epochs, adapt=adapt)
    res = sda.run(lambda t: torch.from_numpy(obs[t]), windows, steps=steps)
    say(f"surrogate DA: {windows} surrogate-assisted windows, {res['times'].mean():.3f} s each, "
        f"{int(res['simulated'].sum())} simulated")
    rmse = lambda w: np.sqrt(((w[:, :, :, -1].mean(1) - obs) ** 2).mean(1))
    out = dict(members=E, voxels=voxels, windows=windows, steps_per_window=steps, build_s=t_build,
               neurons=int(nb.n), synapses=int(nb.nnz),
               surrogate=("per-voxel " if per_voxel else "per-member ") + "MLP " + "x".join(str(h) for h in hidden),
               train_windows=train_windows, train_samples=int(X.shape[0]), collect_s=t_collect, fit_s=t_fit,
               fit_loss_first=hist[0], fit_loss_last=hist[-1], sim_window_s_mean=float(t_sim.mean()),
               sur_window_s_mean=float(res["times"].mean()),
               gain=float(t_sim.sum() / res["times"].sum()), simulated_windows=int(res["simulated"].sum()),
               nrmse=[round(e, 5) for _, e in res["nrmse"]],
               bold_rmse_sim_last=float(rmse(w_sim)[-1]), bold_rmse_sur_last=float(rmse(res["forecasts"])[-1]),
               bold_rmse_sim_mean=float(rmse(w_sim).mean()), bold_rmse_sur_mean=float(rmse(res["forecasts"]).mean()))
    if log:
        log(f"surrogate DA: {out}")
    return out, dict(w_sim=w_sim, sur=res, obs=obs, hist=hist, truth_hp=truth_hp, da=da)
