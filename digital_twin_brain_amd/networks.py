"""Synthetic benchmark workloads (seeded).

plain block     N=10,000 (80% E), g_ui ~ U[0.001, 0.05] per neuron and
                channel, in-degree 100, weights U[0,1], delay 1, noise 10 Hz,
                8,000 steps.
large block     N=5e6, in-degree 1,000 (5e9 synapses, compact int32 index +
                fp32 weight), delays U{1..10}, per-neuron noise 5-20 Hz,
                firing 2-20 Hz, 10,000 steps, BOLD per 1,000-neuron voxel once
                per ms.
ensemble        E members of log-uniform size and uniform in-degree packed in
                one block, plus a hidden "truth" member.
Property columns follow generation/make_block.py:394-434.
"""
from __future__ import annotations

import numpy as np

from .engine import NativeBlock

# conductances of the calibrated small block, test/test_generation.py:313-314
G_CAL = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32)


def base_props(n, dt=0.1):
    p = np.zeros((n, 22), np.float32)
    p[:, 4] = 1.0    # C
    p[:, 5] = 2.0    # T_ref (ms)
    p[:, 6] = 0.03   # g_L
    p[:, 7] = -75.0  # V_L
    p[:, 8] = -50.0  # V_th
    p[:, 9] = -65.0  # V_reset
    p[:, 14:18] = (0.0, 0.0, -70.0, -100.0)
    p[:, 18:22] = (2.0, 40.0, 10.0, 50.0)
    return p


def plain_props(n=10_000, seed=2, dt=0.1):
    rng = np.random.default_rng(seed)
    p = base_props(n, dt)
    n_exc = int(round(0.8 * n))
    p[:, 0] = 10.0 * dt * 1e-3          # 10 Hz -> 1e-3 per step
    p[:, 3] = (np.arange(n) >= n_exc)   # sub-population 0 = E, 1 = I
    p[:, 10:14] = rng.uniform(0.001, 0.05, size=(n, 4)).astype(np.float32)
    return p, n_exc


def large_props(n=5_000_000, seed=2, dt=0.1, voxel=1000):
    """Large-block parameters: inhibition-dominated coupling (0.1 x / 0.05 x the
    calibrated E / I conductances, +-20% per neuron) with I_ext = 5 gives an
    asynchronous state at ~8 Hz under in-degree 1,000."""
    rng = np.random.default_rng(seed)
    p = base_props(n, dt)
    n_exc = int(round(0.8 * n))
    p[:, 0] = rng.uniform(5.0, 20.0, size=n).astype(np.float32) * np.float32(dt * 1e-3)
    p[:, 2] = 5.0
    p[:, 3] = np.arange(n) // voxel
    scale = np.array([0.1, 0.1, 0.05, 0.05], np.float32)
    p[:, 10:14] = G_CAL * scale * rng.uniform(0.8, 1.2, size=(n, 4)).astype(np.float32)
    return p, n_exc


def build_plain(n=10_000, device=0, conn_seed=1, noise_seed=3):
    p, ne = plain_props(n)
    return NativeBlock.fixed_indegree(p, ne, 100, max_delay=1, w_lo=0.0, w_hi=1.0, pair_weights=False,
                                      conn_seed=conn_seed, noise_seed=noise_seed, device=device), p


def build_large(n=5_000_000, in_degree=1000, device=0, conn_seed=1, noise_seed=3):
    p, ne = large_props(n)
    return NativeBlock.fixed_indegree(p, ne, in_degree, max_delay=10, w_lo=0.0, w_hi=1.0, pair_weights=False,
                                      conn_seed=conn_seed, noise_seed=noise_seed, device=device), p


def ensemble_members(E=64, seed=4, n_lo=2000, n_hi=200_000, k_lo=100, k_hi=500, truth_n=20_000, truth_k=300):
    """Ensemble: E members with N log-uniform in [n_lo, n_hi] and in-degree uniform in
    [k_lo, k_hi], plus one hidden "truth" member that produces the observed BOLD."""
    rng = np.random.default_rng(seed)
    ns = np.round(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), size=E))).astype(np.int64)
    ks = rng.integers(k_lo, k_hi + 1, size=E)
    mem = [(int(n), int(round(0.8 * n)), int(k)) for n, k in zip(ns, ks)]
    if truth_n:
        mem.append((int(truth_n), int(round(0.8 * truth_n)), int(truth_k)))
    return mem


def ensemble_props(members, voxels=90, seed=5, dt=0.1, member_jitter=0.3):
    """Per member: `voxels` voxels of an E and an I sub-population
    (sub_idx = (member*voxels + v)*2 + {0: E, 1: I}); conductances of the
    large-block regime rescaled by 1000/in-degree, perturbed +-member_jitter per member
    and channel, +-20% per neuron."""
    rng = np.random.default_rng(seed)
    n = sum(m[0] for m in members)
    p = base_props(n, dt)
    p[:, 0] = rng.uniform(5.0, 20.0, size=n).astype(np.float32) * np.float32(dt * 1e-3)
    p[:, 2] = 5.0
    scale = np.array([0.1, 0.1, 0.05, 0.05], np.float32)
    off = 0
    for m, (nm, ne, k) in enumerate(members):
        ni = nm - ne
        sub = np.empty(nm, np.float32)
        sub[:ne] = (m * voxels + (np.arange(ne) * voxels) // max(ne, 1)) * 2
        sub[ne:] = (m * voxels + (np.arange(ni) * voxels) // max(ni, 1)) * 2 + 1
        p[off:off + nm, 3] = sub
        fac = rng.uniform(1 - member_jitter, 1 + member_jitter, size=4).astype(np.float32)
        p[off:off + nm, 10:14] = (G_CAL * scale * fac * np.float32(1000.0 / max(k, 1)) *
                                  rng.uniform(0.8, 1.2, size=(nm, 4)).astype(np.float32))
        off += nm
    return p


def build_ensemble(E=64, voxels=90, device=0, conn_seed=1, noise_seed=3, max_delay=1, **kw):
    mem = ensemble_members(E=E, **kw)
    p = ensemble_props(mem, voxels=voxels)
    nb = NativeBlock.members(p, mem, max_delay=max_delay, w_lo=0.0, w_hi=1.0, conn_seed=conn_seed,
                             noise_seed=noise_seed, stats_exclude_blocked=True, device=device)
    return nb, p, mem
