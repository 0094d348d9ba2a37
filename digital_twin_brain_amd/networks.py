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
region ensemble E whole-brain members (replicas of one 90-region network with
                a shared connectome) plus a hidden truth replica.
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


# --------------------------------------------------------------------------
# Regions: one GPU's share of a whole-brain model.  R regions with lognormal
# sizes, each neuron receiving k_intra synapses from its own region and
# k_inter from other regions picked through a seeded connectome (Exp(1)
# weights, no self-loops), delays U{1..max_delay}, plus an external Poisson
# drive standing for the regions on other GPUs.

def region_sizes(R=90, total=20_000_000, sigma=0.8, seed=11, min_size=1000):
    """Lognormal(0, sigma) sizes scaled to `total` neurons (exact sum)."""
    rng = np.random.default_rng(seed)
    x = rng.lognormal(0.0, sigma, size=R)
    sizes = np.maximum(np.floor(x / x.sum() * total).astype(np.int64), min_size)
    sizes[np.argmax(sizes)] += total - int(sizes.sum())
    return sizes


def connectome(R=90, seed=12):
    """R x R weights, Exp(1) off the diagonal, 0 on it."""
    rng = np.random.default_rng(seed)
    W = rng.exponential(1.0, size=(R, R))
    np.fill_diagonal(W, 0.0)
    return W


def source_cdf(W):
    """Per target region, cumulative source probabilities scaled to 2^32
    (uint64; rows non-decreasing and ending at exactly 2^32)."""
    W = np.asarray(W, np.float64)
    P = W / W.sum(axis=1, keepdims=True)
    cdf = np.floor(np.cumsum(P, axis=1) * 4294967296.0).astype(np.uint64)
    cdf = np.maximum.accumulate(np.minimum(cdf, np.uint64(1 << 32)), axis=1)
    cdf[:, -1] = np.uint64(1 << 32)
    return cdf


def region_props(sizes, k_total=500, seed=13, dt=0.1):
    """Per region an E (80%) and an I population (sub_idx = 2 region + {0, 1});
    conductances of the large-block regime rescaled by 1000/k_total."""
    rng = np.random.default_rng(seed)
    n = int(np.sum(sizes))
    p = base_props(n, dt)
    p[:, 0] = rng.uniform(5.0, 20.0, size=n).astype(np.float32) * np.float32(dt * 1e-3)
    p[:, 2] = 5.0
    scale = np.array([0.1, 0.1, 0.05, 0.05], np.float32) * np.float32(1000.0 / k_total)
    p[:, 10:14] = G_CAL * scale * rng.uniform(0.8, 1.2, size=(n, 4)).astype(np.float32)
    off = 0
    regions = []
    for r, nr in enumerate(sizes):
        ne = int(round(0.8 * nr))
        p[off:off + ne, 3] = 2 * r
        p[off + ne:off + nr, 3] = 2 * r + 1
        regions.append((int(nr), ne))
        off += nr
    return p, regions


def build_regions(R=90, total=20_000_000, k_intra=400, k_inter=100, max_delay=40, ext_fraction=0.2,
                  target_hz=8.0, dt=0.1, device=0, conn_seed=1, noise_seed=3, sizes=None, W=None):
    """Region network with external drive: the external Poisson stream is
    `ext_fraction` of all input events at `target_hz` (each event the mean
    synaptic weight, 0.5).  Returns (block, props, regions, cdf)."""
    sizes = region_sizes(R, total) if sizes is None else np.asarray(sizes, np.int64)
    W = connectome(len(sizes)) if W is None else W
    cdf = source_cdf(W)
    p, regs = region_props(sizes, k_intra + k_inter, dt=dt)
    spec = [(n, ne, k_intra, k_inter) for n, ne in regs]
    nb = NativeBlock.regions(p, spec, cdf, max_delay=max_delay, w_lo=0.0, w_hi=1.0, conn_seed=conn_seed, dt=dt,
                             noise_seed=noise_seed, stats_exclude_blocked=True, device=device)
    internal = (k_intra + k_inter) * target_hz * dt * 1e-3  # input events per neuron per step
    nb.set_external(ext_fraction / (1.0 - ext_fraction) * internal, 0.5)
    return nb, p, spec, cdf


def build_region_ensemble(E=256, R=90, member_n=20_000, k_intra=80, k_inter=20, max_delay=40, ext_fraction=0.2,
                          target_hz=8.0, truth=True, dt=0.1, device=0, conn_seed=1, noise_seed=3, min_size=40,
                          size_seed=11, conn_matrix_seed=12, props_seed=13):
    """Ensemble of whole-brain members for data assimilation: E replicas (plus
    a hidden truth replica) of one R-region network -- lognormal(sigma 0.8)
    region sizes summing to member_n, k_intra inputs from the own region and
    k_inter from the others through one Exp(1) connectome shared by all
    replicas, delays 1..max_delay, the external Poisson drive of the region
    share.  Replicas differ in their synapse and noise draws and in the +-20%
    per-neuron conductance jitter.  Sub-population of replica m, region r:
    2 (m R + r) + {0: E, 1: I}; one BOLD voxel per region.  Returns
    (block, props, spec, cdf) with spec the (n, n_exc, k_intra, k_inter) of
    one replica's regions."""
    sizes = region_sizes(R, member_n, min_size=min_size, seed=size_seed)
    cdf = source_cdf(connectome(R, seed=conn_matrix_seed))
    reps = E + (1 if truth else 0)
    parts = []
    spec = None
    for m in range(reps):
        p, regs = region_props(sizes, k_intra + k_inter, seed=props_seed + m, dt=dt)
        p[:, 3] += 2 * R * m
        parts.append(p)
        spec = [(n, ne, k_intra, k_inter) for n, ne in regs]
    props = np.concatenate(parts)
    nb = NativeBlock.region_replicas(props, spec, reps, cdf, max_delay=max_delay, w_lo=0.0, w_hi=1.0,
                                     conn_seed=conn_seed, dt=dt, noise_seed=noise_seed, stats_exclude_blocked=True,
                                     device=device)
    internal = (k_intra + k_inter) * target_hz * dt * 1e-3
    nb.set_external(ext_fraction / (1.0 - ext_fraction) * internal, 0.5)
    return nb, props, spec, cdf
