"""CPU checks of the synthetic workload builders (digital_twin_brain_amd/networks.py)
against the configurations they stand for: property columns follow the
reference layout (models/block.py:152-170: noise rate, block, I_ext,
sub-population, C, T_ref, g_L, V_L, V_th, V_reset, g_ui x4, V_ui x4,
tau_ui x4)."""
import numpy as np

from digital_twin_brain_amd import networks

TAU = (2.0, 40.0, 10.0, 50.0)
V_U = (0.0, 0.0, -70.0, -100.0)


def _common_cell(p):
    assert np.all(p[:, 4] == 1.0)        # C
    assert np.all(p[:, 5] == 2.0)        # T_ref, ms
    assert np.all(p[:, 6] == np.float32(0.03))
    assert np.all(p[:, 7] == -75.0)      # V_L
    assert np.all(p[:, 8] == -50.0)      # V_th
    assert np.all(p[:, 9] == -65.0)      # V_reset
    assert np.all(p[:, 14:18] == V_U)
    assert np.all(p[:, 18:22] == TAU)


def test_plain_block_workload():
    """10,000 neurons, 8,000 excitatory then 2,000 inhibitory, g_ui ~ U[0.001, 0.05],
    10 Hz noise at dt = 0.1 ms (p = 1e-3 per step), no external current."""
    p, n_exc = networks.plain_props()
    assert p.shape == (10_000, 22) and n_exc == 8_000
    _common_cell(p)
    assert np.all(p[:8000, 3] == 0) and np.all(p[8000:, 3] == 1)
    g = p[:, 10:14]
    assert g.min() >= 0.001 and g.max() <= 0.05
    assert abs(float(g.mean()) - 0.0255) < 5e-4  # the uniform's mean
    assert np.allclose(p[:, 0], 1e-3) and np.all(p[:, 2] == 0.0)


def test_large_block_workload():
    """Per-neuron noise 5-20 Hz (p = 5e-4 .. 2e-3 per step), 1,000-neuron voxels
    as sub-populations, 80% excitatory (checked on a 200,000-neuron slice of
    the same generator)."""
    n = 200_000
    p, n_exc = networks.large_props(n=n)
    assert n_exc == 160_000
    _common_cell(p)
    assert p[:, 0].min() >= np.float32(5e-4) * (1 - 1e-6) and p[:, 0].max() <= np.float32(2e-3) * (1 + 1e-6)
    assert np.array_equal(p[:, 3], np.arange(n) // 1000)
    ratio = p[:, 10:14] / (networks.G_CAL * np.array([0.1, 0.1, 0.05, 0.05], np.float32))
    assert ratio.min() >= 0.8 - 1e-6 and ratio.max() <= 1.2 + 1e-6


def test_ensemble_workload():
    """64 members, N log-uniform over [2k, 200k], in-degree U{100..500}, 80%
    excitatory, g_ui x (1 +- 0.3) per member and channel, plus the hidden
    truth member; each member's voxels hold an E and an I sub-population."""
    mem = networks.ensemble_members()
    assert len(mem) == 65
    ns = np.array([m[0] for m in mem[:64]])
    ks = np.array([m[2] for m in mem[:64]])
    assert ns.min() >= 2000 and ns.max() <= 200_000 and ks.min() >= 100 and ks.max() <= 500
    assert all(m[1] == round(0.8 * m[0]) for m in mem)
    # log-uniform: log N spread over the whole range, not bunched
    u = (np.log(ns) - np.log(2000)) / (np.log(200_000) - np.log(2000))
    assert 0.35 < u.mean() < 0.65 and u.min() < 0.1 and u.max() > 0.9
    small = [(400, 320, 30), (700, 560, 40)]
    V = 6
    p = networks.ensemble_props(small, voxels=V)
    assert p.shape == (1100, 22)
    _common_cell(p)
    sub = p[:, 3].astype(np.int64)
    off = 0
    for m, (nm, ne, k) in enumerate(small):
        s = sub[off:off + nm]
        assert np.all(s[:ne] % 2 == 0) and np.all(s[ne:] % 2 == 1)
        vox = s // 2 - m * V
        assert vox.min() == 0 and vox.max() == V - 1
        assert np.all(np.diff(vox[:ne]) >= 0) and np.all(np.diff(vox[ne:]) >= 0)
        # per member and channel a factor in [0.7, 1.3], per neuron +-20% around it
        base = networks.G_CAL * np.array([0.1, 0.1, 0.05, 0.05], np.float32) * np.float32(1000.0 / k)
        r = p[off:off + nm, 10:14] / base
        lo, hi = r.min(axis=0), r.max(axis=0)
        assert np.all(lo >= 0.7 * 0.8 - 1e-5) and np.all(hi <= 1.3 * 1.2 + 1e-5)
        assert np.all(hi / lo <= 1.2 / 0.8 + 1e-4)  # one member factor per channel
        off += nm


def test_region_workload():
    """90 regions, lognormal (sigma 0.8) sizes summing to exactly 2e7, a 90 x 90
    Exp(1) connectome without self-loops, its source table, 80/20 E/I per
    region."""
    sizes = networks.region_sizes()
    assert len(sizes) == 90 and int(sizes.sum()) == 20_000_000 and sizes.min() >= 1000
    assert abs(float(np.log(sizes).std()) - 0.8) < 0.2
    W = networks.connectome()
    assert W.shape == (90, 90) and np.all(np.diag(W) == 0.0) and W.min() >= 0.0
    off = W[~np.eye(90, dtype=bool)]
    assert abs(float(off.mean()) - 1.0) < 0.05 and abs(float(off.std()) - 1.0) < 0.08
    cdf = networks.source_cdf(W)
    assert cdf.dtype == np.uint64 and np.all(cdf[:, -1] == 1 << 32)
    assert np.all(np.diff(cdf.astype(np.int64), axis=1) >= 0)
    prob = np.diff(np.concatenate([np.zeros((90, 1), np.int64), cdf.astype(np.int64)], axis=1), axis=1) / 2.0 ** 32
    assert np.all(np.diag(prob) == 0.0)  # a region never draws inter-region inputs from itself
    assert np.allclose(prob, W / W.sum(axis=1, keepdims=True), atol=2.0 ** -31)
    small = np.array([50, 120, 31])
    p, regs = networks.region_props(small, k_total=500)
    assert p.shape == (201, 22)
    _common_cell(p)
    off_n = 0
    for r, (nr, ne) in enumerate(regs):
        assert nr == small[r] and ne == round(0.8 * nr)
        assert np.all(p[off_n:off_n + ne, 3] == 2 * r) and np.all(p[off_n + ne:off_n + nr, 3] == 2 * r + 1)
        off_n += nr
