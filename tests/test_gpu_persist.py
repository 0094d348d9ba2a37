"""Persistent delivery (deliver_mode 2: one launch per batch of steps, a
grid barrier per step) against bucketed fixed-point delivery (mode 1):
the same int32 sums in another order, so every result is bit-identical --
V, J, t_last, rasters, spike lists, per-step event and spike counts,
population counts, attached BOLD (mean V and I_syn, float atomic sums,
within 1e-5) -- across batch sizes, graph
replay, delays, pair weights and switches between the modes mid-run.
Parity with the FP64 oracle then carries over from the bucketed path's
tests; the plain-block test runs the oracle comparison in this mode too."""
import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _props(n, ne, rng, g_scale=3.0, pops=4):
    p = np.zeros((n, 22), np.float32)
    p[:, 0] = rng.uniform(0.002, 0.02, n)
    p[:, 3] = np.arange(n) * pops // n
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32) * g_scale
    p[:, 10:14] = g * rng.uniform(0.8, 1.2, size=(n, 4))
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


def _pair(n, ne, K, D, pair, seed, mode, bold=False):
    from digital_twin_brain_amd import NativeBlock

    rng = np.random.default_rng(seed)
    props = _props(n, ne, rng)
    out = []
    for m in (1, mode):
        nb = NativeBlock.fixed_indegree(props, ne, K, max_delay=D, pair_weights=pair, conn_seed=seed, noise_seed=7)
        nb.set_launch(deliver_mode=1, accum_fixed=1)
        if m != 1:
            nb.set_launch(deliver_mode=m)
        if bold:
            nb.attach_bold(steps_per_sample=10, delta_t=1e-3)
        out.append(nb)
    return out


def _state(nb):
    return {k: nb.get(k) for k in ("V", "J", "t_last", "active", "spike")} | {"list": nb.get("spike_list")}


def _run_both(a, b, steps, **kw):
    recs = []
    for nb in (a, b):
        S = nb.S
        r = dict(freqs=np.zeros((steps, S), np.uint32), vmeans=np.zeros((steps, S), np.float32),
                 events=np.zeros(steps, np.uint64), spikes=np.zeros(steps, np.uint32))
        if getattr(nb, "nvox", 0):
            r["bold"] = np.zeros((steps // nb.bold_sps, nb.nvox), np.float64)
        nb.run(steps, noise_mode=1, **r, **kw)
        recs.append(r)
    for k in recs[0]:
        if k == "vmeans":
            # mean V per population is a float atomic sum: its order varies from run to run
            np.testing.assert_allclose(recs[0][k], recs[1][k], rtol=1e-5, err_msg="record vmeans differs")
        else:
            assert np.array_equal(recs[0][k], recs[1][k]), f"record {k} differs"
    sa, sb = _state(a), _state(b)
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), f"state {k} differs"
    return recs[0]


# the first five have short (row, delay, workgroup) segments (each thread walks
# its own); the last two long ones (scan-balanced chunks: 3,000 x 400 over 12
# workgroups averages 33 synapses per segment)
@pytest.mark.parametrize("n,K,D,pair", [(10_000, 100, 1, False), (2_000, 50, 6, False), (4_000, 80, 4, True),
                                        (40_000, 60, 3, False), (9_000, 100, 63, False), (3_000, 400, 1, False),
                                        (3_000, 400, 2, True)])
def test_persistent_equals_bucketed_fixed_point(n, K, D, pair):
    a, b = _pair(n, int(0.8 * n), K, D, pair, seed=n % 97 + D, mode=2)
    assert b.get_launch()["deliver_mode"] == 2
    tot = 0
    for steps, graph in ((1, False), (7, False), (50, True), (200, True), (30, False)):
        r = _run_both(a, b, steps, graph=graph)
        tot += int(r["events"].sum())
    assert tot > 0, "no events: the comparison is vacuous"
    print(f"n={n} K={K} D={D} pair={pair}: {tot} events, launch {b.get_launch()}")
    a.close(); b.close()


def test_persistent_with_bold_and_mode_switches():
    a, b = _pair(10_000, 8000, 100, 5, False, seed=3, mode=2, bold=True)
    _run_both(a, b, 100, graph=True)
    # mid-run switches: bucketed, atomics-free fallback batches (I_syn means) and back
    b.set_launch(deliver_mode=1)
    _run_both(a, b, 40, graph=True)
    b.set_launch(deliver_mode=2)
    _run_both(a, b, 60, graph=False)
    im_a = np.zeros((20, a.S), np.float32)
    im_b = np.zeros((20, b.S), np.float32)
    a.run(20, noise_mode=1, imeans=im_a)
    b.run(20, noise_mode=1, imeans=im_b)  # runs on the bucketed kernels
    # a float atomic sum of terms of both signs, like the mean V
    np.testing.assert_allclose(im_a, im_b, rtol=1e-5, atol=1e-5 * np.abs(im_a).max())
    _run_both(a, b, 100, graph=True)
    a.close(); b.close()


def test_persistent_matches_oracle_spike_lists_and_events():
    """A short recurrent run in mode 2 against the FP64 oracle directly:
    compacted spike lists and event counts exact, V and J within 1e-5."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K, D = 6000, 4800, 80, 3
    rng = np.random.default_rng(17)
    props = _props(n, ne, rng)
    nb = NativeBlock.fixed_indegree(props, ne, K, max_delay=D, conn_seed=5, noise_seed=11)
    nb.set_launch(deliver_mode=2)
    csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
               w0=nb.get("w0"), w1=None)
    ob = orc.OracleBlock(props, csr, 0.1)
    rp = csr["rowptr"]
    row = np.repeat(np.arange(n), np.diff(rp))
    seg = np.zeros((n, D + 1), np.int64)
    np.add.at(seg, (row, csr["delay"].astype(np.int64)), 1)
    hist = []
    ev = np.zeros(1, np.uint64)
    for s in range(300):
        nb.run(1, noise_mode=1, events=ev, graph=False)
        ob.run(noise_mode=1, seed=11)
        spk = nb.get("spike").astype(bool)
        hist.append(spk)
        assert np.array_equal(nb.get("active").astype(bool), ob.get("active").astype(bool)), f"raster at {s + 1}"
        assert np.array_equal(nb.get("spike_list"), orc.compact(ob.get("spike"))), f"spike list at {s + 1}"
        want = sum(int(seg[hist[s - d + 1], d].sum()) for d in range(1, D + 1) if s - d + 1 >= 0)
        assert int(ev[0]) == want, f"events at {s + 1}"
        np.testing.assert_allclose(nb.get("V"), ob.get("V"), rtol=1e-5)
        Jo = ob.get("J")
        np.testing.assert_allclose(nb.get("J"), Jo, rtol=1e-5, atol=1e-5 * np.abs(Jo).max())
    nb.close()


def test_persistent_refuses_blocks_it_cannot_hold():
    from digital_twin_brain_amd import NativeBlock, _native as nat

    rng = np.random.default_rng(2)
    n = 300_000  # above 32 workgroups x 8,192 neurons of LDS slice
    props = _props(n, int(0.8 * n), rng)
    nb = NativeBlock.fixed_indegree(props, int(0.8 * n), 20, max_delay=1, conn_seed=1, noise_seed=1)
    with pytest.raises(nat.DTBError):
        nb.set_launch(deliver_mode=2)
    assert nb.get_launch()["deliver_mode"] != 2
    nb.close()
