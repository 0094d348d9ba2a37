"""Segment lists of bucketed delivery (launch option segment_lists): each
spike's (row, delay) segments are written once into the lists of the steps
that deliver them, instead of being read from the CSR's row pointer and
segment ends at every delay.  With int32 fixed-point accumulation the sums do
not depend on the order of the events, so every result must be bit-identical
to the spike-list path -- V, J, t_last, rasters, spike lists, per-step event
and spike counts, population counts -- including steps whose list overflows a
small cap (they read the spike lists), the first D - 1 steps after the lists
restart (a mode switch, a state load), graph replay and pair weights.  Parity
with the FP64 oracle then carries over from the spike-list path's tests (the
oracle tests with delays run with segment lists on, the default)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _props(n, rng, g_scale=3.0, pops=4):
    p = np.zeros((n, 22), np.float32)
    p[:, 0] = rng.uniform(0.002, 0.02, n)
    p[:, 3] = np.arange(n) * pops // n
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32) * g_scale
    p[:, 10:14] = g * rng.uniform(0.8, 1.2, size=(n, 4))
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


def _state(nb):
    return {k: nb.get(k) for k in ("V", "J", "t_last", "active", "spike")} | {"list": nb.get("spike_list")}


def _run_all(blocks, steps, **kw):
    recs = []
    for nb in blocks:
        S = nb.S
        r = dict(freqs=np.zeros((steps, S), np.uint32), events=np.zeros(steps, np.uint64),
                 spikes=np.zeros(steps, np.uint32))
        nb.run(steps, noise_mode=1, **r, **kw)
        recs.append(r)
    s0 = _state(blocks[0])
    for i in range(1, len(blocks)):
        for k in recs[0]:
            assert np.array_equal(recs[0][k], recs[i][k]), f"block {i}: record {k} differs"
        si = _state(blocks[i])
        for k in s0:
            assert np.array_equal(s0[k], si[k]), f"block {i}: state {k} differs"
    return recs[0]


def _fixed_indegree_blocks(n, K, D, pair, seed, caps):
    from digital_twin_brain_amd import NativeBlock

    props = _props(n, np.random.default_rng(seed))
    out = []
    for cap in caps:
        nb = NativeBlock.fixed_indegree(props, int(0.8 * n), K, max_delay=D, pair_weights=pair, conn_seed=seed,
                                        noise_seed=7)
        nb.set_launch(deliver_mode=1, accum_fixed=1, segment_lists=cap)
        out.append(nb)
    return out


@pytest.mark.parametrize("n,K,D,pair", [(20_000, 100, 10, False), (9_000, 100, 63, False), (12_000, 80, 5, True)])
def test_segment_lists_equal_spike_list_path(n, K, D, pair):
    # off, on (default capacity), on with a cap small enough to overflow on busy steps
    blocks = _fixed_indegree_blocks(n, K, D, pair, seed=n % 89 + D, caps=(0, 1, 64))
    assert blocks[0].get_launch()["segment_lists"] == 0
    assert blocks[1].get_launch()["segment_lists"] > 64
    assert blocks[2].get_launch()["segment_lists"] == 64
    tot = 0
    for steps, graph in ((1, False), (D + 3, False), (100, True), (200, True), (17, False)):
        tot += int(_run_all(blocks, steps, graph=graph)["events"].sum())
    assert tot > 0, "no events: the comparison is vacuous"
    for nb in blocks:
        nb.close()


def test_segment_lists_restart_after_mode_switch_and_state_load():
    blocks = _fixed_indegree_blocks(16_000, 100, 12, False, seed=5, caps=(0, 1))
    a, b = blocks
    _run_all(blocks, 150, graph=True)
    # batches without segment lists in between (persistent delivery, the same
    # int32 sums): the lists restart afterwards
    for nb in blocks:
        nb.set_launch(deliver_mode=2)
    _run_all(blocks, 30, graph=True)
    for nb in blocks:
        nb.set_launch(deliver_mode=1, accum_fixed=1)
    _run_all(blocks, 60, graph=True)
    # a state snapshot taken here, restored after more steps: both paths replay the same steps
    snaps = [nb.save_state() for nb in blocks]
    ref = _run_all(blocks, 40, graph=False)
    for nb, s in zip(blocks, snaps):
        nb.load_state(s)
    again = _run_all(blocks, 40, graph=False)
    for k in ref:
        assert np.array_equal(ref[k], again[k]), f"replay after state load: {k} differs"
    a.close(); b.close()


def test_segment_lists_on_short_segment_region_network():
    """Region network with delays 1-40 (segments of a few synapses): the
    configuration the lists are for."""
    from digital_twin_brain_amd import networks

    blocks = []
    for cap in (0, 1, 512):
        nb = networks.build_regions(R=12, total=150_000, k_intra=60, k_inter=15, max_delay=40)[0]
        nb.set_launch(deliver_mode=1, accum_fixed=1, segment_lists=cap)
        blocks.append(nb)
    tot = 0
    for steps, graph in ((45, False), (200, True)):
        tot += int(_run_all(blocks, steps, graph=graph)["events"].sum())
    assert tot > 0
    for nb in blocks:
        nb.close()
