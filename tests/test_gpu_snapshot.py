"""State checkpoint and restore (dtb_block_save_state / dtb_block_load_state).

The reference keeps no V, J or t_last across runs, so there is no golden
vector: the property checked is the one a restart needs (the oracle only
generates the network). After a restore the
block continues bit-for-bit as the saved one did, with delays still in flight,
for both delivery modes, both weight forms, host and device snapshots, into
the same block or into a fresh block with another noise seed (whose captured
batches must then be dropped), and with BOLD attached."""
import numpy as np
import pytest
import torch

from digital_twin_brain_amd import NativeBlock
from digital_twin_brain_amd import _native as nat
from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _props(n, ne, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 22), np.float32)
    p[:, 3] = np.arange(n) // 1000
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g_cal = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32)
    p[:, 10:14] = g_cal * 2.0 * rng.uniform(0.8, 1.2, size=(n, 4)).astype(np.float32)
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


_NETS = {}


def _net(pair):
    """Seeded network with delays 1-10 and weights on a 1/64 grid, so that
    arrival sums are exact whatever the order of the direct mode's atomics."""
    if pair not in _NETS:
        ref = orc.gen_fixed_indegree(6000, 4800, 120, 10, 0.0, 1.0, pair, seed=7)
        for key in ("w0", "w1"):
            if ref.get(key) is not None:
                ref[key] = (np.floor(ref[key] * 64) / 64).astype(np.float32)
        _NETS[pair] = ref
    return _NETS[pair]


def _block(pair, mode, noise_seed, bold=False):
    ref = _net(pair)
    nb = NativeBlock.from_csr(_props(6000, 4800, 3), ref["rowptr"], ref["post"], ref["kind"], ref["w0"],
                              ref["w1"] if pair else None, ref["delay"], dt=0.1, noise_seed=noise_seed)
    if mode == 1:
        nb.set_launch(deliver_mode=1, deliver_grid=256, stage_variant=2)
    if bold:
        nb.attach_bold(steps_per_sample=10, delta_t=1e-3)
    return nb


def _window(nb, steps, graph):
    """Run `steps` steps; everything a caller can observe of them."""
    ev = np.zeros(steps, np.uint64)
    sp = np.zeros(steps, np.uint32)
    fr = np.zeros((steps, nb.S), np.uint32)
    vm = np.zeros((steps, nb.S), np.float32)
    bold = np.zeros((steps // 10, nb.info()[8]), np.float64) if nb.info()[8] else None
    nb.run(steps, noise_rate=0.02, graph=graph, events=ev, spikes=sp, freqs=fr, vmeans=vm, bold=bold)
    out = dict(ev=ev, sp=sp, fr=fr, vm=vm, V=nb.get("V"), J=nb.get("J"), t_last=nb.get("t_last"),
               spikes=nb.get("spike_list"), steps=nb.steps, t=nb.t)
    if bold is not None:
        out["bold"] = bold
    return out


def _same(a, b):
    """Bit-exact, except the per-population mean V: its sums are float
    atomics across workgroups, so their last bits follow the order of
    arrival (relative error ~1e-7 per sum)."""
    assert a.keys() == b.keys()
    for k in a:
        if k == "vm":
            continue
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k], equal_nan=True), k
        else:
            assert a[k] == b[k], k
    np.testing.assert_allclose(a["vm"], b["vm"], rtol=1e-5, atol=0)


@pytest.mark.parametrize("pair,mode,where", [(False, 0, "host"), (False, 1, "device"), (True, 0, "device"),
                                             (True, 1, "host")])
def test_restore_continues_exactly(pair, mode, where):
    nb = _block(pair, mode, noise_seed=4)
    try:
        nb.run(37, noise_rate=0.02, graph=True)
        nbytes = nb.state_bytes()
        buf = np.zeros(nbytes, np.uint8) if where == "host" else torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        nb.save_state(buf)
        steps0, t0 = nb.steps, nb.t
        first = _window(nb, 40, graph=True)
        assert first["ev"][:10].min() > 0  # delays 1-10 were in flight at the snapshot
        nb.load_state(buf)
        assert (nb.steps, nb.t) == (steps0, t0)
        _same(first, _window(nb, 40, graph=True))
        # once more, eagerly launched this time
        nb.load_state(buf)
        _same(first, _window(nb, 40, graph=False))
    finally:
        nb.close()


def test_restore_into_another_block_with_another_seed():
    """A fresh block built from the same network with another noise seed,
    which has already captured a batch of the same length under that seed,
    continues as the saved block after the load."""
    a = _block(False, 1, noise_seed=4)
    b = _block(False, 1, noise_seed=99)
    try:
        a.run(23, noise_rate=0.02, graph=True)
        snap = a.save_state(torch.empty(a.state_bytes(), dtype=torch.uint8, device="cuda"))
        want = _window(a, 30, graph=True)
        other = _window(b, 30, graph=True)
        assert not np.array_equal(other["ev"], want["ev"])
        b.load_state(snap)
        _same(want, _window(b, 30, graph=True))
    finally:
        a.close()
        b.close()


def test_restore_with_bold_attached():
    nb = _block(True, 1, noise_seed=6, bold=True)
    try:
        nb.run(50, noise_rate=0.02, graph=True)
        snap = nb.save_state()
        first = _window(nb, 60, graph=True)
        assert np.isfinite(first["bold"]).all() and first["bold"].std() > 0
        nb.load_state(snap)
        _same(first, _window(nb, 60, graph=True))
    finally:
        nb.close()


def test_restore_with_external_drive():
    """A region network (delays 1-40) under the external Poisson drive, whose
    draws are keyed by seed and step like the noise: restored, it continues
    exactly (bucketed delivery, fixed-point sums)."""
    from digital_twin_brain_amd import networks

    nb, _, _, _ = networks.build_regions(R=6, total=60000, k_intra=80, k_inter=20, max_delay=40)
    try:
        nb.set_launch(deliver_mode=1, deliver_grid=256, stage_variant=2)
        assert nb.get_launch()["accum_fixed"] >= 16
        nb.run(60, noise_rate=0.0, graph=True)
        snap = nb.save_state()
        first = _window(nb, 50, graph=True)
        assert first["ev"].min() > 0 and first["fr"].sum() > 0
        nb.load_state(snap)
        _same(first, _window(nb, 50, graph=True))
    finally:
        nb.close()


def test_restore_full_size_block():
    """The benchmark's 5e9-synapse block in its benchmark launch shape
    (bucketed, LDS-sorted staging, fixed-point sums): a device snapshot
    restores it, and 100 further steps repeat bit-for-bit."""
    from digital_twin_brain_amd import networks

    nb, _ = networks.build_large()
    try:
        nb.set_launch(deliver_mode=1, deliver_grid=1024, stage_segs=256, stage_ept=16, stage_variant=2)
        assert nb.get_launch()["accum_fixed"] >= 16
        nb.run(100, noise_mode=1, graph=True)
        snap = nb.save_state(torch.empty(nb.state_bytes(), dtype=torch.uint8, device="cuda"))

        def window():
            ev = np.zeros(100, np.uint64)
            fr = np.zeros((100, nb.S), np.uint32)
            nb.run(100, noise_mode=1, graph=True, events=ev, freqs=fr)
            return ev, fr, nb.get("V"), nb.get("J"), nb.get("spike_list")

        first = window()
        assert first[0].min() > 1_000_000
        nb.load_state(snap)
        for x, y in zip(first, window()):
            assert np.array_equal(x, y)
    finally:
        nb.close()


def test_load_rejects_foreign_or_short_snapshots():
    a = _block(False, 0, noise_seed=4)
    b = _block(False, 0, noise_seed=4, bold=True)
    c = NativeBlock.fixed_indegree(_props(3000, 2400, 3), 2400, 50, max_delay=10, conn_seed=7, noise_seed=4)
    assert c.max_delay == a.max_delay == 10
    try:
        snap = a.save_state()
        with pytest.raises(nat.DTBError, match="BOLD"):
            b.load_state(snap)
        with pytest.raises(nat.DTBError, match="n=6000"):
            c.load_state(snap)
        with pytest.raises(nat.DTBError):
            a.load_state(snap[:-256])
        with pytest.raises(nat.DTBError):
            a.save_state(np.zeros(a.state_bytes() - 1, np.uint8))
        bad = snap.copy()
        bad[:8] = 0
        with pytest.raises(nat.DTBError, match="not a state snapshot"):
            a.load_state(bad)
        # the failed loads left the block usable and unchanged
        a.load_state(snap)
        assert a.steps == 0
        a.run(5, noise_rate=0.02)
    finally:
        a.close()
        b.close()
        c.close()
