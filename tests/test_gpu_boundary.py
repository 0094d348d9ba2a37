"""GPU parity of the drop-in boundary's remaining pieces, through the C ABI:

- block(src=...) forced spikes (models/block.py:38-50,104-109) against the
  golden vectors the reference block.py wrote (tests/golden/block_src64.npz);
- frozen sources (the neurons of an npz input block, cuda/src/block.cpp:195-221)
  and block.py sources against the oracle;
- block.update_conn_weight (models/block.py:172-177): single channels, both
  channels, and a synapse that did not exist, against the oracle;
- LocalBlockWrapper on a directory with an input block, and its run()
  outputs vmean, imean and sample_for_show (cuda/src/stage/stat_result.cpp:
  223-304, cuda/src/stage/sample.cpp:8-34) against the oracle;
- simulation.run end to end.

Tolerances as the other parity tests: rasters identical (until a reported
first divergence where stated), V within 1e-5 relative, J within 1e-5 of
max|J|, counts exact.
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _load(golden_dir, name):
    z = np.load(os.path.join(golden_dir, f"block_{name}.npz"))
    return {k: z[k] for k in z.files}


def _sparse_w(c):
    """[4, N_post, N_pre] sparse w_uij of a golden case (generation/read_block.py:64-90)."""
    n = int(c["n"])
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(c["rowptr"]))
    post = c["post"].astype(np.int64)
    off = c["kind"].astype(np.int64) * 2
    idx = np.stack([np.concatenate([off, off + 1]), np.concatenate([post, post]), np.concatenate([pre, pre])])
    vals = np.concatenate([c["w0"], c["w1"]]).astype(np.float32)
    return torch.sparse_coo_tensor(torch.from_numpy(idx), torch.from_numpy(vals), (4, n, n)).coalesce()


def _lif_props(n, rng, g_scale=3.0, sub=None, noise=(0.002, 0.01)):
    p = np.zeros((n, 22), np.float32)
    p[:, 0] = rng.uniform(*noise, size=n)
    p[:, 3] = 0 if sub is None else sub
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32) * g_scale
    p[:, 10:14] = g * rng.uniform(0.8, 1.2, size=(n, 4))
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


def _drop_inputs(n, csr, targets):
    row = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    keep = ~np.isin(csr["post"], targets)
    out = {k: (v[keep] if isinstance(v, np.ndarray) and v.shape[:1] == keep.shape else v) for k, v in csr.items()}
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, row[keep] + 1, 1)
    out["rowptr"] = np.cumsum(rp)
    return out


def _keep_synapses(n, csr, keep):
    row = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    out = {k: (v[keep] if isinstance(v, np.ndarray) and v.shape[:1] == keep.shape else v) for k, v in csr.items()}
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, row[keep] + 1, 1)
    out["rowptr"] = np.cumsum(rp)
    return out


def test_dropin_block_src_matches_reference_golden(golden_dir):
    from digital_twin_brain_amd.block import block

    c = _load(golden_dir, "src64")
    n, steps = int(c["n"]), int(c["steps"])
    B = block(torch.from_numpy(c["props"]), _sparse_w(c), delta_t=float(c["dt"]),
              src=torch.from_numpy(c["src"].astype(bool)), seed=int(c["seed"]))
    assert np.array_equal(B.src_neuron.numpy(), c["src_neuron"])
    act = np.unpackbits(c["active"], axis=1)[:, :n]
    spk = np.unpackbits(c["spike"], axis=1)[:, :n]
    snaps = set(int(s) for s in c["snap_steps"])
    diverged = None
    for s in range(1, steps + 1):
        sa, mv = B.run(noise_rate=float(c["noise_rate"]), isolated=bool(c["isolated"]))
        a = B.active.numpy().astype(np.uint8)
        if diverged is None and (not np.array_equal(a, act[s - 1]) or
                                 not np.array_equal(B._nb.get("spike"), spk[s - 1])):
            diverged = s
        if diverged is not None:
            continue
        # the forced spikes of the sources are exactly the schedule's column s - 1
        assert np.array_equal(a[c["src_neuron"]], c["src"][:, s - 1]), f"step {s}"
        assert np.array_equal(sa.numpy(), c["sum_act"][s - 1]), f"step {s}"
        np.testing.assert_allclose(mv.numpy(), c["mean_v"][s - 1], rtol=RTOL)
        np.testing.assert_allclose(B.V_i.numpy(), c["V"][s - 1], rtol=RTOL)
        if s in snaps:
            J = c[f"J_{s}"]
            np.testing.assert_allclose(B.J_ui.numpy(), J, rtol=RTOL, atol=RTOL * np.abs(J).max())
            I = c[f"I_ui_{s}"]
            np.testing.assert_allclose(B.I_ui.numpy(), I, rtol=RTOL, atol=RTOL * np.abs(I).max())
    print(f"src64: first divergence {diverged}")
    assert diverged is None or diverged > 150, f"rasters diverge at step {diverged}"
    assert B.iter == steps
    # the reference raised at its step T (= steps + 1): here column T - 1 still drives step T, and
    # only a step past the schedule's columns raises
    B.run(noise_rate=float(c["noise_rate"]))
    assert np.array_equal(B.active.numpy()[c["src_neuron"]], c["src"][:, steps])
    with pytest.raises(IndexError):
        B.run(noise_rate=float(c["noise_rate"]))


def test_sources_and_frozen_sources_match_oracle():
    """Plain sources (membrane dynamics and noise kept) and frozen ones (an
    input block's neurons: V held, no noise) against the oracle for 300
    steps: every raster identical, the frozen V unchanged, spikes of frozen
    sources exactly the schedule."""
    from digital_twin_brain_amd import NativeBlock

    rng = np.random.default_rng(21)
    n, ne = 2000, 1600
    srcs = np.arange(0, 2000, 10)  # 200 sources, no afferents
    frozen = (np.arange(srcs.shape[0]) % 2).astype(np.uint8)
    csr = _drop_inputs(n, orc.gen_fixed_indegree(n, ne, 40, 3, 0.0, 1.0, False, seed=22), srcs)
    # dyadic weights: float sums are exact in any order, so graph and direct runs agree bitwise
    csr["w0"] = (np.floor(csr["w0"] * 256) / 256).astype(np.float32)
    p = _lif_props(n, rng, sub=np.arange(n) // 500)
    T = 300
    ev = np.nonzero(rng.random((srcs.shape[0], T)) < 0.05)
    steps_, neurons_ = ev[1].astype(np.uint32), srcs[ev[0]].astype(np.uint32)
    nb = NativeBlock.from_csr(p, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], None, csr["delay"], dt=0.1,
                              noise_seed=8)
    nb.set_sources(srcs, steps=steps_, neurons=neurons_, frozen=frozen)
    ob = orc.OracleBlock(p, nb_csr(nb), 0.1)
    ob.set_sources(srcs, steps_, neurons_, frozen=frozen)
    V0 = nb.get("V")
    sched = np.zeros((T, n), bool)
    sched[steps_, neurons_] = True
    freqs = np.zeros(nb.S, np.uint32)
    for s in range(1, T + 1):
        nb.run(1, noise_mode=1, graph=False, freqs=freqs)
        sa, _ = ob.run(noise_mode=1, seed=8)
        a, sp = nb.get("active").astype(bool), nb.get("spike").astype(bool)
        assert np.array_equal(a, ob.get("active").astype(bool)), f"threshold raster at step {s}"
        assert np.array_equal(sp, ob.get("spike").astype(bool)), f"delivered raster at step {s}"
        assert np.array_equal(a[srcs], sched[s - 1, srcs]), f"forced spikes at step {s}"
        fz = srcs[frozen == 1]
        assert np.array_equal(sp[fz], sched[s - 1, fz]), f"frozen sources drew noise at step {s}"
        assert np.array_equal(freqs, sa.astype(np.uint32))
        V = nb.get("V")
        np.testing.assert_allclose(V, ob.get("V"), rtol=RTOL)
        J = ob.get("J")
        np.testing.assert_allclose(nb.get("J"), J, rtol=RTOL, atol=RTOL * np.abs(J).max())
    assert np.array_equal(nb.get("V")[srcs[frozen == 1]], V0[srcs[frozen == 1]])
    # a graph-replayed batch reproduces the direct launches
    nb2 = NativeBlock.from_csr(p, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], None, csr["delay"], dt=0.1,
                               noise_seed=8)
    nb2.set_sources(srcs, steps=steps_, neurons=neurons_, frozen=frozen)
    nb2.run(T, noise_mode=1, graph=True)
    assert np.array_equal(nb2.get("V"), nb.get("V")) and np.array_equal(nb2.get("J"), nb.get("J"))


def nb_csr(nb):
    return dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
                w0=nb.get("w0"), w1=nb.get("w1") if nb.pair else None)


def test_update_conn_weight_matches_oracle():
    """block.update_conn_weight on a dense w_uij: single channels (the block's
    pair weights change independently), both channels of a synapse, and an
    entry with no synapse yet (the network is rebuilt and the state carries
    over), against the oracle given the same connectivity change."""
    from digital_twin_brain_amd.block import block

    rng = np.random.default_rng(31)
    n, ne = 64, 48
    csr = orc.gen_fixed_indegree(n, ne, 12, 1, 0.0, 1.0, True, seed=32)
    for k in ("w0", "w1"):
        csr[k] = (np.floor(csr[k] * 256) / 256).astype(np.float32)
    # a dense w_uij holds one entry per (channel, post, pre): keep the first of repeated pairs
    pre = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    _, first = np.unique(pre * n + csr["post"], return_index=True)
    keep = np.zeros(pre.shape[0], bool)
    keep[first] = True
    csr = _keep_synapses(n, csr, keep)
    p = _lif_props(n, rng, g_scale=1.5)
    pre = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    W = torch.zeros(4, n, n, dtype=torch.float32)
    ch0 = csr["kind"].astype(np.int64) * 2
    W[ch0, csr["post"], pre] = torch.from_numpy(csr["w0"])
    W[ch0 + 1, csr["post"], pre] = torch.from_numpy(csr["w1"])
    B = block(torch.from_numpy(p), W, delta_t=0.1, seed=5)
    ob = orc.OracleBlock(p, csr, 0.1)
    rate = 2.0 ** -5

    def compare(steps, tag):
        for s in range(steps):
            B.run(noise_rate=rate)
            ob.run(noise_rate=rate, seed=5)
            assert np.array_equal(B.active.numpy(), ob.get("active").astype(bool)), f"{tag} step {s + 1}"
            np.testing.assert_allclose(B.V_i.numpy(), ob.get("V"), rtol=RTOL, err_msg=f"{tag} step {s + 1}")
            J = ob.get("J")
            np.testing.assert_allclose(B.J_ui.numpy(), J, rtol=RTOL, atol=RTOL * np.abs(J).max())

    compare(50, "before")
    # requests: 8 single channels, 4 synapses with both channels, one new synapse (AMPA only)
    sel = rng.choice(pre.shape[0], 12, replace=False)
    chs, posts, pres, ws = [], [], [], []
    for j, e in enumerate(sel):
        c0 = int(csr["kind"][e]) * 2
        for c in ([c0 + int(rng.integers(0, 2))] if j < 8 else [c0, c0 + 1]):
            chs.append(c); posts.append(int(csr["post"][e])); pres.append(int(pre[e]))
            ws.append(float(rng.integers(0, 512)) / 256)
    taken = set(zip(pre.tolist(), csr["post"].tolist()))
    new_pre, new_post = next((a, b) for a in range(ne) for b in range(n) if a != b and (a, b) not in taken)
    chs.append(0); posts.append(new_post); pres.append(new_pre); ws.append(0.75)
    B.update_conn_weight((torch.tensor(chs), torch.tensor(posts), torch.tensor(pres)), torch.tensor(ws))
    # the same change on the oracle's CSR
    w0, w1 = csr["w0"].copy(), csr["w1"].copy()
    for c, q, a, x in zip(chs[:-1], posts[:-1], pres[:-1], ws[:-1]):
        hit = np.flatnonzero((pre == a) & (csr["post"] == q) & (csr["kind"] == c // 2))
        arr = w0 if c % 2 == 0 else w1
        arr[hit] = 0.0
        arr[hit[0]] = x
    pre2 = np.concatenate([pre, [new_pre]])
    post2 = np.concatenate([csr["post"], [new_post]]).astype(np.int32)
    kind2 = np.concatenate([csr["kind"], [0]]).astype(np.uint8)
    w02 = np.concatenate([w0, [0.75]]).astype(np.float32)
    w12 = np.concatenate([w1, [0.0]]).astype(np.float32)
    o = np.lexsort((kind2, post2, pre2))
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, pre2 + 1, 1)
    want = dict(rowptr=np.cumsum(rp), post=post2[o], kind=kind2[o], w0=w02[o], w1=w12[o])
    ob.replace_csr(want)
    got = nb_csr(B._nb)
    gpre = np.repeat(np.arange(n), np.diff(got["rowptr"]))
    wpre = np.repeat(np.arange(n), np.diff(want["rowptr"]))
    assert sorted(zip(gpre, got["post"], got["kind"], got["w0"], got["w1"])) == \
        sorted(zip(wpre, want["post"], want["kind"], want["w0"], want["w1"]))
    assert float(B.w_uij[0, new_post, new_pre]) == 0.75
    compare(150, "after")


def _input_block_dir(tmp, rng, n_in=60, sizes=(400, 300)):
    """Reference-format directory: block 0 an input block, blocks 1-2 LIF
    blocks fed by it and by each other."""
    from digital_twin_brain_amd.wrapper import save_block_dir

    n = n_in + sum(sizes)
    p = _lif_props(n, rng, g_scale=3.0)
    p[:, 3] = np.concatenate([np.zeros(n_in), 1 + (np.arange(sizes[0]) >= 0.8 * sizes[0]),
                              3 + (np.arange(sizes[1]) >= 0.8 * sizes[1])])
    K = 30
    post = np.repeat(np.arange(n_in, n), K)
    pre = rng.integers(0, n, size=post.shape[0])
    pre = np.where(pre == post, (pre + 1) % n, pre)
    # populations 1, 3 excitatory, 2, 4 inhibitory; the input block (0) excitatory
    sp = p[pre, 3].astype(np.int64)
    kind = ((sp > 0) & (sp % 2 == 0)).astype(np.uint8)
    o = np.lexsort((post, pre))
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, pre + 1, 1)
    csr = dict(rowptr=np.cumsum(rp), post=post[o].astype(np.int32), kind=kind[o],
               w0=rng.random(post.shape[0]).astype(np.float32)[o], w1=None, delay=None)
    ts = np.sort(rng.integers(0, 400, size=400)).astype(np.uint32)
    nid = rng.integers(0, n_in, size=400).astype(np.uint32)
    save_block_dir(tmp, p, csr, [n_in, *sizes], inputs={0: (ts, nid)})
    return n, n_in, ts, nid


def test_local_wrapper_with_input_block_runs_like_oracle(tmp_path):
    from digital_twin_brain_amd.wrapper import LocalBlockWrapper

    rng = np.random.default_rng(41)
    n, n_in, ts, nid = _input_block_dir(str(tmp_path), rng)
    bw = LocalBlockWrapper("local", str(tmp_path), 0.1, noise_seed=6)
    nb = bw.nb
    ob = orc.OracleBlock(nb.get("props"), nb_csr(nb), 0.1)
    ob.set_sources(np.arange(n_in), ts, nid, frozen=1)
    steps = 400
    out = list(bw.run(steps, freqs=True, vmean=True))
    for s in range(steps):
        sa, mv = ob.run(noise_mode=1, seed=6)
        assert np.array_equal(out[s][0].cpu().numpy(), sa.astype(np.int32)), f"step {s + 1}"
        np.testing.assert_allclose(out[s][1].cpu().numpy(), mv, rtol=RTOL)
    inp = np.stack([o[0].cpu().numpy()[0] for o in out])  # population 0 = the input block
    # forced spikes counted per step (duplicates of one neuron at one timestamp fire once)
    want = np.array([np.unique(nid[ts == t]).shape[0] for t in range(steps)])
    assert np.array_equal(inp, want)
    assert sum(int(o[0][1:].sum()) for o in out) > 0


def test_wrapper_run_vmean_imean_samples_match_oracle(tmp_path):
    """LocalBlockWrapper.run(freqs, vmean, sample_for_show, imean): per-step
    population counts, mean V and mean I_syn (after the step's delivery, as
    the engine's stat_Imeans follows update_I_synaptic), and the delivered
    spike and V of sampled neurons in the caller's order."""
    from digital_twin_brain_amd.wrapper import LocalBlockWrapper

    rng = np.random.default_rng(51)
    n, n_in, ts, nid = _input_block_dir(str(tmp_path), rng)
    bw = LocalBlockWrapper("local", str(tmp_path), 0.1, noise_seed=7)
    nb = bw.nb
    ob = orc.OracleBlock(nb.get("props"), nb_csr(nb), 0.1)
    ob.set_sources(np.arange(n_in), ts, nid, frozen=1)
    samples = torch.tensor([700, 3, 455, 61, 62, 300, 3], dtype=torch.int64)  # unsorted, one repeat
    bw.set_samples(samples)
    sub = nb.get("props")[:, 3].astype(np.int64)
    steps = 300
    out = list(bw.run(steps, freqs=True, vmean=True, sample_for_show=True, imean=True))
    assert len(out[0]) == 5
    for s in range(steps):
        sa, mv = ob.run(noise_mode=1, seed=7)
        f, vm, ss, sv, im = (x.cpu().numpy() for x in out[s])
        assert np.array_equal(f, sa.astype(np.int32)), f"step {s + 1}"
        np.testing.assert_allclose(vm, mv, rtol=RTOL)
        I = ob.get("I_syn")
        want_i = np.bincount(sub, weights=I) / np.bincount(sub)
        scale = np.bincount(sub, weights=np.abs(I)) / np.bincount(sub)
        assert (np.abs(im - want_i) <= RTOL * scale + 1e-12).all(), f"imean at step {s + 1}"
        assert np.array_equal(ss, ob.get("spike")[samples.numpy()]), f"sampled spikes at step {s + 1}"
        np.testing.assert_allclose(sv, ob.get("V")[samples.numpy()], rtol=RTOL)


def test_simulation_run_end_to_end(tmp_path):
    """simulation.run (simulation/simulation.py:301-395): the 3200-step warm-up,
    then observation windows with hyper-parameter pushes, freqs and BOLD
    written as .npy; the outputs agree with a second simulation driven
    through evolve with the same pushes."""
    from digital_twin_brain_amd import NativeBlock
    from digital_twin_brain_amd.simulation import simulation

    rng = np.random.default_rng(61)
    n = 2000
    sub = np.arange(n) // 100  # 20 populations -> 2 voxels of 10 (column layout)
    p = _lif_props(n, rng, sub=sub)
    csr = orc.gen_fixed_indegree(n, 1600, 40, 1, 0.0, 1.0, False, seed=62)

    def make():
        nb = NativeBlock.from_csr(p, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], dt=1.0, noise_seed=3,
                                  stats_exclude_blocked=True)
        return simulation(network=nb, column=True, dt=1.0, write_path=str(tmp_path))

    hp = torch.from_numpy(rng.uniform(0.8, 1.2, size=(1 + 3, 20, 1)).astype(np.float32))
    sim = make()
    assert sim.num_voxels == 2
    bolds = sim.run(step=40, observation_time=3, hp_index=[10], hp_total=hp)
    assert bolds.shape == (3, 2) and np.isfinite(bolds).all()
    fa = np.load(os.path.join(str(tmp_path), "freqs_after_assim.npy"))
    assert fa.shape == (3, 40, 20)
    # the same sequence by hand
    sim2 = make()
    for k in [10]:
        sim2.gamma_initialize(k, sim2.population_id)
    info = torch.stack(torch.meshgrid(sim2.population_id.cpu(), torch.tensor([10]), indexing="ij"),
                       dim=-1).reshape((-1, 2))
    sim2.block_model.mul_property_by_subblk(info, hp[0].reshape(-1).float())
    sim2.evolve(3200)
    for i in range(3):
        sim2.block_model.mul_property_by_subblk(info, hp[1 + i].reshape(-1).float())
        out = sim2.evolve(40)
        assert np.array_equal(out[0].cpu().numpy(), fa[i])
        np.testing.assert_array_equal(out[-1].cpu().numpy().astype(np.float32), bolds[i])


def test_dropin_run_steps_equals_repeated_run():
    """block.run_steps(K) (one graph-replayed library call) returns row by row
    what K calls of block.run() return, and leaves the same state (dyadic
    weights: every delivery order gives the same sums)."""
    from digital_twin_brain_amd.block import block

    rng = np.random.default_rng(41)
    n, ne = 3000, 2400
    csr = orc.gen_fixed_indegree(n, ne, 60, 1, 0.0, 1.0, True, seed=42)
    for k in ("w0", "w1"):
        csr[k] = (np.floor(csr[k] * 256) / 256).astype(np.float32)
    pre = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    _, first = np.unique(pre * n + csr["post"], return_index=True)
    keep = np.zeros(pre.shape[0], bool)
    keep[first] = True
    csr = _keep_synapses(n, csr, keep)
    pre = np.repeat(np.arange(n), np.diff(csr["rowptr"]))
    p = _lif_props(n, rng, g_scale=1.5)
    idx = torch.from_numpy(np.stack([csr["kind"].astype(np.int64) * 2, csr["post"], pre]))
    idx = torch.cat([idx, idx + torch.tensor([[1], [0], [0]])], dim=1)
    W = torch.sparse_coo_tensor(idx, torch.from_numpy(np.concatenate([csr["w0"], csr["w1"]])), (4, n, n))
    a = block(torch.from_numpy(p), W, delta_t=0.1, seed=9)
    b = block(torch.from_numpy(p), W, delta_t=0.1, seed=9)
    K = 300
    rows = [a.run(noise_rate=2.0 ** -6) for _ in range(K)]
    fa = torch.stack([r[0] for r in rows]).numpy()
    va = torch.stack([r[1] for r in rows]).numpy()
    fb, vb = (x.numpy() for x in b.run_steps(K, noise_rate=2.0 ** -6))
    assert fa.shape == fb.shape == (K, a.sub_idx.max().item() + 1)
    assert fa.sum() > 0
    assert np.array_equal(fa, fb)
    np.testing.assert_allclose(va, vb, rtol=1e-6, atol=1e-6)  # float atomic sums per population
    assert np.array_equal(a.V_i.numpy(), b.V_i.numpy())
    assert np.array_equal(a.J_ui.numpy(), b.J_ui.numpy())
    assert float(a.t) == float(b.t)
