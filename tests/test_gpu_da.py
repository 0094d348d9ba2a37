"""GPU tests of the in-process BlockWrapper / simulation / DataAssimilation
layer against the oracle: reference npz loading, hyper-parameter push
semantics, the GPU voxel merge + BOLD, and one EnKF window."""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


def _props(n, rng, sub):
    p = np.zeros((n, 22), np.float32)
    p[:, 0] = rng.uniform(0.002, 0.01, size=n)
    p[:, 3] = sub
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32) * 3
    p[:, 10:14] = g * rng.uniform(0.8, 1.2, size=(n, 4))
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


def _write_blocks(tmp, rng):
    """Two reference-format block files (generation/make_block.py:473-479)."""
    sizes = [300, 200]
    bases = [0, 300]
    for i, n in enumerate(sizes):
        sub = np.where(np.arange(n) < int(0.8 * n), 2 * i, 2 * i + 1)
        K = 20
        out_idx = np.repeat(np.arange(n, dtype=np.uint32), K)
        blk = rng.integers(0, 2, size=n * K).astype(np.int16) - i  # relative block index
        src_sizes = np.array(sizes)[blk + i]
        in_idx = (rng.random(n * K) * src_sizes).astype(np.uint32)
        chan = np.where(in_idx < (0.8 * src_sizes), 0, 2).astype(np.uint8)
        w = rng.random((n * K, 2)).astype(np.float32)
        np.savez(os.path.join(tmp, f"block_{i}.npz"), property=_props(n, rng, sub), output_neuron_idx=out_idx,
                 input_block_idx=blk, input_neuron_idx=in_idx, input_channel_offset=chan, weight=w)
    return bases


def test_local_wrapper_loads_reference_npz_and_runs_like_oracle(tmp_path):
    from digital_twin_brain_amd.wrapper import LocalBlockWrapper

    rng = np.random.default_rng(0)
    bases = _write_blocks(str(tmp_path), rng)
    bw = LocalBlockWrapper("local", str(tmp_path), 0.1, noise_seed=6)
    assert int(bw.total_neurons) == 500 and bw.total_subblks == 4
    assert bw.neurons_per_subblk.cpu().tolist() == [240, 60, 160, 40]
    # CSR matches the reference reader's (pre, post, channel) triples
    nb = bw.nb
    rp = nb.get("rowptr")
    pre = np.repeat(np.arange(500), np.diff(rp))
    got = set(zip(pre.tolist(), nb.get("post").tolist(), nb.get("kind").tolist(), nb.get("w0").tolist()))
    want = set()
    for i in range(2):
        z = np.load(os.path.join(str(tmp_path), f"block_{i}.npz"))
        gp = (np.array(bases)[z["input_block_idx"] + i] + z["input_neuron_idx"]).tolist()
        gq = (bases[i] + z["output_neuron_idx"].astype(np.int64)).tolist()
        want |= set(zip(gp, gq, (z["input_channel_offset"] // 2).tolist(), z["weight"][:, 0].tolist()))
    assert got == want
    csr = dict(rowptr=rp, post=nb.get("post"), kind=nb.get("kind"), w0=nb.get("w0"), w1=nb.get("w1"))
    ob = orc.OracleBlock(nb.get("props"), csr, 0.1)
    steps = 200
    f = [x[0].cpu().numpy() for x in bw.run(steps, freqs=True)]
    for s in range(steps):
        sa, _ = ob.run(noise_mode=1, seed=6)
        assert np.array_equal(f[s], sa.astype(np.int32)), f"step {s + 1}"
    assert sum(x.sum() for x in f) > 0


def test_mul_property_by_subblk_is_relative_to_the_original():
    """cuda/python/dist_blockwrapper_pytorch.py:446-458: without `accumulate`
    the property ends at original * hp; with it, factors compound."""
    from digital_twin_brain_amd import NativeBlock
    from digital_twin_brain_amd.wrapper import LocalBlockWrapper

    rng = np.random.default_rng(1)
    n = 400
    p = _props(n, rng, np.arange(n) // 100)
    nb = NativeBlock.from_csr(p, np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint8),
                              np.zeros(0, np.float32), stats_exclude_blocked=True)
    bw = LocalBlockWrapper(network=nb)
    idx = torch.tensor([[1, 10], [3, 13]], dtype=torch.int64)
    bw.mul_property_by_subblk(idx, torch.tensor([2.0, 0.5]))
    bw.mul_property_by_subblk(idx, torch.tensor([3.0, 0.25]))
    got = nb.get("props")
    np.testing.assert_allclose(got[100:200, 10], p[100:200, 10] * 3.0, rtol=1e-6)
    np.testing.assert_allclose(got[300:400, 13], p[300:400, 13] * 0.25, rtol=1e-6)
    bw.mul_property_by_subblk(idx, torch.tensor([2.0, 2.0]), accumulate=True)
    got = nb.get("props")
    np.testing.assert_allclose(got[100:200, 10], p[100:200, 10] * 6.0, rtol=1e-6)
    assert np.array_equal(got[:100], p[:100])


def test_simulation_evolve_voxel_bold_matches_oracle():
    """simulation.evolve: per-step voxel merge (np.histogram over population ids)
    and BOLD, run on the GPU, against the oracle BOLD on the same counts."""
    from digital_twin_brain_amd import NativeBlock
    from digital_twin_brain_amd.simulation import simulation

    rng = np.random.default_rng(2)
    n = 3000
    sub = np.arange(n) // 250  # 12 populations -> 6 voxels of 2
    p = _props(n, rng, sub)
    csr = orc.gen_fixed_indegree(n, 2400, 60, 1, 0.0, 1.0, False, seed=3)
    nb = NativeBlock.from_csr(p, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], dt=1.0, noise_seed=2,
                              stats_exclude_blocked=True)
    sim = simulation(network=nb, column=False, dt=1.0, write_path="/tmp")
    assert sim.num_voxels == 6
    out = sim.evolve(150)
    freqs = out[0].cpu().numpy()
    ob = orc.OracleBOLD(6)
    sizes = np.full(6, 500.0)
    for s in range(150):
        act = np.histogram(np.arange(12), weights=freqs[s], bins=6, range=(0, 12))[0] / sizes
        want = ob.run(np.maximum(act, 1e-5))
    np.testing.assert_allclose(out[-1].cpu().numpy(), want, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(sim.bold.q.cpu().numpy(), ob.q, rtol=1e-10)
    assert freqs.sum() > 0


def test_da_window_analysis_matches_oracle_enkf():
    """One DataAssimilation window: hidden state layout, the FP64 EnKF against
    the oracle with the same noise, and the BOLD state written back."""
    from digital_twin_brain_amd import networks
    from digital_twin_brain_amd.data_assimilation import DataAssimilation

    E, V = 4, 3
    mem = networks.ensemble_members(E=E, seed=1, n_lo=400, n_hi=1500, k_lo=30, k_hi=60, truth_n=600, truth_k=40)
    props = networks.ensemble_props(mem, voxels=V)
    nb = networks.NativeBlock.members(props, mem, max_delay=2, stats_exclude_blocked=True)
    da = DataAssimilation(network=nb, column=False, ensemble=E, dt=0.1, bold_steps=10, da_populations=E * 2 * V,
                          da_voxels=E * V, write_path="/tmp")
    low, high = torch.full((2 * V, 2), 0.5), torch.full((2 * V, 2), 2.0)
    hp = da.hp_random_initialize(low, high, rng=np.random.default_rng(0))
    da.hp_index2hidden_state()
    da.da_property_initialize([10, 12], None, hp)
    da.da_evolve(200)
    hs = da._hidden_state.cpu().numpy()
    assert hs.shape == (E, V, 2 * 2 + 6)
    np.testing.assert_allclose(hs[:, :, :4].reshape(E, -1), da._hp_log.numpy().reshape(E, -1), rtol=1e-6)
    xi = np.random.default_rng(3).normal(size=(E, V))
    y = np.full((1, V), 0.02)
    da.da_filter(torch.from_numpy(y), bold_sigma=1e-6, solo_rate=0.8, xi=torch.from_numpy(xi))
    want = orc.diffusion_enkf(hs, 1e-6, y, 0.8, xi)
    np.testing.assert_allclose(da._hidden_state.cpu().numpy(), want, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(da.bold.q.cpu().numpy()[: E * V], np.maximum(want[:, :, -4], 1e-5).reshape(-1),
                               rtol=1e-12)


def test_da_rest_run_resumes_from_a_checkpoint(tmp_path):
    """A DA run stopped after 2 of 4 windows and resumed from its checkpoint
    by a freshly built block and DataAssimilation (RNG streams scrambled)
    gives the uninterrupted run's hidden states, before and after every
    analysis, bit-for-bit: bucketed delivery with fixed-point sums makes the
    windows independent of atomic order."""
    from digital_twin_brain_amd import networks
    from digital_twin_brain_amd.data_assimilation import DataAssimilation

    E, V = 4, 3

    def make():
        mem = networks.ensemble_members(E=E, seed=1, n_lo=400, n_hi=1500, k_lo=30, k_hi=60, truth_n=600, truth_k=40)
        props = networks.ensemble_props(mem, voxels=V)
        nb = networks.NativeBlock.members(props, mem, max_delay=3, stats_exclude_blocked=True)
        nb.set_launch(deliver_mode=1, deliver_grid=256)
        assert nb.get_launch()["accum_fixed"] >= 16
        da = DataAssimilation(network=nb, column=False, ensemble=E, dt=0.1, bold_steps=10, da_populations=E * 2 * V,
                              da_voxels=E * V, write_path=str(tmp_path))
        low, high = torch.full((2 * V, 2), 0.5), torch.full((2 * V, 2), 2.0)
        hp = da.hp_random_initialize(low, high, rng=np.random.default_rng(0))
        da.hp_index2hidden_state()
        da.da_property_initialize([10, 12], None, hp)
        return da, nb

    y = np.random.default_rng(5).uniform(0.01, 0.03, size=(4, V))
    kw = dict(steps=200, bold_sigma=1e-6, hp_sigma=0.25, save=False)
    da, nb = make()
    torch.manual_seed(11)
    w_all, f_all, _ = da.da_rest_run(y, str(tmp_path), **kw)
    nb.close()
    assert not np.array_equal(w_all[0], w_all[3])

    ck = str(tmp_path / "da_checkpoint.npz")
    da, nb = make()
    torch.manual_seed(11)
    w2, f2, _ = da.da_rest_run(y[:2], str(tmp_path), checkpoint=ck, **kw)
    nb.close()
    assert np.array_equal(w2, w_all[:2]) and np.array_equal(f2, f_all[:2])

    da, nb = make()
    torch.manual_seed(12345)
    w3, f3, _ = da.da_rest_run(y, str(tmp_path), checkpoint=ck, **kw)
    assert np.array_equal(w3, w_all) and np.array_equal(f3, f_all)
    assert nb.steps == 4 * 200
    nb.close()


def test_ensemble_da_runner_small():
    from digital_twin_brain_amd.da_runner import run_ensemble_da

    out, da, rec = run_ensemble_da(E=6, voxels=4, windows=3, window_ms=50.0, n_lo=500, n_hi=3000, k_lo=40, k_hi=80,
                          truth_n=1000, truth_k=60)
    assert out["windows"] == 3 and np.isfinite(out["window_s_mean"])
    assert rec["w_save"].shape == (3, 6, 4, 2 * 2 + 6)
    assert np.all(np.isfinite(rec["obs"]))


@pytest.mark.parametrize("adapt", [0.0, 1.5])
def test_surrogate_da_small(adapt):
    """Surrogate-assisted DA: fallback every 5th window, and after any window
    over the NRMSE limit (with adapt > 0, over adapt x the median of the last 5
    simulated windows too); finite forecasts, surrogate windows faster than
    simulated ones."""
    from digital_twin_brain_amd.da_runner import run_surrogate_da

    out, rec = run_surrogate_da(E=8, voxels=4, train_windows=12, windows=7, window_ms=50.0, hidden=(128, 128), epochs=60,
                      finetune_epochs=5, n_lo=500, n_hi=3000, k_lo=40, k_hi=80, truth_n=1000, truth_k=60, adapt=adapt)
    sim = rec["sur"]["simulated"]
    assert sim[0] and sim[5]
    assert out["simulated_windows"] == int(sim.sum()) and len(out["nrmse"]) == out["simulated_windows"]
    errs = rec["sur"]["nrmse"]
    for i, (t, e) in enumerate(errs):
        floor = float(np.median([x for _, x in errs[max(0, i - 5):i]])) if (adapt > 0 and i) else 0.0
        force = e > max(0.02, adapt * floor)
        if t + 1 < 7:
            assert bool(sim[t + 1]) == (force or (t + 1) % 5 == 0)
    assert np.all(np.isfinite(rec["sur"]["forecasts"]))
    times = rec["sur"]["times"]
    if (~sim).any():
        assert times[~sim].mean() < times[sim].mean()
    assert out["fit_loss_last"] < out["fit_loss_first"]
    # both runs start from the same snapshot: the first window is simulated in
    # both, from the same state and hyper-parameters (float atomics may round
    # differently, so not bit for bit)
    b_sim, b_sur = rec["w_sim"][0][..., -1], rec["sur"]["forecasts"][0][..., -1]
    assert np.linalg.norm(b_sur - b_sim) <= 0.05 * np.linalg.norm(b_sim) + 1e-9
    # the assisted run tracks the truth about as well as the simulator-only run
    assert out["bold_rmse_sur_mean"] <= 2.0 * out["bold_rmse_sim_mean"] + 1e-3, out


def test_bold_run_equals_the_oracle_stepped_with_a_constant_input():
    """dtb_bold_run_f64: K Balloon-Windkessel steps with a constant input per
    voxel in one launch, against the FP64 oracle stepped K times."""
    from digital_twin_brain_amd import _native as nat

    rng = np.random.default_rng(7)
    n, K = 1000, 800
    s0, q0, v0, f0 = (rng.uniform(-0.05, 0.05, n), rng.uniform(0.8, 1.1, n), rng.uniform(0.9, 1.1, n),
                      rng.uniform(0.9, 1.2, n))
    u = rng.uniform(1e-5, 0.02, n)
    ob = orc.OracleBOLD(n, s=s0, q=q0, v=v0, f_in=f0)
    for _ in range(K):
        want = ob.run(u)
    dev = torch.device("cuda")
    st = [torch.from_numpy(x.copy()).to(dev) for x in (s0, q0, v0, f0)]
    ud = torch.from_numpy(u).to(dev)
    out = torch.zeros(n, dtype=torch.float64, device=dev)
    par = torch.from_numpy(ob.par.copy()).to(dev)
    nat.check(nat.lib().dtb_bold_run_f64(n, K, par.data_ptr(), st[0].data_ptr(), st[1].data_ptr(), st[2].data_ptr(),
                                          st[3].data_ptr(), ud.data_ptr(), out.data_ptr(), None))
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-11, atol=1e-14)
    for got, w in zip(st, (ob.s, ob.q, ob.v, ob.f_in)):
        np.testing.assert_allclose(got.cpu().numpy(), w, rtol=1e-11, atol=1e-14)


def test_window_input_is_the_mean_bold_input_of_the_window():
    """simulation.evolve(..., window_input=True): the window's mean BOLD input
    per voxel equals the mean over its BOLD steps of (the voxel's spikes in
    the step's 10 network steps) / (its neurons), from the per-step counts of
    the same window replayed from a snapshot."""
    from digital_twin_brain_amd import networks
    from digital_twin_brain_amd.data_assimilation import DataAssimilation

    E, V = 3, 4
    mem = networks.ensemble_members(E=E, seed=2, n_lo=400, n_hi=1200, k_lo=30, k_hi=60, truth_n=500, truth_k=40)
    props = networks.ensemble_props(mem, voxels=V)
    nb = networks.NativeBlock.members(props, mem, max_delay=2, stats_exclude_blocked=True)
    nb.set_launch(deliver_mode=1, accum_fixed=1)  # integer sums: the replay is exact
    da = DataAssimilation(network=nb, column=False, ensemble=E, dt=0.1, bold_steps=10, da_populations=E * 2 * V,
                          da_voxels=E * V, write_path="/tmp")
    da.evolve(200, bold_detail=True)
    snap = nb.save_state()
    steps = 2300  # more than one 1000-step batch, not a multiple of it
    da.evolve(steps, bold_detail=True, window_input=True)
    got = da.last_window_input.cpu().numpy()
    nb.load_state(snap)
    f = da.evolve(steps)[0].cpu().numpy().astype(np.int64)  # [steps, populations]
    pops = da.population_id_cpu
    sizes = da.num_neurons_per_voxel_cpu.astype(np.float64)
    want = np.zeros(da.num_voxels)
    for j, p in enumerate(pops):
        v = da._vop[p]
        if v >= 0:
            want[v] += f[:, j].sum()
    want = want / sizes / (steps // 10)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert want.sum() > 0


@pytest.mark.parametrize("surrogate", ["window", "rate"])
def test_region_surrogate_da_small(surrogate):
    """Surrogate-assisted DA on whole-brain members (4 g_ui per region, a
    judge simulating copies of the first members every 5th window, full
    simulation after a judged window over the NRMSE threshold) against
    simulator-only DA from the same state, on a small configuration: the
    judge's floor (simulation against simulation) is measured; the
    surrogate's error on that window is a few percent; surrogate windows cost
    a fraction of simulated ones; the fallback pattern follows the policy.
    The full configuration (256 members of 10,000 neurons, 50,176 training
    windows, 800 ms windows) runs in bench.py."""
    from digital_twin_brain_amd.region_da import run_region_surrogate_da

    out, rec = run_region_surrogate_da(E=32, R=20, member_n=4000, train_windows=32, windows=10, window_ms=200.0,
                                       subset=4, every=5, seed=3, surrogate=surrogate)
    cal = out["calibration"]
    # 200 ms windows: the BOLD state at the window end is mostly the start state's, so two
    # simulations agree closely; the surrogate (1,024 training windows) is off by a few percent
    assert 0 < cal["sim_vs_sim"] < 0.2
    assert cal["sur_vs_full"] < 0.05, cal
    assert out["nrmse_threshold"] >= 0.02
    kinds = out["kinds"]
    assert kinds[0] == "judge"
    for t, k in enumerate(kinds):
        prev = [(kk, e) for (tt, kk, e) in out["nrmse"] if tt == t - 1]
        if prev and prev[0][0] == "judge" and prev[0][1] > out["nrmse_threshold"]:
            assert k == "full", (t, kinds)  # a judged window over the threshold: the simulator next
        else:
            assert k == ("judge" if t % 5 == 0 else "surrogate"), (t, kinds)
    if "surrogate" in out["window_s_by_kind"]:
        assert out["window_s_by_kind"]["surrogate"] < 0.5 * out["sim_window_s_mean"]
    assert out["train_samples"] == 32 * 32
    rec["da"].block_model.nb.close()
    rec["judge"].nb.close()
