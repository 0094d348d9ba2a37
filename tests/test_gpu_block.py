"""GPU parity of the block step (libdtb through the C ABI) against the
reference golden vectors and the CPU oracle.

Tolerances (FP32 GPU state against an FP64 reference):
- threshold/delivered rasters identical until the first threshold-crossing
  divergence, which is reported; the first divergence must come late;
- V within 1e-5 relative and J_ui within 1e-5 of max|J| before it;
- CSR indexing, spike compaction and spike counts bit-exact.
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


def rows_by_delay(csr):
    """Stable per-row reorder by delay: the library's row layout."""
    rp = csr["rowptr"]
    n = rp.shape[0] - 1
    row = np.repeat(np.arange(n), np.diff(rp))
    order = np.lexsort((np.arange(row.shape[0]), csr["delay"], row))
    out = dict(csr)
    for k in ("post", "kind", "delay", "w0", "w1"):
        if out.get(k) is not None:
            out[k] = out[k][order]
    return out


def arrivals_per_step(csr, spikes_by_step, D):
    """Synaptic events arriving at each step: sum over d of the delay-d
    segments of the rows that spiked at step - d + 1."""
    rp = csr["rowptr"]
    n = rp.shape[0] - 1
    row = np.repeat(np.arange(n), np.diff(rp))
    cnt = np.zeros((n, D + 1), np.int64)
    np.add.at(cnt, (row, csr["delay"].astype(np.int64)), 1)
    out = []
    for t in range(len(spikes_by_step)):
        e = 0
        for d in range(1, D + 1):
            if t - d + 1 >= 0:
                e += int(cnt[spikes_by_step[t - d + 1].astype(bool), d].sum())
        out.append(e)
    return np.array(out, np.int64)


def _native_from_case(c, seed):
    from digital_twin_brain_amd import NativeBlock

    w1 = c["w1"] if not np.array_equal(c["w0"], c["w1"]) else None
    return NativeBlock.from_csr(c["props"], c["rowptr"], c["post"], c["kind"], c["w0"], w1, dt=float(c["dt"]),
                                noise_seed=seed)


@pytest.mark.parametrize("name,min_agree", [("dense64", 150), ("isolated64", 200), ("sparse400", 500)])
def test_block_matches_reference_golden(golden_dir, name, min_agree):
    c = _load(golden_dir, name)
    n, steps, keep = int(c["n"]), int(c["steps"]), int(c["keep_every"])
    nb = _native_from_case(c, int(c["seed"]))
    act = np.unpackbits(c["active"], axis=1)[:, :n]
    spk = np.unpackbits(c["spike"], axis=1)[:, :n]
    snaps = set(int(s) for s in c["snap_steps"])
    freqs = np.zeros(nb.S, np.uint32)
    vmeans = np.zeros(nb.S, np.float32)
    diverged = None
    kv = 0
    for s in range(1, steps + 1):
        nb.run(1, noise_mode=0, noise_rate=float(c["noise_rate"]), isolated=bool(c["isolated"]), graph=False,
               freqs=freqs, vmeans=vmeans)
        a = nb.get("active")
        if not np.array_equal(a, act[s - 1]):
            diverged = s
            break
        assert np.array_equal(nb.get("spike"), spk[s - 1]), f"delivered raster differs at step {s}"
        assert np.array_equal(freqs.astype(np.float64), c["sum_act"][s - 1])
        np.testing.assert_allclose(vmeans, c["mean_v"][s - 1], rtol=RTOL)
        if s % keep == 0:
            np.testing.assert_allclose(nb.get("V"), c["V"][kv], rtol=RTOL, atol=0)
            kv += 1
        if s in snaps:
            Jg = c[f"J_{s}"]
            np.testing.assert_allclose(nb.get("J"), Jg, rtol=RTOL, atol=RTOL * (np.abs(Jg).max() + 1e-30))
            Ig = c[f"I_ui_{s}"]
            np.testing.assert_allclose(nb.get("I_ui"), Ig, rtol=RTOL, atol=RTOL * (np.abs(Ig).max() + 1e-30))
            Isg = c[f"I_syn_{s}"]
            np.testing.assert_allclose(nb.get("I_syn"), Isg, rtol=RTOL, atol=RTOL * (np.abs(Isg).max() + 1e-30))
    print(f"{name}: first threshold-raster divergence at step {diverged} of {steps}")
    assert diverged is None or diverged > min_agree, f"raster diverged at step {diverged}"
    assert abs(nb.t - float(c["t_final"])) == 0.0 or diverged is not None


def test_dropin_block_class_matches_golden(golden_dir):
    """digital_twin_brain_amd.block with the reference constructor and run()."""
    from digital_twin_brain_amd import block
    from tests.golden.make_golden import csr_to_torch_sparse

    c = _load(golden_dir, "dense64")
    n = int(c["n"])
    W = csr_to_torch_sparse(n, dict(rowptr=c["rowptr"], post=c["post"], kind=c["kind"], w0=c["w0"], w1=c["w1"]))
    B = block(node_property=torch.from_numpy(c["props"]), w_uij=W.to_dense(), delta_t=float(c["dt"]),
              seed=int(c["seed"]))
    act = np.unpackbits(c["active"], axis=1)[:, :n]
    for s in range(1, 101):
        sa, mv = B.run(noise_rate=float(c["noise_rate"]), isolated=False)
        assert np.array_equal(B.active.numpy().astype(np.uint8), act[s - 1])
        assert np.array_equal(sa.numpy().astype(np.float64), c["sum_act"][s - 1])
        np.testing.assert_allclose(B.V_i.numpy(), c["V"][s - 1], rtol=RTOL)
    J = B.J_ui.numpy()
    assert J.shape == (4, n) and B.I_ui.shape == (4, n) and B.I_syn.shape == (n,)
    assert B.sub_idx.shape == (n,) and float(B.t) == pytest.approx(100 * float(c["dt"]))


@pytest.mark.parametrize("pair,D", [(False, 1), (True, 1), (False, 5), (True, 7)])
def test_gpu_generator_csr_bit_exact(pair, D):
    from digital_twin_brain_amd import NativeBlock

    n, ne, K = 3000, 2400, 60
    props = np.zeros((n, 22), np.float32)
    props[:, 4] = 1; props[:, 5] = 2; props[:, 6] = 0.03; props[:, 7] = -75; props[:, 8] = -50; props[:, 9] = -65
    props[:, 18:22] = (2, 40, 10, 50)
    nb = NativeBlock.fixed_indegree(props, ne, K, max_delay=D, w_lo=0.0, w_hi=1.0, pair_weights=pair, conn_seed=77)
    ref = rows_by_delay(orc.gen_fixed_indegree(n, ne, K, D, 0.0, 1.0, pair, seed=77))
    assert np.array_equal(nb.get("rowptr"), ref["rowptr"])
    assert np.array_equal(nb.get("post"), ref["post"])
    assert np.array_equal(nb.get("kind"), ref["kind"])
    assert np.array_equal(nb.get("delay"), ref["delay"])
    assert np.array_equal(nb.get("w0"), ref["w0"])
    assert np.array_equal(nb.get("w1"), ref["w1"] if pair else ref["w0"])


def _plain_like_props(n, n_exc, rng, g_scale=3.0):
    p = np.zeros((n, 22), np.float32)
    p[:, 3] = (np.arange(n) >= n_exc).astype(np.float32)
    p[:, 4] = 1; p[:, 5] = 2; p[:, 6] = 0.03; p[:, 7] = -75; p[:, 8] = -50; p[:, 9] = -65
    g_cal = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], np.float32)
    p[:, 10:14] = g_cal * g_scale * rng.uniform(0.8, 1.2, size=(n, 4)).astype(np.float32)
    p[:, 14:18] = (0, 0, -70, -100)
    p[:, 18:22] = (2, 40, 10, 50)
    return p


def test_compaction_and_delivery_bit_exact_on_fixed_spike_vector():
    """Isolated step with injected uniforms: the spiking set is fixed, so the
    compacted list and (dyadic weights) the delivered J are exact."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K = 5000, 4000, 40
    rng = np.random.default_rng(5)
    ref = orc.gen_fixed_indegree(n, ne, K, 1, 0.0, 1.0, True, seed=12)
    for k in ("w0", "w1"):
        ref[k] = (np.floor(ref[k] * 256) / 256).astype(np.float32)
    props = _plain_like_props(n, ne, rng)
    nb = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], ref["w1"], dt=0.1)
    ob = orc.OracleBlock(props, ref, 0.1)
    for s in range(1, 6):
        u = np.ones(n, np.float32)
        chosen = rng.choice(n, size=50 + 100 * s, replace=False)
        u[chosen] = 0.0
        nb.run(1, noise_mode=2, noise_rate=0.5, uniforms=u, isolated=True, graph=False)
        ob.run(noise_rate=0.5, uniforms=u, isolated=True)
        got = nb.get("spike_list")
        assert np.array_equal(got, orc.compact(ob.get("spike")))
        assert np.array_equal(np.sort(chosen).astype(np.int32), got)
    # same-step delivery of dyadic weights: the ring sum is exact; J differs
    # from the FP64 oracle only by the FP32 decay of earlier steps
    Jg, Jo = nb.get("J"), ob.get("J")
    np.testing.assert_allclose(Jg, Jo, rtol=1e-6, atol=1e-6 * np.abs(Jo).max())


@pytest.mark.parametrize("pair,D,graph,mode", [(False, 1, True, None), (True, 1, False, None), (False, 6, True, None),
                                              (True, 4, True, None), (False, 63, True, 1), (True, 63, False, 0)])
def test_network_run_matches_oracle(pair, D, graph, mode):
    """Plain-block-like recurrent network (fixed in-degree, delays 1..D, up to
    63, the generator's limit) against the FP64 push oracle on the same CSR and
    the same Philox noise; mode forces direct (0) or bucketed (1) delivery."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K, steps = 4000, 3200, 100, 600
    rng = np.random.default_rng(11)
    props = _plain_like_props(n, ne, rng)
    props[:, 3] = np.arange(n) // 500  # 8 sub-populations
    nb = NativeBlock.fixed_indegree(props, ne, K, max_delay=D, pair_weights=pair, conn_seed=3, noise_seed=9)
    if mode is not None:
        nb.set_launch(deliver_mode=mode, deliver_grid=64)
    csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
               w0=nb.get("w0"), w1=nb.get("w1") if pair else None)
    ob = orc.OracleBlock(props, csr, 0.1)
    rate = 2.0 ** -7
    B = 50
    freqs = np.zeros((B, nb.S), np.uint32)
    vmeans = np.zeros((B, nb.S), np.float32)
    events = np.zeros(B, np.uint64)
    nspk = np.zeros(B, np.uint32)
    diverged = None
    total_spikes = 0
    ospk = []
    for b0 in range(0, steps, B):
        nb.run(B, noise_mode=0, noise_rate=rate, graph=graph, freqs=freqs, vmeans=vmeans, events=events,
               spikes=nspk)
        oa = []
        for k in range(B):
            sa, mv = ob.run(noise_rate=rate, seed=9)
            oa.append((sa, mv))
            ospk.append(ob.get("spike"))
        # per-step records of the batch against the oracle (bit-exact counts)
        for k in range(B):
            if diverged is None and not np.array_equal(freqs[k].astype(np.float64), oa[k][0]):
                diverged = b0 + k + 1
        if diverged is not None:
            break
        for k in range(B):
            np.testing.assert_allclose(vmeans[k], oa[k][1], rtol=RTOL)
        np.testing.assert_allclose(nb.get("V"), ob.get("V"), rtol=RTOL)
        Jo = ob.get("J")
        np.testing.assert_allclose(nb.get("J"), Jo, rtol=RTOL, atol=RTOL * np.abs(Jo).max())
        sl = nb.get("spike_list")
        assert np.array_equal(sl, orc.compact(ob.get("spike")))
        # events = synapses of the (row, delay) segments arriving this step
        arr = arrivals_per_step(csr, ospk, D)[b0:b0 + B]
        assert np.array_equal(events.astype(np.int64), arr)
        assert int(nspk[-1]) == sl.shape[0]
        assert np.array_equal(nspk.astype(np.int64), np.array([x.sum() for x in ospk[b0:b0 + B]]))
        total_spikes += int(nspk.sum())
    print(f"pair={pair} D={D}: first divergence {diverged}, delivered spikes {total_spikes}")
    assert total_spikes > 0
    assert diverged is None or diverged > 300


def test_graph_and_direct_launch_agree_bitwise():
    from digital_twin_brain_amd import NativeBlock

    n, ne, K = 3000, 2400, 50
    rng = np.random.default_rng(2)
    props = _plain_like_props(n, ne, rng)
    ref = orc.gen_fixed_indegree(n, ne, K, 3, 0.0, 1.0, False, seed=4)
    ref["w0"] = (np.floor(ref["w0"] * 256) / 256).astype(np.float32)
    a = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], None, ref["delay"], dt=0.1,
                             noise_seed=1)
    b = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], None, ref["delay"], dt=0.1,
                             noise_seed=1)
    fa = np.zeros((400, a.S), np.uint32)
    fb = np.zeros((400, b.S), np.uint32)
    a.run(400, noise_rate=0.01, graph=True, freqs=fa)
    for k in range(4):
        b.run(100, noise_rate=0.01, graph=False, freqs=fb[k * 100:(k + 1) * 100])
    assert np.array_equal(fa, fb)
    assert np.array_equal(a.get("V"), b.get("V"))
    assert np.array_equal(a.get("J"), b.get("J"))
    assert fa.sum() > 0


def test_property_edits_match_reference_semantics():
    from digital_twin_brain_amd import NativeBlock

    n = 1000
    rng = np.random.default_rng(3)
    props = _plain_like_props(n, 800, rng)
    props[:, 3] = np.arange(n) // 250
    csr = orc.gen_fixed_indegree(n, 800, 20, 1, 0.0, 1.0, False, seed=5)
    nb = NativeBlock.from_csr(props, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], None, dt=0.1)
    # mul_property_by_subblk: column 10 (g_AMPA) of sub-population 2 times 1.5, column 4 (C) of 0 times 2
    nb.mul_props_by_pop([2, 0], [10, 4], [1.5, 2.0])
    want = props.copy()
    want[500:750, 10] *= np.float32(1.5)
    want[0:250, 4] *= np.float32(2.0)
    assert np.array_equal(nb.get("props"), want)
    nb.set_props([7, 999], [9, 0], [-60.0, 0.25])
    want[7, 9] = -60.0
    want[999, 0] = 0.25
    assert np.array_equal(nb.get("props"), want)
    # gamma(alpha, beta) draws: mean alpha/beta, variance alpha/beta^2
    big = 200000
    p2 = np.tile(props[:1], (big, 1))
    p2[:, 3] = np.arange(big) // (big // 2)
    c2 = dict(rowptr=np.zeros(big + 1, np.int64), post=np.zeros(0, np.int32), kind=np.zeros(0, np.uint8),
              w0=np.zeros(0, np.float32))
    nb2 = NativeBlock.from_csr(p2, c2["rowptr"], c2["post"], c2["kind"], c2["w0"], None, dt=0.1)
    nb2.gamma_props_by_pop([0, 1], [10, 12], [5.0, 0.5], [5.0, 2.0], seed=17)
    g = nb2.get("props")
    a0 = g[: big // 2, 10]
    a1 = g[big // 2:, 12]
    assert abs(a0.mean() - 1.0) < 0.01 and abs(a0.var() - 0.2) < 0.01
    assert abs(a1.mean() - 0.25) < 0.005 and abs(a1.var() - 0.125) < 0.006
    assert np.array_equal(g[big // 2:, 10], p2[big // 2:, 10])  # untouched


def test_gamma_draws_follow_the_gamma_distribution():
    """The whole distribution of the per-neuron Gamma(alpha, beta) property
    draws (cuda/src/stage/property.cpp:312-345, Marsaglia-Tsang), not only its
    moments: a Kolmogorov-Smirnov test against scipy's Gamma CDF with 100,000
    draws per case, which resolves a CDF error of about 0.006.  Covers the
    alpha < 1 boost (alpha = 0.3, 0.5), alpha = 1 and a large alpha."""
    from scipy import stats

    from digital_twin_brain_amd import NativeBlock

    cases = [(0.3, 1.0), (0.5, 2.0), (1.0, 0.5), (5.0, 5.0), (40.0, 8.0)]
    per = 100000
    big = per * len(cases)
    rng = np.random.default_rng(1)
    p2 = np.tile(_plain_like_props(1, 1, rng), (big, 1))
    p2[:, 3] = np.arange(big) // per
    nb = NativeBlock.from_csr(p2, np.zeros(big + 1, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint8),
                              np.zeros(0, np.float32), None, dt=0.1)
    try:
        nb.gamma_props_by_pop(list(range(len(cases))), [11] * len(cases), [a for a, _ in cases],
                              [b for _, b in cases], seed=23)
        g = nb.get("props")[:, 11].astype(np.float64)
    finally:
        nb.close()
    for k, (a, b) in enumerate(cases):
        x = g[k * per:(k + 1) * per]
        assert (x > 0).all() and np.isfinite(x).all(), (a, b)
        res = stats.kstest(x, "gamma", args=(a, 0.0, 1.0 / b))
        assert res.pvalue > 1e-3, (a, b, res.statistic, res.pvalue)


def test_bold_kernel_matches_reference_golden(golden_dir):
    from digital_twin_brain_amd import BOLD

    z = np.load(os.path.join(golden_dir, "bold_v18.npz"))
    for dtype, tol in ((torch.float64, 1e-12), (torch.float32, 2e-5)):
        b = BOLD(epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02)
        k = 0
        for s in range(z["u"].shape[0]):
            out = b.run(torch.from_numpy(z["u"][s]).to(dtype).cuda())
            np.testing.assert_allclose(out.cpu().numpy(), z["out"][s], rtol=tol, atol=tol * 0.05)
            if (s + 1) % 10 == 0:
                st = torch.stack([b.s, b.q, b.v, b.f_in]).cpu().numpy()
                np.testing.assert_allclose(st, z["state"][k], rtol=tol, atol=tol * 0.05)
                k += 1
        b.state_update(torch.from_numpy(z["su_w"]).to(dtype).cuda())
        st = torch.stack([b.s, b.q, b.v, b.f_in]).cpu().numpy()
        np.testing.assert_allclose(st, z["su_state"], rtol=tol, atol=1e-12)
        nv3 = z["su_w"].shape[0] * z["su_w"].shape[1]
        out = b.run(torch.from_numpy(z["u"][0][:nv3]).to(dtype).cuda())
        np.testing.assert_allclose(out.cpu().numpy(), z["su_out"], rtol=tol, atol=tol * 0.05)


def test_enkf_kernel_matches_oracle():
    from digital_twin_brain_amd import diffusion_enkf

    rng = np.random.default_rng(8)
    for E, V, S in ((64, 90, 10), (256, 90, 10), (7, 3, 2)):
        w = rng.normal(size=(E, V, S)) * 0.1
        y = rng.normal(size=V) * 0.05
        xi = rng.normal(size=(E, V))
        for rho in (0.0, 0.8, 1.0):
            got = diffusion_enkf(torch.from_numpy(w).cuda(), 1e-4, torch.from_numpy(y).reshape(1, -1), rho,
                                 xi=torch.from_numpy(xi))
            want = orc.diffusion_enkf(w, 1e-4, y, rho, xi)
            np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-10, atol=1e-12)


def test_invalid_inputs_raise_cleanly():
    from digital_twin_brain_amd import NativeBlock
    from digital_twin_brain_amd._native import DTBError

    props = _plain_like_props(10, 8, np.random.default_rng(0))
    with pytest.raises(DTBError):
        NativeBlock.from_csr(props, np.array([0] * 10 + [1], np.int64), np.array([10], np.int32),
                             np.array([0], np.uint8), np.array([1.0], np.float32))
    with pytest.raises(DTBError):
        NativeBlock.from_csr(props, np.array([0] * 10 + [1], np.int64), np.array([3], np.int32),
                             np.array([2], np.uint8), np.array([1.0], np.float32))
    nb = NativeBlock.from_csr(props, np.zeros(11, np.int64), np.zeros(0, np.int32), np.zeros(0, np.uint8),
                              np.zeros(0, np.float32))
    with pytest.raises(DTBError) as e:
        nb.run(3, sample_v=np.zeros(3, np.float32))
    assert e.value.code == 2  # UNSET_SAMPLE_FAILURE
    with pytest.raises(DTBError):
        nb.mul_props_by_pop([5], [10], [2.0])  # sub-population out of range


def test_attached_bold_matches_oracle_on_same_counts():
    """Per-voxel BOLD fed by the block's own spike counts, every 10 steps
    (1 ms at dt 0.1), against the oracle's FP64 Balloon-Windkessel step."""
    from digital_twin_brain_amd import NativeBlock

    n, ne = 6000, 4800
    rng = np.random.default_rng(4)
    props = _plain_like_props(n, ne, rng)
    props[:, 3] = np.arange(n) // 1000
    props[:2000, 1] = 1.0  # blocked_in_stat: excluded when stats_exclude_blocked
    nb = NativeBlock.fixed_indegree(props, ne, 80, max_delay=3, conn_seed=2, noise_seed=5,
                                    stats_exclude_blocked=True)
    nb.attach_bold(steps_per_sample=10, delta_t=1e-3)
    steps = 600
    freqs = np.zeros((steps, nb.S), np.uint32)
    bold = np.zeros((steps // 10, nb.S), np.float64)
    nb.run(steps, noise_rate=2.0 ** -7, graph=True, freqs=freqs, bold=bold)
    sizes = nb.get("pop_sizes").astype(np.float64)
    assert list(sizes) == [0, 0, 1000, 1000, 1000, 1000]
    ob = orc.OracleBOLD(nb.S)
    for w in range(steps // 10):
        with np.errstate(invalid="ignore", divide="ignore"):
            u = freqs[w * 10:(w + 1) * 10].sum(0) / sizes
        u = np.where(u > 1e-5, u, np.where(np.isnan(u), u, 1e-5))
        want = ob.run(u)
        # empty voxels get u = 0/0; the f_in >= 0 clamp (models/bold_model_pytorch.py:74)
        # maps the NaN back to 0, so both sides stay finite and must agree too
        np.testing.assert_allclose(bold[w], want, rtol=1e-10, atol=1e-14, equal_nan=True)
    np.testing.assert_allclose(nb.bold_state("q")[2:], ob.q[2:], rtol=1e-10)
    assert freqs.sum() > 0


def test_results_independent_of_delivery_launch_shape():
    """Dyadic weights make every accumulation exact, so any launch shape of
    the delivery kernel must give bit-identical state."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K, D = 3000, 2400, 120, 8
    rng = np.random.default_rng(6)
    props = _plain_like_props(n, ne, rng)
    ref = orc.gen_fixed_indegree(n, ne, K, D, 0.0, 1.0, False, seed=8)
    ref["w0"] = (np.floor(ref["w0"] * 256) / 256).astype(np.float32)
    outs = []
    for mode, lanes, grid, sg, ept, ps, occ in ((0, 64, 4, 256, 32, 0, 1), (0, 32, 37, 256, 32, 0, 1),
                                                (0, 16, 300, 256, 32, 0, 1), (0, 64, 2048, 256, 32, 0, 1),
                                                (1, 64, 1, 256, 32, 0, 1), (1, 64, 300, 256, 16, 0, 2),
                                                (1, 64, 7, 64, 16, 1, 1), (1, 64, 1000, 128, 32, 1, 2),
                                                (1, 64, 3, 256, 32, 1, 1), (1, 64, 5, 128, 16, 2, 2),
                                                (1, 64, 900, 64, 16, 2, 1)):
        nb = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], None, ref["delay"],
                                  dt=0.1, noise_seed=2)
        nb.set_launch(deliver_lanes=lanes, deliver_grid=grid, deliver_mode=mode, neuron_psets=mode, stage_segs=sg,
                      stage_ept=ept, accum_occ=occ, stage_variant=ps)
        want = {"deliver_grid": grid, "deliver_lanes": lanes, "deliver_mode": mode, "neuron_psets": mode,
                "stage_segs": sg, "stage_ept": ept, "stage_variant": ps, "accum_occ": occ}
        got = nb.get_launch()  # neuron_psets reports one shared parameter set when enabled
        assert {k: got[k] for k in want} == want
        ev = np.zeros(300, np.uint64)
        nb.run(300, noise_rate=0.02, graph=True, events=ev)
        outs.append((nb.get("V"), nb.get("J"), ev))
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)
    assert outs[0][2].sum() > 0


@pytest.mark.parametrize("pair", [False, True])
def test_bucketed_delivery_matches_atomics_including_overflow(pair):
    """Bucketed delivery (staging + LDS sums, with the atomic fallback of full
    buckets) against direct atomics: dense activity across several buckets,
    dyadic weights, so both must agree bitwise."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K, D = 20000, 16000, 400, 3
    rng = np.random.default_rng(9)
    props = _plain_like_props(n, ne, rng, g_scale=0.2)
    ref = orc.gen_fixed_indegree(n, ne, K, D, 0.0, 1.0, pair, seed=10)
    for key in ("w0", "w1"):
        if ref.get(key) is not None:
            ref[key] = (np.floor(ref[key] * 64) / 64).astype(np.float32)
    outs = []
    for mode, var, occ in ((0, 0, 1), (1, 0, 1), (1, 1, 2), (1, 2, 1), (1, 2, 2)):
        nb = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], ref["w1"] if pair else None,
                                  ref["delay"], dt=0.1, noise_seed=4)
        nb.set_launch(deliver_mode=mode, deliver_grid=512, accum_occ=occ, stage_variant=var)
        ev = np.zeros(40, np.uint64)
        # 20 steps of heavy noise (up to ~10^7 events per step: buckets overflow), then 20 light ones
        nb.run(20, noise_rate=0.25, graph=True, events=ev[:20])
        nb.run(20, noise_rate=0.01, graph=False, events=ev[20:])
        outs.append((nb.get("V"), nb.get("J"), ev))
    for o in outs[1:]:
        for x, y in zip(outs[0], o):
            assert np.array_equal(x, y)
    assert outs[0][2][:20].max() > 3 * 8192 * 3  # some bucket saw more events than its staging room


def test_large_block_full_size_event_identity():
    """The benchmark's single large block at full size (5e6 neurons, 5e9
    synapses, delays 1-10): with noise-only spikes during 4 steps and then
    10 silent steps, every spike's whole row is delivered inside the window
    and nothing else is, so the delivered events equal the spiking rows'
    out-degrees exactly; the state stays finite."""
    from digital_twin_brain_amd import networks

    nb, props = networks.build_large()
    try:
        D = nb.max_delay
        outdeg = np.diff(nb.get("rowptr"))
        # the launch made at creation, then the benchmark's tuned shape (LDS-sorted staging)
        for launch in ({}, dict(deliver_mode=1, deliver_grid=1024, stage_segs=256, stage_ept=16, stage_variant=2)):
            if launch:
                nb.set_launch(**launch)
            spiked = 0
            delivered = 0
            for k in range(4 + D):
                ev = np.zeros(1, np.uint64)
                nb.run(1, noise_mode=0, noise_rate=0.002 if k < 4 else 0.0, isolated=True, graph=False, events=ev)
                delivered += int(ev[0])
                if k < 4:
                    sl = nb.get("spike_list")
                    assert sl.size > 5000
                    spiked += int(outdeg[sl].sum())
                else:
                    assert nb.get("spike_list").size == 0
            assert delivered == spiked > 0, launch
        assert np.isfinite(nb.get("V")).all() and np.isfinite(nb.get("J")).all()
    finally:
        nb.close()


def test_large_block_full_size_weight_identity():
    """Full-size linearity: from rest (J = 0), one step's arrivals summed over
    all 5e6 neurons equal, per receptor group, the weights of that step's
    spiking rows' delay-1 segments, summed on the GPU from the block's own CSR
    (int32 fixed-point sums: within half a quantum per event)."""
    import torch

    from digital_twin_brain_amd import networks

    nb, props = networks.build_large()
    try:
        S = nb.get_launch()["accum_fixed"]
        assert S >= 16
        ev = np.zeros(1, np.uint64)
        nb.run(1, noise_mode=0, noise_rate=0.002, isolated=True, graph=False, events=ev)
        J = nb.get("J").astype(np.float64)
        dev = torch.device("cuda")
        spk = torch.from_numpy(nb.get("spike_list").astype(np.int64)).to(dev)
        w = torch.empty(nb.nnz, dtype=torch.float32, device=dev)
        dl = torch.empty(nb.nnz, dtype=torch.uint8, device=dev)
        kd = torch.empty(nb.nnz, dtype=torch.uint8, device=dev)
        nb.get("w0", out=w)
        nb.get("delay", out=dl)
        nb.get("kind", out=kd)
        rp = torch.from_numpy(nb.get("rowptr")).to(dev)
        starts = rp[spk]
        lens = rp[spk + 1] - starts
        total = int(lens.sum())
        offs = torch.cumsum(lens, 0) - lens
        idx = starts.repeat_interleave(lens) + (torch.arange(total, device=dev) - offs.repeat_interleave(lens))
        ws, d1, k = w[idx].double(), dl[idx] == 1, kd[idx]
        exc = float(ws[d1 & (k == 0)].sum())
        inh = float(ws[d1 & (k == 1)].sum())
        n_ev = int(d1.sum())
        del w, dl, kd, idx, ws
        torch.cuda.empty_cache()
        assert n_ev == int(ev[0]) > 0
        tol = n_ev * 2.0 ** -(S + 1) + 1e-6 * (exc + inh)
        for ch, want in ((0, exc), (1, exc), (2, inh), (3, inh)):
            assert abs(J[ch].sum() - want) <= tol, (ch, J[ch].sum(), want, tol)
    finally:
        nb.close()


def test_region_share_full_size_event_identity():
    """The 90-region share at full size (2e7 neurons, 1e10 synapses, delays
    1-40, external drive): the same closed-window identity as for the large
    block, for direct scattered and LDS-sorted staging."""
    from digital_twin_brain_amd import networks

    nb, props, spec, _ = networks.build_regions()
    try:
        D = nb.max_delay
        outdeg = np.diff(nb.get("rowptr"))
        for var in (0, 2):
            nb.set_launch(deliver_mode=1, deliver_grid=1024, stage_segs=128, stage_ept=16, stage_variant=var)
            spiked = 0
            delivered = 0
            for k in range(3 + D):
                ev = np.zeros(1, np.uint64)
                nb.run(1, noise_mode=0, noise_rate=0.001 if k < 3 else 0.0, isolated=True, graph=False, events=ev)
                delivered += int(ev[0])
                if k < 3:
                    spiked += int(outdeg[nb.get("spike_list")].sum())
            assert delivered == spiked > 0, var
        assert np.isfinite(nb.get("V")).all()
    finally:
        nb.close()


def test_autotuner_contract_and_cache(tmp_path, monkeypatch):
    """The tuner re-measures its finalists, keeps the best of them, leaves
    the block on that shape, and a second call is served from the cache
    (which restores the same shape on a fresh block)."""
    from digital_twin_brain_amd import NativeBlock, autotune

    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    n, ne, K, D = 20000, 16000, 100, 4
    rng = np.random.default_rng(21)
    props = _plain_like_props(n, ne, rng, g_scale=0.3)
    ref = orc.gen_fixed_indegree(n, ne, K, D, 0.0, 1.0, False, seed=22)
    kw = dict(noise_mode=0, noise_rate=0.01, isolated=True)  # stationary spike counts: a stable cache key
    keys = ("deliver_mode", "deliver_lanes", "deliver_grid", "stage_segs", "stage_ept", "stage_variant", "accum_occ")

    def block():
        return NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"], None, ref["delay"],
                                    dt=0.1, noise_seed=2)

    nb = block()
    best, table = autotune.tune_delivery(nb, kw, steps=10, use_cache=True)
    finalists = {k[:-1]: v for k, v in table.items() if k.endswith("*")}
    assert 1 <= len(finalists) <= 6
    # when the persistent launch wins, best keeps the winning per-kernel shape and its mode
    shape = dict(best, deliver_mode=best.get("shape_mode", best["deliver_mode"]))
    assert best["deliver_mode"] in (0, 1) or ("shape_mode" in best and best["deliver_mode"] == 2)
    key = "/".join(str(shape[k]) for k in keys)
    assert finalists[key] == min(finalists.values()) == best["us_per_1e6_events"]
    got = nb.get_launch()
    assert all(got[k] == best[k] for k in keys)
    nb2 = block()
    best2, table2 = autotune.tune_delivery(nb2, kw, steps=10, use_cache=True)
    assert "cached" in table2 and best2 == best
    got2 = nb2.get_launch()
    assert all(got2[k] == best[k] for k in keys)


def test_gpu_members_generator_csr_bit_exact():
    """Uneven packed members (the batched ensemble layout) against the oracle."""
    from digital_twin_brain_amd import NativeBlock

    members = [(500, 400, 30), (2, 1, 1), (2500, 2000, 120), (1, 1, 0), (900, 700, 60)]
    n = sum(m[0] for m in members)
    props = np.zeros((n, 22), np.float32)
    props[:, 4] = 1; props[:, 5] = 2; props[:, 6] = 0.03; props[:, 7] = -75; props[:, 8] = -50; props[:, 9] = -65
    props[:, 18:22] = (2, 40, 10, 50)
    for pair, D in ((False, 1), (True, 6)):
        nb = NativeBlock.members(props, members, max_delay=D, pair_weights=pair, conn_seed=21)
        ref = rows_by_delay(orc.gen_members(members, D, 0.0, 1.0, pair, seed=21))
        for k in ("rowptr", "post", "kind", "delay", "w0"):
            assert np.array_equal(nb.get(k), ref[k]), k
        assert np.array_equal(nb.get("w1"), ref["w1"] if pair else ref["w0"])


def _region_case():
    from digital_twin_brain_amd import networks

    sizes = np.array([700, 1500, 400, 1200, 900])
    W = networks.connectome(5, seed=4)
    cdf = networks.source_cdf(W)
    props, regs = networks.region_props(sizes, k_total=60)
    spec = [(n, ne, 40, 20) for n, ne in regs]
    return props, spec, cdf


def test_gpu_regions_generator_csr_bit_exact():
    """Regions with inter-region synapses drawn through the source table."""
    from digital_twin_brain_amd import NativeBlock

    props, spec, cdf = _region_case()
    for pair, D in ((False, 40), (True, 7)):
        nb = NativeBlock.regions(props, spec, cdf, max_delay=D, pair_weights=pair, conn_seed=17)
        ref = rows_by_delay(orc.gen_regions(spec, cdf, D, 0.0, 1.0, pair, seed=17))
        for k in ("rowptr", "post", "kind", "delay", "w0"):
            assert np.array_equal(nb.get(k), ref[k]), k
        assert np.array_equal(nb.get("w1"), ref["w1"] if pair else ref["w0"])


def test_region_replicas_equal_the_block_diagonal_region_network():
    """dtb_block_create_region_replicas (one source table shared by every
    replica, sources drawn inside the replica) builds the same CSR, bit for
    bit, as dtb_block_create_regions given the block-diagonal table of all
    replicas -- which the test above pins to the oracle generator."""
    from digital_twin_brain_amd import NativeBlock

    props, spec, cdf = _region_case()
    reps = 3
    R = len(spec)
    big = np.zeros((reps * R, reps * R), np.uint64)
    for r in range(reps):
        big[r * R:(r + 1) * R, r * R:(r + 1) * R] = cdf
        big[r * R:(r + 1) * R, (r + 1) * R:] = np.uint64(1 << 32)
    allp = np.concatenate([props] * reps)
    for pair, D in ((False, 12), (True, 3)):
        a = NativeBlock.region_replicas(allp, spec, reps, cdf, max_delay=D, pair_weights=pair, conn_seed=23)
        b = NativeBlock.regions(allp, spec * reps, big, max_delay=D, pair_weights=pair, conn_seed=23)
        assert a.n == b.n == reps * int(sum(x[0] for x in spec)) and a.nnz == b.nnz
        for k in ("rowptr", "post", "kind", "delay", "w0", "w1"):
            assert np.array_equal(a.get(k), b.get(k)), k
        # no synapse leaves its replica
        rp = a.get("rowptr")
        pre = np.repeat(np.arange(a.n), np.diff(rp))
        m = int(sum(x[0] for x in spec))
        assert np.array_equal(pre // m, a.get("post") // m)
        a.close()
        b.close()


@pytest.mark.parametrize("mode", [0, 1])
def test_region_network_with_external_drive_matches_oracle(mode):
    """Region network, delays 1..40, external Poisson drive: the GPU against
    the FP64 push oracle with the same thresholds and the same Philox draws."""
    from digital_twin_brain_amd import NativeBlock
    from digital_twin_brain_amd.engine import poisson_thresholds

    props, spec, cdf = _region_case()
    nb = NativeBlock.regions(props, spec, cdf, max_delay=40, conn_seed=5, noise_seed=8)
    nb.set_launch(deliver_mode=mode)
    thr = nb.set_external(0.15, 0.5)
    assert np.array_equal(thr, poisson_thresholds(0.15))
    csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
               w0=nb.get("w0"), w1=None)
    ob = orc.OracleBlock(props, csr, 0.1)
    ob.set_external(thr, 0.5)
    steps, B = 400, 50
    freqs = np.zeros((B, nb.S), np.uint32)
    events = np.zeros(B, np.uint64)
    diverged = None
    ospk = []
    for b0 in range(0, steps, B):
        nb.run(B, noise_mode=1, graph=True, freqs=freqs, events=events)
        for k in range(B):
            sa, _ = ob.run(noise_mode=1, seed=8)
            ospk.append(ob.get("spike"))
            if diverged is None and not np.array_equal(freqs[k].astype(np.float64), sa):
                diverged = b0 + k + 1
        if diverged is not None:
            break
        np.testing.assert_allclose(nb.get("V"), ob.get("V"), rtol=RTOL)
        Jo = ob.get("J")
        np.testing.assert_allclose(nb.get("J"), Jo, rtol=RTOL, atol=RTOL * np.abs(Jo).max())
        arr = arrivals_per_step(csr, ospk, 40)[b0:b0 + B]
        assert np.array_equal(events.astype(np.int64), arr)
    print(f"regions mode={mode}: first divergence {diverged}, spikes {int(np.sum(ospk))}")
    assert diverged is None or diverged > 200
    assert np.sum(ospk) > 0
    # the external drive is what keeps J_AMPA of an otherwise silent network up
    assert ob.get("J")[0].mean() > 0


@pytest.mark.parametrize("pair", [False, True])
def test_fixed_point_accumulation_matches_float(pair):
    """Bucketed delivery with int32 fixed-point sums (the default when the
    block's largest per-step input sum allows 2^S with S >= 16) against fp32
    LDS sums: bit-identical on dyadic weights, within 1e-6 on random ones."""
    from digital_twin_brain_amd import NativeBlock

    n, ne, K, D = 20000, 16000, 200, 4
    rng = np.random.default_rng(12)
    props = _plain_like_props(n, ne, rng, g_scale=0.5)
    for dyadic in (True, False):
        ref = orc.gen_fixed_indegree(n, ne, K, D, 0.0, 1.0, pair, seed=13)
        if dyadic:
            for key in ("w0", "w1"):
                if ref.get(key) is not None:
                    ref[key] = (np.floor(ref[key] * 64) / 64).astype(np.float32)
        outs = []
        for fixed in (1, 0):
            nb = NativeBlock.from_csr(props, ref["rowptr"], ref["post"], ref["kind"], ref["w0"],
                                      ref["w1"] if pair else None, ref["delay"], dt=0.1, noise_seed=4)
            nb.set_launch(deliver_mode=1, deliver_grid=256, accum_fixed=fixed)
            S = nb.get_launch()["accum_fixed"]
            assert (S >= 16) if fixed else S == 0
            ev = np.zeros(100, np.uint64)
            # isolated noise spikes: the arrivals do not feed back into the spikes
            nb.run(100, noise_rate=0.01, isolated=True, graph=True, events=ev)
            outs.append((nb.get("J"), nb.get("V"), ev))
        assert np.array_equal(outs[0][2], outs[1][2]) and outs[0][2].sum() > 0
        if dyadic:
            assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        else:
            Jf, Jr = outs[0][0], outs[1][0]
            np.testing.assert_allclose(Jf, Jr, rtol=1e-6, atol=1e-6 * np.abs(Jr).max())
            np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-5)


def _heterogeneous_csr(n, ne, seed):
    """Per-target in-degree log-uniform in [10, 1000] (sources uniform, not
    the target itself), delays 1-3, weights U[0, 1) times a per-target scale
    log-uniform in [1e-3, 1]: per-step input sums that differ by about six
    orders of magnitude between the neurons of one block."""
    rng = np.random.default_rng(seed)
    k_of = np.exp(rng.uniform(np.log(10), np.log(1000), n)).astype(np.int64)
    scale = np.exp(rng.uniform(np.log(1e-3), 0.0, n))
    post = np.repeat(np.arange(n), k_of)
    pre = rng.integers(0, n - 1, post.shape[0])
    pre = pre + (pre >= post)
    o = np.argsort(pre, kind="stable")
    pre, post = pre[o], post[o]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, pre + 1, 1)
    w = (rng.random(post.shape[0]) * scale[post]).astype(np.float32)
    csr = dict(rowptr=np.cumsum(rp), post=post.astype(np.int32), kind=(pre >= ne).astype(np.uint8),
               delay=rng.integers(1, 4, post.shape[0]).astype(np.uint8), w0=w, w1=None)
    return csr, k_of, scale


@pytest.mark.parametrize("mode", [1, 0])
def test_fixed_point_per_neuron_scale_on_heterogeneous_inputs(mode):
    """Targets whose inputs differ in scale by 1e3 and in count by 1e2 within
    one block: every neuron's J is within 1e-5 of the FP64 oracle relative to
    that neuron's own input scale (its mean incoming weight on the channel),
    for bucketed fixed-point delivery (mode 1, the default for large blocks)
    and direct float atomics (mode 0)."""
    from digital_twin_brain_amd import NativeBlock

    n, ne = 30000, 24000
    csr, k_of, scale = _heterogeneous_csr(n, ne, 21)
    rng = np.random.default_rng(22)
    props = _plain_like_props(n, ne, rng, g_scale=0.5)
    nb = NativeBlock.from_csr(props, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], None, csr["delay"], dt=0.1,
                              noise_seed=5)
    nb.set_launch(deliver_mode=mode, deliver_grid=256)
    if mode == 1:
        assert nb.get_launch()["accum_fixed"] >= 16, "fixed-point accumulation not selected"
    # the stored weights read back unscaled, bit for bit
    lib_csr = dict(rowptr=nb.get("rowptr"), post=nb.get("post"), kind=nb.get("kind"), delay=nb.get("delay"),
                   w0=nb.get("w0"), w1=None)
    ref = rows_by_delay(csr)
    for k in ("rowptr", "post", "kind", "delay", "w0"):
        assert np.array_equal(lib_csr[k], ref[k]), f"CSR {k} read back differs"
    ob = orc.OracleBlock(props, lib_csr, 0.1)
    # per (neuron, receptor group): mean incoming weight = the neuron's own input scale
    post = lib_csr["post"].astype(np.int64)
    grp = lib_csr["kind"].astype(np.int64)
    wsum = np.zeros((2, n))
    wcnt = np.zeros((2, n))
    np.add.at(wsum, (grp, post), lib_csr["w0"].astype(np.float64))
    np.add.at(wcnt, (grp, post), 1.0)
    wbar = np.where(wcnt > 0, wsum / np.maximum(wcnt, 1), 0.0)
    wbar4 = np.stack([wbar[0], wbar[0], wbar[1], wbar[1]])
    ev = np.zeros(50, np.uint64)
    worst = 0.0
    for b in range(4):
        # noise spikes only: the arrivals do not feed back into the raster
        nb.run(50, noise_rate=0.02, isolated=True, graph=True, events=ev)
        for _ in range(50):
            ob.run(noise_rate=0.02, seed=5, isolated=True)
        Jg, Jo = nb.get("J").astype(np.float64), ob.get("J")
        err = np.abs(Jg - Jo)
        tol = RTOL * (np.abs(Jo) + wbar4)
        assert (err <= tol).all(), (
            f"batch {b}: J beyond 1e-5 of the neuron's own scale at {np.argwhere(err > tol)[:4].tolist()}, worst "
            f"{float((err / np.maximum(tol, 1e-300)).max()):.3g} x tolerance")
        worst = max(worst, float((err / np.maximum(tol, 1e-300)).max()))
        np.testing.assert_allclose(nb.get("V"), ob.get("V"), rtol=RTOL)
    # the test spans the scales it claims: small- and large-scale neurons both received input
    got = np.abs(Jo).sum(0) > 0
    assert scale[got].min() < 3e-3 and scale[got].max() > 0.3 and k_of[got].min() < 30
    print(f"mode {mode}: worst J error {worst:.3g} x tolerance, launch {nb.get_launch()}")
    nb.close()
