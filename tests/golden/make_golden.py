"""Write the golden vectors of the reference block step and BOLD model.

Run once, in the build container only (the reference tree is not present on
GPU machines):  python tests/golden/make_golden.py /path/to/reference

It imports the reference models/block.py and models/bold_model_pytorch.py,
runs them in FP64 (torch default dtype float64) on tiny seeded networks, with
torch.rand replaced by the Philox noise uniforms of the oracle (the same
noise-injection technique as cuda/python/test_demo.py:103-110), and stores
inputs and outputs as .npz fixtures next to this script.  Nothing from the
reference is copied into the fixtures except the numbers it computed.
"""
from __future__ import annotations

import os
import sys
from unittest import mock

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from oracle import oracle as orc  # noqa: E402


def make_props(n, n_exc, rng, g_lo, g_hi, tau=(2, 40, 10, 50), T_ref=2.0, g_L=0.03, noise=0.0, n_sub=None):
    """22-column property matrix (generation/make_block.py:394-434)."""
    p = np.zeros((n, 22), dtype=np.float32)
    p[:, 0] = noise
    p[:, 2] = 0.0
    if n_sub is None:
        p[:n_exc, 3] = 0
        p[n_exc:, 3] = 1
    else:
        p[:, 3] = np.arange(n) * n_sub // n
    p[:, 4] = 1.0
    p[:, 5] = T_ref
    p[:, 6] = g_L
    p[:, 7] = -75.0
    p[:, 8] = -50.0
    p[:, 9] = -65.0
    p[:, 10:14] = rng.uniform(g_lo, g_hi, size=(n, 4)).astype(np.float32)
    p[:, 14:18] = np.array([0, 0, -70, -100], dtype=np.float32)
    p[:, 18:22] = np.array(tau, dtype=np.float32)
    return p


def csr_to_torch_sparse(n, csr):
    """[4, N_post, N_pre] COO from presynaptic-major CSR, channel pairs as in
    generation/read_block.py:64-90."""
    rowptr = csr["rowptr"]
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    post = csr["post"].astype(np.int64)
    off = csr["kind"].astype(np.int64) * 2
    # float32 values, as generation/read_block.py loads them; models/block.py:72
    # multiplies by active.float(), so W.s is a float32 product even when the
    # torch default dtype is float64.
    w0 = csr["w0"].astype(np.float32)
    w1 = (csr["w1"] if csr.get("w1") is not None else csr["w0"]).astype(np.float32)
    ch = np.concatenate([off, off + 1])
    rows = np.concatenate([post, post])
    cols = np.concatenate([pre, pre])
    vals = np.concatenate([w0, w1])
    idx = torch.from_numpy(np.stack([ch, rows, cols]))
    return torch.sparse_coo_tensor(idx, torch.from_numpy(vals), (4, n, n)).coalesce()


def drop_inputs(n, csr, targets):
    """The CSR without the synapses onto `targets` (neurons left with no
    afferents, which models/block.py:38-48 takes as the spike sources)."""
    rowptr = csr["rowptr"]
    row = np.repeat(np.arange(n), np.diff(rowptr))
    keep = ~np.isin(csr["post"], targets)
    out = {k: (v[keep] if isinstance(v, np.ndarray) and v.shape[:1] == keep.shape else v) for k, v in csr.items()}
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, row[keep] + 1, 1)
    out["rowptr"] = np.cumsum(rp)
    return out


def dyadic(csr, bits=8):
    """Round weights to multiples of 2**-bits so every float32 partial sum of
    W.s is exact (the reference's float32 product then equals the FP64 sum)."""
    out = dict(csr)
    for k in ("w0", "w1"):
        if out.get(k) is not None:
            out[k] = (np.floor(out[k] * 2 ** bits) / 2 ** bits).astype(np.float32)
    return out


def run_block_case(ref_block, name, n, n_exc, csr, props, steps, noise_rate, isolated, dt, seed, dense,
                   keep_every=1, snap_steps=(), src=None):
    torch.set_default_dtype(torch.float64)
    W = csr_to_torch_sparse(n, csr)
    if dense:
        W = W.to_dense()
    kw = {} if src is None else dict(src=torch.from_numpy(src.astype(bool)))
    B = ref_block(node_property=torch.from_numpy(props.astype(np.float64)), w_uij=W, delta_t=dt, **kw)
    step_box = [0]

    def fake_rand(*size, **kw):
        return torch.from_numpy(orc.noise_uniforms(seed, step_box[0], n).astype(np.float64))

    V_hist, act_hist, spk_hist, sum_hist, mean_hist = [], [], [], [], []
    snaps = {}
    with mock.patch("torch.rand", side_effect=fake_rand):
        for s in range(1, steps + 1):
            step_box[0] = s
            u = orc.noise_uniforms(seed, s, n)
            sa, mv = B.run(noise_rate=noise_rate, isolated=isolated)
            if s % keep_every == 0:
                V_hist.append(B.V_i.numpy().copy())
            act_hist.append(B.active.numpy().astype(np.uint8))
            thr = B.active.numpy()
            noise = u < np.float64(noise_rate)
            spk_hist.append((noise | (thr & (not isolated))).astype(np.uint8))
            sum_hist.append(sa.numpy())
            mean_hist.append(mv.numpy())
            if s in snap_steps:
                snaps[f"J_{s}"] = B.J_ui.numpy().copy()
                snaps[f"I_ui_{s}"] = B.I_ui.numpy().copy()
                snaps[f"I_syn_{s}"] = B.I_syn.numpy().copy()
        extra = {}
        t_final = float(B.t)
        if src is not None:
            # the step at which the reference stops: update_Vi reads src[:, iter] once more
            # after advancing iter (models/block.py:105-107)
            extra = dict(src=src.astype(np.uint8), src_neuron=B.src_neuron.numpy().astype(np.int64), raises_at=-1)
            step_box[0] = steps + 1
            try:
                B.run(noise_rate=noise_rate, isolated=isolated)
            except IndexError:
                extra["raises_at"] = steps + 1
    out = dict(
        n=n, n_exc=n_exc, dt=dt, seed=seed, steps=steps, noise_rate=noise_rate, isolated=int(isolated),
        keep_every=keep_every, dense=int(dense), props=props, rowptr=csr["rowptr"], post=csr["post"],
        kind=csr["kind"], w0=csr["w0"], w1=csr["w1"] if csr.get("w1") is not None else csr["w0"],
        V=np.stack(V_hist), active=np.packbits(np.stack(act_hist), axis=1),
        spike=np.packbits(np.stack(spk_hist), axis=1), sum_act=np.stack(sum_hist), mean_v=np.stack(mean_hist),
        t_final=t_final, snap_steps=np.array(sorted(snap_steps), dtype=np.int64), **snaps, **extra)
    np.savez_compressed(os.path.join(HERE, f"block_{name}.npz"), **out)
    nsp = int(np.stack(act_hist).sum())
    print(f"{name}: n={n} steps={steps} threshold spikes={nsp} delivered={int(np.stack(spk_hist).sum())}")


def run_bold_case(ref_bold, name, nv, steps, seed):
    torch.set_default_dtype(torch.float64)
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 0.05, size=(steps, nv))
    u[:, 0] = 1e-5  # near-silent voxel
    u[steps // 2:, 1] = 0.2  # strong step input
    params = dict(epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02)
    bold = ref_bold(**params)
    outs, states = [], []
    for s in range(steps):
        o = bold.run(torch.from_numpy(u[s]))
        outs.append(o.numpy().copy())
        if (s + 1) % 10 == 0:
            states.append(np.stack([bold.s.numpy(), bold.q.numpy(), bold.v.numpy(), bold.f_in.numpy()]).copy())
    # state_update from an ensemble-shaped hidden state (DataAssimilation.py:400)
    w = rng.uniform(-0.5, 1.5, size=(3, nv // 3, 4))
    w[0, 0, 1] = -1.0
    w_saved = w.copy()  # state_update keeps a view of w[:, :, 0] as s, which run() then updates in place
    bold.state_update(torch.from_numpy(w))
    su = np.stack([bold.s.numpy(), bold.q.numpy(), bold.v.numpy(), bold.f_in.numpy()])
    o2 = bold.run(torch.from_numpy(u[0][: (nv // 3) * 3]))
    np.savez_compressed(os.path.join(HERE, f"bold_{name}.npz"), u=u, out=np.stack(outs), state=np.stack(states),
                        su_w=w_saved, su_state=su, su_out=o2.numpy(), **{k: float(v) for k, v in params.items()})
    print(f"bold {name}: nv={nv} steps={steps}")


def main(ref_root):
    sys.path.insert(0, ref_root)
    from models.block import block as ref_block  # noqa: E402
    from models.bold_model_pytorch import BOLD as ref_bold  # noqa: E402

    rng = np.random.default_rng(7)
    # 1) dense W, 64 neurons (48 E / 16 I), pair weights, strong coupling
    n, ne = 64, 48
    csr = dyadic(orc.gen_fixed_indegree(n, ne, 12, 1, 0.0, 1.0, True, seed=11))
    props = make_props(n, ne, rng, 0.01, 0.08)
    run_block_case(ref_block, "dense64", n, ne, csr, props, steps=400, noise_rate=2.0 ** -5, isolated=False,
                   dt=0.1, seed=5, dense=True, snap_steps=(1, 2, 50, 100, 200, 400))
    # 2) sparse W, 400 neurons, compact weights, 4 sub-populations, plain-block parameters
    n, ne = 400, 320
    csr = orc.gen_fixed_indegree(n, ne, 100, 1, 0.0, 1.0, False, seed=21)
    props = make_props(n, ne, rng, 0.001, 0.05, n_sub=4)
    # conductances near the calibrated small block of test/test_generation.py:313-314
    # (x3, +-20% per neuron) so the network fires at a few Hz instead of saturating
    g_cal = np.array([2.13037975e-02, 1.39240506e-04, 1.58227848e-01, 1.89873418e-02], dtype=np.float32)
    props[:, 10:14] = g_cal * 3.0 * rng.uniform(0.8, 1.2, size=(n, 4)).astype(np.float32)
    run_block_case(ref_block, "sparse400", n, ne, csr, props, steps=500, noise_rate=2.0 ** -7, isolated=False,
                   dt=0.1, seed=9, dense=False, keep_every=5, snap_steps=(1, 250, 500))
    # 3) isolated mode (noise only drives synapses), T_ref=5, dt=0.5
    n, ne = 64, 48
    csr = dyadic(orc.gen_fixed_indegree(n, ne, 16, 1, 0.0, 1.0, True, seed=31))
    props = make_props(n, ne, rng, 0.02, 0.1, T_ref=5.0)
    run_block_case(ref_block, "isolated64", n, ne, csr, props, steps=200, noise_rate=2.0 ** -4, isolated=True,
                   dt=0.5, seed=13, dense=True, snap_steps=(1, 100, 200))
    # 4) forced-spike sources (models/block.py:38-50,104-109): six neurons with no afferents driven by a
    # spike schedule of 201 columns; the reference runs 200 steps and raises at the 201st.  W is sparse:
    # with current torch, isinstance(w, torch.sparse.Tensor) holds for dense tensors too, and the
    # dense branch of models/block.py:42-45 is never taken
    n, ne = 64, 48
    csr = dyadic(orc.gen_fixed_indegree(n, ne, 12, 1, 0.0, 1.0, True, seed=41))
    srcs = np.array([0, 7, 20, 33, 50, 63])
    csr = drop_inputs(n, csr, srcs)
    props = make_props(n, ne, rng, 0.01, 0.08)
    sched = (rng.random((srcs.shape[0], 201)) < 0.08).astype(np.uint8)
    run_block_case(ref_block, "src64", n, ne, csr, props, steps=200, noise_rate=2.0 ** -6, isolated=False,
                   dt=0.1, seed=17, dense=False, snap_steps=(1, 100, 200), src=sched)
    run_bold_case(ref_bold, "v18", 18, 600, seed=3)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
