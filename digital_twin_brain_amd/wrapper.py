"""In-process replacement of the gRPC `BlockWrapper`
(cuda/python/dist_blockwrapper_pytorch.py:30-543).

`LocalBlockWrapper` keeps the client's surface -- `subblk_id`,
`total_subblks`, `total_neurons`, `neurons_per_subblk`, the `run()`
generator, `mul_property_by_subblk`, `gamma_property_by_subblk`,
`update_property`, `set_samples`, `shutdown` -- and runs the step in a local
libdtb block instead of a remote MPI/gRPC server.  Semantics follow the HIP
engine the client talks to: per-neuron noise from property column 0,
statistics exclude `blocked_in_stat` neurons.

Networks come from a directory of reference `block_{i}.npz` files
(generation/make_block.py:473-479, read like generation/read_block.py:12-97)
or from an already built `NativeBlock`.
"""
from __future__ import annotations

import os
import re
import time

import numpy as np
import torch

from . import _native as nat
from .engine import NativeBlock


def load_block_dir(path):
    """Merge the reference block_{i}.npz files of `path` into one network.

    Returns (props [N,22] float32, csr dict) with a presynaptic-major CSR:
    pre = bases[i + input_block_idx] + input_neuron_idx, post = bases[i] +
    output_neuron_idx, kind = input_channel_offset // 2, weights (w0, w1) the
    two columns of `weight` (generation/read_block.py:64-90).

    Input blocks (src_timestamp, src_neuron_idx instead of connections,
    cuda/src/configuration.cpp:753-762; generation/read_block.py:53-61) are
    neurons driven by a spike schedule: csr["inputs"] then holds
    dict(sources, steps, neurons) -- every neuron of an input block is a
    source, and source neuron bases[i] + src_neuron_idx[k] fires at the step
    of 0-based index src_timestamp[k] (cuda/src/block.cpp:195-221).  Without
    input blocks csr["inputs"] is None.
    """
    pat = re.compile(r"block_([0-9]+)\.npz")
    ids = sorted(int(pat.fullmatch(f).group(1)) for f in os.listdir(path) if pat.fullmatch(f))
    if not ids or ids != list(range(len(ids))):
        raise FileNotFoundError(f"no contiguous block_0..block_k .npz files in {path}")
    files = [np.load(os.path.join(path, f"block_{i}.npz")) for i in ids]
    is_input = []
    for i, f in enumerate(files):
        has = ("src_timestamp" in f.files, "src_neuron_idx" in f.files)
        if has[0] != has[1]:
            raise ValueError(f"block_{i}.npz has only one of src_timestamp and src_neuron_idx")
        is_input.append(has[0])
    sizes = [f["property"].shape[0] for f in files]
    bases = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    props, pre, post, kind, w0, w1 = [], [], [], [], [], []
    srcs, ev_t, ev_n = [], [], []
    for i, f in enumerate(files):
        props.append(np.asarray(f["property"], dtype=np.float32)[:, :nat.NPROP])
        if is_input[i]:
            ts = np.asarray(f["src_timestamp"]).astype(np.int64)
            nid = np.asarray(f["src_neuron_idx"]).astype(np.int64)
            if ts.shape != nid.shape:
                raise ValueError(f"block_{i}.npz: src_timestamp and src_neuron_idx differ in length")
            if nid.size and (nid.min() < 0 or nid.max() >= sizes[i] or ts.min() < 0):
                raise ValueError(f"block_{i}.npz: src_neuron_idx outside the block or negative timestamps")
            srcs.append(bases[i] + np.arange(sizes[i], dtype=np.int64))
            ev_t.append(ts)
            ev_n.append(bases[i] + nid)
            continue
        blk = np.asarray(f["input_block_idx"]).astype(np.int64)
        pre.append(bases[i + blk] + np.asarray(f["input_neuron_idx"]).astype(np.int64))
        post.append(bases[i] + np.asarray(f["output_neuron_idx"]).astype(np.int64))
        kind.append((np.asarray(f["input_channel_offset"]).astype(np.int64) // 2).astype(np.uint8))
        w = np.asarray(f["weight"], dtype=np.float32).reshape(-1, 2)
        w0.append(w[:, 0])
        w1.append(w[:, 1])
    props = np.concatenate(props)
    pre = np.concatenate(pre)
    post = np.concatenate(post)
    kind = np.concatenate(kind)
    w0 = np.concatenate(w0)
    w1 = np.concatenate(w1)
    n = props.shape[0]
    order = np.lexsort((post, pre))
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, pre + 1, 1)
    rowptr = np.cumsum(rowptr)
    w0, w1 = w0[order], w1[order]
    inputs = None
    if srcs:
        inputs = dict(sources=np.concatenate(srcs), steps=np.concatenate(ev_t), neurons=np.concatenate(ev_n))
    return props, dict(rowptr=rowptr, post=post[order].astype(np.int32), kind=kind[order], w0=w0,
                       w1=None if np.array_equal(w0, w1) else w1, inputs=inputs)


def save_block_dir(path, props, csr, sizes, inputs=None):
    """Write a network as reference block_{i}.npz files (the fields and dtypes
    of generation/make_block.py:473-479, read back by
    cuda/src/configuration.cpp:586-594,783-792,972-975 and load_block_dir).

    `csr` is a presynaptic-major CSR dict (rowptr, post, kind, w0, optional w1,
    optional delay); `sizes` splits the neurons into consecutive blocks.  Each
    block holds the synapses onto its own neurons, ordered by (post, pre):
    output_neuron_idx (u32, local), input_block_idx (i16, source block minus
    this block), input_neuron_idx (u32, local in the source block),
    input_channel_offset (u8: 0 excitatory, 2 inhibitory), weight (f32 [nnz,2])
    and property (f32 [n,22]).  The format has no delay field, so delays other
    than 1 are refused.  `inputs` maps a block index to its spike schedule
    (src_timestamp, src_neuron_idx as local ids): that block is written as an
    input block (property, src_timestamp u32, src_neuron_idx u32) and must
    receive no synapses."""
    props = np.asarray(props, dtype=np.float32)
    sizes = np.asarray(sizes, dtype=np.int64)
    n = props.shape[0]
    if sizes.ndim != 1 or (sizes <= 0).any() or int(sizes.sum()) != n:
        raise ValueError(f"block sizes must be positive and sum to {n}")
    if len(sizes) > 32767:
        raise ValueError("at most 32767 blocks (input_block_idx is int16)")
    rowptr = np.asarray(csr["rowptr"], dtype=np.int64)
    post = np.asarray(csr["post"], dtype=np.int64)
    if rowptr.shape != (n + 1,) or rowptr[-1] != post.shape[0]:
        raise ValueError("rowptr does not match props and post")
    delay = csr.get("delay")
    if delay is not None and (np.asarray(delay) != 1).any():
        raise ValueError("block npz files have no delay field: every delay must be 1")
    kind = np.asarray(csr["kind"]).astype(np.uint8)
    w0 = np.asarray(csr["w0"], dtype=np.float32)
    w1 = w0 if csr.get("w1") is None else np.asarray(csr["w1"], dtype=np.float32)
    pre = np.repeat(np.arange(n, dtype=np.int64), np.diff(rowptr))
    bases = np.concatenate([[0], np.cumsum(sizes)])
    blk_pre = np.searchsorted(bases, pre, side="right") - 1
    order = np.lexsort((pre, post))
    os.makedirs(path, exist_ok=True)
    bounds = np.searchsorted(post[order], bases)
    inputs = inputs or {}
    for i in range(len(sizes)):
        sel = order[bounds[i]:bounds[i + 1]]
        if i in inputs:
            if sel.size:
                raise ValueError(f"block {i} is an input block but receives {sel.size} synapses")
            ts, nid = (np.asarray(x) for x in inputs[i])
            np.savez(os.path.join(path, f"block_{i}.npz"),
                     property=np.ascontiguousarray(props[bases[i]:bases[i + 1]]),
                     src_timestamp=ts.astype(np.uint32), src_neuron_idx=nid.astype(np.uint32))
            continue
        np.savez(os.path.join(path, f"block_{i}.npz"),
                 property=np.ascontiguousarray(props[bases[i]:bases[i + 1]]),
                 output_neuron_idx=(post[sel] - bases[i]).astype(np.uint32),
                 input_block_idx=(blk_pre[sel] - i).astype(np.int16),
                 input_neuron_idx=(pre[sel] - bases[blk_pre[sel]]).astype(np.uint32),
                 input_channel_offset=(2 * kind[sel]).astype(np.uint8),
                 weight=np.ascontiguousarray(np.stack([w0[sel], w1[sel]], axis=1)))


class LocalBlockWrapper:
    """`BlockWrapper` surface over an in-process libdtb block."""

    MAX_MESSAGE_LENGTH = 2147483647

    def __init__(self, address=None, path=None, delta_t=1.0, route_path=None, print_stat=False,
                 force_rebase=False, allow_rebase=True, overlap=2, network=None, device=0, noise_seed=0):
        if network is None:
            props, csr = load_block_dir(path)
            network = NativeBlock.from_csr(props, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], csr["w1"],
                                           dt=float(delta_t), noise_seed=noise_seed, stats_exclude_blocked=True,
                                           device=device)
            if csr["inputs"] is not None:
                # input blocks: neurons without membrane dynamics driven by their schedule
                inp = csr["inputs"]
                network.set_sources(inp["sources"], steps=inp["steps"], neurons=inp["neurons"], frozen=True)
        self.nb = network
        self.print_stat = print_stat
        self.delta_t = float(delta_t)
        self._dev = torch.device("cuda", network.device)
        props = network.get("props")
        sub = props[:, 3].astype(np.int64)
        ids, counts = np.unique(sub, return_counts=True)
        self._subblk_id = torch.from_numpy(ids).to(self._dev)
        self._neurons_per_subblk = torch.from_numpy(counts.astype(np.int64)).to(self._dev)
        self._ids_np = ids
        self._block_id = torch.zeros(1, dtype=torch.int64, device=self._dev)
        self._neurons_per_block = torch.tensor([network.n], dtype=torch.int64, device=self._dev)
        self._reset_hyper_parameter()
        self._sample_order = None
        self._sample_num = 0
        self._iterations = 0
        self._stat = []
        self._gamma_seed = 1

    # ------------------------------------------------------------ topology
    def _reset_hyper_parameter(self):
        # last multiplicative hyper-parameter per (sub-population, property column)
        self._last_hyper_parameter = torch.ones([self.nb.S, nat.NPROP], dtype=torch.float32)

    @property
    def total_neurons(self):
        return self._neurons_per_block.sum()

    @property
    def block_id(self):
        return self._block_id

    @property
    def subblk_id(self):
        return self._subblk_id

    @property
    def total_subblks(self):
        return self._subblk_id.shape[0]

    @property
    def neurons_per_subblk(self):
        return self._neurons_per_subblk

    @property
    def neurons_per_block(self):
        return self._neurons_per_block

    # ----------------------------------------------------------------- run
    def run(self, iterations, freqs=True, vmean=False, sample_for_show=False, imean=False, strategy=None,
            chunk=1000):
        """Per-step tuples (freqs[S], [vmean[S]], [spike[n_s], V[n_s]], [imean[S]]) as GPU tensors,
        cuda/python/dist_blockwrapper_pytorch.py:204-303."""
        if sample_for_show and self._sample_num == 0:
            raise RuntimeError("set_samples() first")
        S = self.nb.S
        sel = torch.from_numpy(self._ids_np).to(self._dev)
        full = len(self._ids_np) == S and np.array_equal(self._ids_np, np.arange(S))
        done = 0
        while done < iterations:
            k = min(chunk, iterations - done)
            t0 = time.perf_counter()
            # torch.empty launches nothing, so no fill on torch's stream can race
            # the library's copies on the block's stream; every element is written
            f = torch.empty((k, S), dtype=torch.int32, device=self._dev) if freqs else None
            vm = torch.empty((k, S), dtype=torch.float32, device=self._dev) if vmean else None
            im = torch.empty((k, S), dtype=torch.float32, device=self._dev) if imean else None
            ss = torch.empty((k, self._sample_num), dtype=torch.uint8, device=self._dev) if sample_for_show else None
            sv = torch.empty((k, self._sample_num), dtype=torch.float32, device=self._dev) if sample_for_show else None
            self.nb.run(k, noise_mode=nat.NOISE_PER_NEURON, graph=True, freqs=f, vmeans=vm, imeans=im,
                        sample_spike=ss, sample_v=sv)
            self._stat.append(time.perf_counter() - t0)
            outs = []
            if freqs:
                outs.append(f if full else f.index_select(1, sel))
            if vmean:
                outs.append(vm if full else vm.index_select(1, sel))
            if sample_for_show:
                if self._sample_order is not None:
                    ss = ss.index_select(1, self._sample_order)
                    sv = sv.index_select(1, self._sample_order)
                outs += [ss, sv]
            if imean:
                outs.append(im if full else im.index_select(1, sel))
            for j in range(k):
                yield tuple(o[j] for o in outs)
            done += k
        self._iterations += iterations

    def last_time_stat(self):
        import pandas as pd

        d = np.array(self._stat, dtype=np.float64)
        table = pd.DataFrame([[d.mean() if d.size else 0.0, d.std() if d.size else 0.0]],
                             index=pd.Index(["computing_duration"]), columns=["total_mean", "total_std"])
        return d, table

    # ---------------------------------------------------------- properties
    def _split(self, property_idx, values):
        assert isinstance(property_idx, torch.Tensor) and property_idx.dtype == torch.int64
        assert len(property_idx.shape) == 2 and property_idx.shape[1] == 2
        idx = property_idx.detach().cpu().numpy()
        return idx[:, 0].astype(np.int64), idx[:, 1].astype(np.int64), values

    def mul_property_by_subblk(self, property_idx, property_hyper_parameter, accumulate=False):
        """cuda/python/dist_blockwrapper_pytorch.py:446-458: unless `accumulate`,
        the factor applied is hp / previous hp, so the property ends at
        original * hp."""
        assert property_hyper_parameter.dtype == torch.float32
        pop, col, _ = self._split(property_idx, None)
        hp = property_hyper_parameter.detach().cpu().reshape(-1).clone()
        last = self._last_hyper_parameter
        if not accumulate:
            hp = hp / last[pop, col]
        last[pop, col] *= hp
        self.nb.mul_props_by_pop(pop.astype(np.uint32), col.astype(np.uint32), hp.numpy().astype(np.float32))

    def gamma_property_by_subblk(self, property_idx, gamma_concentration, gamma_rate, debug=False):
        """cuda/python/dist_blockwrapper_pytorch.py:460-503: property ~ Gamma(alpha, beta)."""
        if gamma_concentration is None:
            gamma_concentration = torch.ones_like(gamma_rate)
        pop, col, _ = self._split(property_idx, None)
        a = gamma_concentration.detach().cpu().numpy().astype(np.float32).reshape(-1)
        b = gamma_rate.detach().cpu().numpy().astype(np.float32).reshape(-1)
        assert (a >= 0).all() and (b > 0).all()
        self._last_hyper_parameter[pop, col] = 1.0
        self.nb.gamma_props_by_pop(pop.astype(np.uint32), col.astype(np.uint32), a, b, seed=self._gamma_seed)
        self._gamma_seed += 1
        return True

    # ---------------------------------------------------------- checkpoint
    def state_dict(self):
        """Everything later steps and property pushes depend on: the block's
        state snapshot, its current properties and the factors that relative
        property pushes divide by (numpy arrays)."""
        return {"block_state": self.nb.save_state(), "props": self.nb.get("props"),
                "last_hyper_parameter": self._last_hyper_parameter.numpy().copy(),
                "gamma_seed": np.int64(self._gamma_seed)}

    def load_state_dict(self, d):
        # the snapshot is validated before anything changes; the properties
        # then re-derive the refractory ends from the restored t_last
        self.nb.load_state(d["block_state"])
        self.nb.replace_props(d["props"])
        self._last_hyper_parameter = torch.from_numpy(np.array(d["last_hyper_parameter"], dtype=np.float32))
        self._gamma_seed = int(d["gamma_seed"])

    def update_property(self, property_idx, property_weight, bid=None):
        """cuda/python/dist_blockwrapper_pytorch.py:322-382: point-wise assignment."""
        assert property_weight.dtype == torch.float32
        neuron, col, _ = self._split(property_idx, None)
        self.nb.set_props(neuron.astype(np.uint32), col.astype(np.uint32),
                          property_weight.detach().cpu().numpy().astype(np.float32))
        self._reset_hyper_parameter()

    def set_samples(self, sample_idx, bid=None):
        """cuda/python/dist_blockwrapper_pytorch.py:505-528."""
        assert isinstance(sample_idx, torch.Tensor) and sample_idx.dtype == torch.int64
        order, recover = torch.unique(sample_idx, return_inverse=True)
        if order.shape[0] == sample_idx.shape[0] and bool((order == sample_idx).all()):
            self._sample_order = None
        else:
            self._sample_order = recover.to(self._dev)
        self._sample_num = int(order.shape[0])
        self.nb.set_samples(order.cpu().numpy().astype(np.uint32))

    # ------------------------------------------------------------- closing
    def close(self):
        pass

    def shutdown(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.close()
