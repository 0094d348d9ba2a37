"""Drop-in replacement of the reference `models.block.block` (models/block.py:5-177).

Same constructor, same `run(noise_rate, isolated)` returning
(sum_activate[S], mean_Vi[S]), same public attributes (`active`, `V_i`,
`J_ui`, `I_ui`, `I_syn`, `t`, `sub_idx`, and the property attributes), with
the step executed on the GPU by libdtb.  Outputs and attributes are torch
tensors on the device of `w_uij` (CPU by default, as in the reference).

Differences that follow from running on the GPU:
- the noise uniforms come from Philox4x32-10 keyed by (seed, step, neuron)
  instead of torch.rand; pass `seed=` to choose the stream;
- the state is FP32 and time is accumulated in double, so the refractory test
  `t >= t_last + T_ref` (models/block.py:100) is decided exactly as in a
  float64 run of the reference;
- `src` forced spikes (models/block.py:38-50,104-109): the sources are the
  neurons with no afferents, and column t of `src` drives the step whose
  0-based index is t, as in the reference.  The reference also reads column
  iter + 1 after advancing iter, for a write to V that the next line
  discards (:105-107), so it raises IndexError at its T-th step with T
  columns; this block runs all T steps and raises IndexError at step T + 1;
- the weights are stored per channel (pair form), so `update_conn_weight`
  can set AMPA and NMDA (or GABAa and GABAb) of one synapse independently.
  Setting a w_uij entry that has no synapse yet rebuilds the network on the
  GPU with that synapse added; the dynamic state carries over exactly.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .engine import NativeBlock


def csr_from_w(w_uij):
    """Presynaptic-major CSR from a [4, N, N] dense or sparse weight tensor.

    Channels (0, 1) become kind-0 synapses with weights (w0, w1) and channels
    (2, 3) kind-1 synapses (generation/read_block.py:68-74 builds the same
    pairing in the opposite direction).  A (pre, post) pair present in only
    one channel of a pair gets weight 0 on the other.  Returns a dict of
    numpy arrays; w1 is None when every pair carries equal weights.
    """
    if w_uij.dim() != 3 or w_uij.shape[0] != 4 or w_uij.shape[1] != w_uij.shape[2]:
        raise ValueError(f"w_uij must be [4, N, N], got {tuple(w_uij.shape)}")
    n = int(w_uij.shape[1])
    w = w_uij.detach()
    if w.is_sparse:
        w = w.coalesce()
        idx = w.indices().cpu().numpy()
        val = w.values().cpu().numpy().astype(np.float64)
    else:
        w = w.cpu()
        nz = torch.nonzero(w, as_tuple=False).T.numpy()
        idx = nz
        val = w.numpy()[nz[0], nz[1], nz[2]].astype(np.float64)
    ch, post, pre = idx[0].astype(np.int64), idx[1].astype(np.int64), idx[2].astype(np.int64)
    kinds, pres, posts, w0s, w1s = [], [], [], [], []
    for kind in (0, 1):
        sel = (ch // 2) == kind
        if not np.any(sel):
            continue
        key = pre[sel] * n + post[sel]
        uk, inv = np.unique(key, return_inverse=True)
        a0 = np.zeros(uk.shape[0])
        a1 = np.zeros(uk.shape[0])
        first = (ch[sel] % 2) == 0
        np.add.at(a0, inv[first], val[sel][first])
        np.add.at(a1, inv[~first], val[sel][~first])
        kinds.append(np.full(uk.shape[0], kind, np.uint8))
        pres.append(uk // n)
        posts.append(uk % n)
        w0s.append(a0)
        w1s.append(a1)
    if kinds:
        kind = np.concatenate(kinds)
        pre_a = np.concatenate(pres)
        post_a = np.concatenate(posts)
        w0 = np.concatenate(w0s)
        w1 = np.concatenate(w1s)
        order = np.lexsort((kind, post_a, pre_a))
        kind, pre_a, post_a, w0, w1 = kind[order], pre_a[order], post_a[order], w0[order], w1[order]
    else:
        kind = np.zeros(0, np.uint8)
        pre_a = post_a = np.zeros(0, np.int64)
        w0 = w1 = np.zeros(0)
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, pre_a + 1, 1)
    rowptr = np.cumsum(rowptr)
    w0 = w0.astype(np.float32)
    w1 = w1.astype(np.float32)
    return dict(n=n, rowptr=rowptr, post=post_a.astype(np.int32), kind=kind, w0=w0,
                w1=None if np.array_equal(w0, w1) else w1)


class block:
    """GPU LIF block with the interface of models/block.py:5-177."""

    def __init__(self, node_property, w_uij, delta_t=1, src=None, seed=0, device=0):
        assert len(w_uij.shape) == 3
        self.w_uij = w_uij
        self.src = src
        self.delta_t = delta_t
        self._out_device = w_uij.device
        self._dtype = torch.float32
        self._seed = int(seed)
        self._device = device
        csr = csr_from_w(w_uij)
        if csr["w1"] is None:
            csr["w1"] = csr["w0"].copy()
        props = node_property.detach().cpu().numpy().astype(np.float32)
        self._nb = NativeBlock.from_csr(props, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], csr["w1"],
                                        dt=float(delta_t), noise_seed=self._seed, device=device)
        self.N = self._nb.n
        if src is None:
            self.src_neuron = None
            self.iter = None
        else:
            # sources: the neurons no synapse targets (models/block.py:42-48)
            has_in = np.zeros(self.N, bool)
            has_in[csr["post"]] = True
            src_neuron = np.flatnonzero(~has_in)
            sched = torch.as_tensor(src).detach().cpu().numpy().astype(bool)
            assert sched.ndim == 2 and src_neuron.shape[0] == sched.shape[0]
            self.src_neuron = torch.from_numpy(src_neuron.astype(np.int64))
            self.iter = 0
            self._sched = sched
            self._apply_sources(self._nb)
        self._set_property_attrs(node_property)
        self._freq = np.zeros(self._nb.S, np.uint32)
        self._vmean = np.zeros(self._nb.S, np.float32)

    def _apply_sources(self, nb):
        k, t = np.nonzero(self._sched)
        nb.set_sources(self.src_neuron.numpy(), steps=t, neurons=self.src_neuron.numpy()[k])

    # reference: update_property (models/block.py:152-170)
    def _set_property_attrs(self, node_property):
        p = node_property.detach().to(self._out_device)
        E_I, blocked_in_stat, I_extern_Input, sub_idx, C, T_ref, g_Li, V_L, V_th, V_reset, g_ui, V_ui, tau_ui = \
            p.transpose(0, 1)[:22].split([1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 4, 4, 4])
        self.I_extern_Input = I_extern_Input.squeeze(0)
        self.V_ui = V_ui
        self.tau_ui = tau_ui
        self.g_ui = g_ui
        self.g_Li = g_Li.squeeze(0)
        self.V_L = V_L.squeeze(0)
        self.C = C.squeeze(0)
        self.sub_idx = sub_idx.squeeze(0)
        self.V_th = V_th.squeeze(0)
        self.V_reset = V_reset.squeeze(0)
        self.T_ref = T_ref.squeeze(0)

    def update_property(self, node_property):
        self._nb.replace_props(node_property.detach().cpu().numpy().astype(np.float32))
        self._set_property_attrs(node_property)
        return True

    def update_conn_weight(self, conn_idx, conn_weight):
        """w_uij[conn_idx] = conn_weight (models/block.py:172-177).  conn_idx
        holds the (channel, post, pre) indices of the w_uij entries, as three
        index arrays (a tuple/list, or one [3, X'] array); conn_weight is [X']."""
        if isinstance(conn_idx, (tuple, list)):
            idx = [torch.as_tensor(x).detach().cpu().numpy().reshape(-1) for x in conn_idx]
        else:
            a = torch.as_tensor(conn_idx).detach().cpu().numpy()
            if a.ndim != 2 or a.shape[0] != 3:
                raise ValueError("conn_idx must index the three dimensions of w_uij: (channel, post, pre)")
            idx = list(a)
        if len(idx) != 3:
            raise ValueError("conn_idx must index the three dimensions of w_uij: (channel, post, pre)")
        ch, post, pre = (np.asarray(x, np.int64) for x in idx)
        w = torch.as_tensor(conn_weight).detach().cpu().numpy().astype(np.float32).reshape(-1)
        w = np.broadcast_to(w, ch.shape).copy()
        if ((ch < 0) | (ch > 3)).any() or ((post < 0) | (post >= self.N) | (pre < 0) | (pre >= self.N)).any():
            raise IndexError("conn_idx out of range")
        try:
            self._nb.set_weights(pre, post, ch, w)
        except nat.DTBError as e:
            if e.code != nat.DTB_INVALID_PARAMETER or "no synapse" not in str(e):
                raise
            self._rebuild_with(ch, post, pre, w)
        if not self.w_uij.is_sparse:
            with torch.no_grad():
                self.w_uij[torch.from_numpy(ch), torch.from_numpy(post), torch.from_numpy(pre)] = \
                    torch.from_numpy(w).to(self.w_uij.dtype).to(self.w_uij.device)
        return True

    def _rebuild_with(self, ch, post, pre, w):
        """New network from the current one with the requested entries set
        (missing synapses added), continuing from the current state."""
        nb = self._nb
        rp = nb.get("rowptr")
        cur = dict(post=nb.get("post").astype(np.int64), kind=nb.get("kind"), w0=nb.get("w0"), w1=nb.get("w1"))
        pre_c = np.repeat(np.arange(self.N, dtype=np.int64), np.diff(rp))
        key = (pre_c * self.N + cur["post"]) * 2 + cur["kind"]
        w0, w1 = cur["w0"].copy(), cur["w1"].copy()
        add = {}
        order = np.argsort(key, kind="stable")
        for c, q, p, x in zip(ch, post, pre, w):
            k = (int(p) * self.N + int(q)) * 2 + int(c) // 2
            lo = np.searchsorted(key[order], k, side="left")
            hi = np.searchsorted(key[order], k, side="right")
            if hi > lo:
                hit = order[lo:hi]
                arr = w0 if c % 2 == 0 else w1
                arr[hit] = 0.0
                arr[hit[0]] = x
            else:
                a = add.setdefault(k, [0.0, 0.0])
                a[int(c) % 2] = x
        ks = np.array(sorted(add), dtype=np.int64)
        pre_all = np.concatenate([pre_c, ks // 2 // self.N])
        post_all = np.concatenate([cur["post"], ks // 2 % self.N])
        kind_all = np.concatenate([cur["kind"], (ks % 2).astype(np.uint8)])
        w0_all = np.concatenate([w0, np.array([add[k][0] for k in ks], np.float32)])
        w1_all = np.concatenate([w1, np.array([add[k][1] for k in ks], np.float32)])
        o = np.lexsort((kind_all, post_all, pre_all))
        rowptr = np.zeros(self.N + 1, np.int64)
        np.add.at(rowptr, pre_all + 1, 1)
        rowptr = np.cumsum(rowptr)
        snap = nb.save_state()
        new = NativeBlock.from_csr(nb.get("props"), rowptr, post_all[o].astype(np.int32), kind_all[o], w0_all[o],
                                   w1_all[o], dt=float(self.delta_t), noise_seed=self._seed, device=self._device)
        if self.src_neuron is not None:
            self._apply_sources(new)
        new.load_state(snap)
        nb.close()
        self._nb = new

    def _t(self, a):
        return torch.from_numpy(a).to(self._out_device)

    def run(self, noise_rate=0.01, isolated=False, uniforms=None):
        """One step (models/block.py:116-150).

        `uniforms` (optional, [N] float32) replaces the Philox noise uniforms
        of this step: spike iff uniforms[i] < noise_rate.
        """
        mode = nat.NOISE_SCALAR
        if uniforms is not None:
            mode = nat.NOISE_UNIFORMS
            uniforms = np.ascontiguousarray(np.asarray(uniforms, dtype=np.float32).reshape(-1))
        if self.src_neuron is not None and self.iter >= self._sched.shape[1]:
            raise IndexError(f"src has {self._sched.shape[1]} columns: no forced spikes for step {self.iter + 1}")
        self._nb.run(1, noise_mode=mode, noise_rate=float(noise_rate), uniforms=uniforms, isolated=isolated,
                     graph=False, freqs=self._freq, vmeans=self._vmean)
        if self.src_neuron is not None:
            self.iter += 1
        return self._t(self._freq.astype(np.float32)), self._t(self._vmean.copy())

    def run_steps(self, steps, noise_rate=0.01, isolated=False):
        """`steps` calls of run() in one library call (replayed as a captured
        graph, no host work between steps): returns (sum_activate[steps, S],
        mean_Vi[steps, S]), row t being what the t-th run() would return.
        run() pays a host round trip per step, since it hands back that step's
        statistics; this form is for callers that only need them afterwards."""
        steps = int(steps)
        if steps < 0:
            raise ValueError("steps must be >= 0")
        if self.src_neuron is not None and self.iter + steps > self._sched.shape[1]:
            raise IndexError(f"src has {self._sched.shape[1]} columns: no forced spikes past step "
                             f"{self._sched.shape[1]}")
        S = self._nb.S
        freq = np.zeros((steps, S), np.uint32)
        vmean = np.zeros((steps, S), np.float32)
        if steps:
            self._nb.run(steps, noise_mode=nat.NOISE_SCALAR, noise_rate=float(noise_rate), isolated=isolated,
                         graph=True, freqs=freq, vmeans=vmean)
        if self.src_neuron is not None:
            self.iter += steps
        return self._t(freq.astype(np.float32)), self._t(vmean)

    # state attributes, read from the device on access
    @property
    def active(self):
        return self._t(self._nb.get("active").astype(bool))

    @property
    def V_i(self):
        return self._t(self._nb.get("V"))

    @property
    def J_ui(self):
        return self._t(self._nb.get("J"))

    @property
    def I_ui(self):
        return self._t(self._nb.get("I_ui"))

    @property
    def I_syn(self):
        return self._t(self._nb.get("I_syn"))

    @property
    def t(self):
        return torch.tensor(self._nb.t, dtype=torch.float64, device=self._out_device)

    @property
    def t_ik_last(self):
        return self._t(self._nb.get("t_last"))

    @staticmethod
    def expand(t, size):
        t = torch.tensor(t)
        shape = list(t.shape) + [1] * (len(size) - len(t.shape))
        return t.reshape(shape).expand(size)
