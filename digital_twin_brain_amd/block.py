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
- `src` forced-spike inputs are not supported yet (NotImplementedError).
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
        if src is not None:
            raise NotImplementedError("forced-spike `src` inputs (models/block.py:38-50) are not supported yet")
        assert len(w_uij.shape) == 3
        self.w_uij = w_uij
        self.src = src
        self.src_neuron = None
        self.iter = None
        self.delta_t = delta_t
        self._out_device = w_uij.device
        self._dtype = torch.float32
        csr = csr_from_w(w_uij)
        self._csr = csr
        props = node_property.detach().cpu().numpy().astype(np.float32)
        self._nb = NativeBlock.from_csr(props, csr["rowptr"], csr["post"], csr["kind"], csr["w0"], csr["w1"],
                                        dt=float(delta_t), noise_seed=int(seed), device=device)
        self.N = self._nb.n
        self._set_property_attrs(node_property)
        self._freq = np.zeros(self._nb.S, np.uint32)
        self._vmean = np.zeros(self._nb.S, np.float32)

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
        raise NotImplementedError("update_conn_weight needs the CSR weight-update entry point (not built yet)")

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
        self._nb.run(1, noise_mode=mode, noise_rate=float(noise_rate), uniforms=uniforms, isolated=isolated,
                     graph=False, freqs=self._freq, vmeans=self._vmean)
        return self._t(self._freq.astype(np.float32)), self._t(self._vmean.copy())

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
