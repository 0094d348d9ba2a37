"""NativeBlock: a Python handle over one libdtb block (include/dtb.h).

Arrays passed in may be numpy arrays (host) or torch tensors (host or
device); outputs are written into numpy arrays or torch tensors the caller
provides, or into freshly allocated ones.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native as nat


def _as_np(x, dtype):
    if x is None:
        return None
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(x, dtype=dtype)


def poisson_thresholds(lam, kmax=16):
    """thr[j] = floor(2^32 P(X <= j)) for X ~ Poisson(lam), j < kmax (uint32):
    a 32-bit uniform r gives the count #{j : r >= thr[j]}."""
    lam = float(lam)
    p = math.exp(-lam)
    cdf = 0.0
    thr = np.empty(kmax, np.uint32)
    for j in range(kmax):
        cdf += p
        thr[j] = min(int(cdf * 4294967296.0), 0xFFFFFFFF)
        p *= lam / (j + 1)
    return thr


class NativeBlock:
    """One simulated network on one GPU (handle of dtb_block_*)."""

    def __init__(self, handle, keep=None):
        self._h = ctypes.c_void_p(handle)
        self._keep = keep
        info = self.info()
        self.n, self.nnz, self.S, self.max_delay = info[0], info[1], info[2], info[3]
        self.pair = bool(info[6])
        self.device = int(info[7])

    # ------------------------------------------------------------ creation
    @classmethod
    def from_csr(cls, props, rowptr, post, kind, w0, w1=None, delay=None, dt=0.1, noise_seed=0,
                 stats_exclude_blocked=False, device=0):
        """Network from presynaptic-major CSR (cuda/src/configuration.cpp:747-828)."""
        nat.require_gpu()
        props = _as_np(props, np.float32)
        if props.ndim != 2 or props.shape[1] < nat.NPROP:
            raise ValueError(f"props must be [n, {nat.NPROP}], got {props.shape}")
        props = np.ascontiguousarray(props[:, :nat.NPROP])
        n = props.shape[0]
        rowptr = _as_np(rowptr, np.int64)
        post = _as_np(post, np.int32)
        kind = _as_np(kind, np.uint8)
        w0 = _as_np(w0, np.float32)
        w1 = _as_np(w1, np.float32)
        delay = _as_np(delay, np.uint8)
        if rowptr.shape != (n + 1,):
            raise ValueError("rowptr must have n+1 entries")
        nnz = int(post.shape[0])
        for name, a in (("kind", kind), ("w0", w0), ("w1", w1), ("delay", delay)):
            if a is not None and a.shape[0] != nnz:
                raise ValueError(f"{name} must have nnz={nnz} entries")
        d = nat.BlockDesc(n=n, nnz=nnz, props=nat.ptr(props), rowptr=nat.ptr(rowptr), post=nat.ptr(post),
                          kind=nat.ptr(kind), delay=nat.ptr(delay), w0=nat.ptr(w0), w1=nat.ptr(w1), dt=float(dt),
                          noise_seed=int(noise_seed), stats_exclude_blocked=int(bool(stats_exclude_blocked)),
                          device=int(device))
        h = ctypes.c_void_p()
        nat.check(nat.lib().dtb_block_create(ctypes.byref(h), ctypes.byref(d)))
        return cls(h.value)

    @classmethod
    def fixed_indegree(cls, props, n_exc, in_degree, max_delay=1, w_lo=0.0, w_hi=1.0, pair_weights=False,
                       conn_seed=1, dt=0.1, noise_seed=3, stats_exclude_blocked=False, device=0):
        """Seeded fixed in-degree network generated on the GPU."""
        nat.require_gpu()
        props = _as_np(props, np.float32)
        props = np.ascontiguousarray(props[:, :nat.NPROP])
        spec = nat.FixedIndegreeSpec(n=props.shape[0], n_exc=int(n_exc), in_degree=int(in_degree),
                                     max_delay=int(max_delay), w_lo=float(w_lo), w_hi=float(w_hi),
                                     pair_weights=int(bool(pair_weights)), seed=int(conn_seed))
        h = ctypes.c_void_p()
        nat.check(nat.lib().dtb_block_create_fixed_indegree(ctypes.byref(h), ctypes.byref(spec), nat.ptr(props),
                                                            float(dt), int(noise_seed),
                                                            int(bool(stats_exclude_blocked)), int(device)))
        return cls(h.value)

    @classmethod
    def members(cls, props, members, max_delay=1, w_lo=0.0, w_hi=1.0, pair_weights=False, conn_seed=1, dt=0.1,
                noise_seed=3, stats_exclude_blocked=False, device=0):
        """Uneven members packed in one block; members = [(n, n_exc, in_degree), ...]."""
        nat.require_gpu()
        props = np.ascontiguousarray(_as_np(props, np.float32)[:, :nat.NPROP])
        arr = (nat.MemberSpec * len(members))(*[nat.MemberSpec(int(n), int(ne), int(k), 0) for n, ne, k in members])
        h = ctypes.c_void_p()
        nat.check(nat.lib().dtb_block_create_members(ctypes.byref(h), arr, len(members), int(max_delay), float(w_lo),
                                                     float(w_hi), int(bool(pair_weights)), int(conn_seed),
                                                     nat.ptr(props), float(dt), int(noise_seed),
                                                     int(bool(stats_exclude_blocked)), int(device)))
        return cls(h.value)

    @classmethod
    def regions(cls, props, regions, source_cdf, max_delay=1, w_lo=0.0, w_hi=1.0, pair_weights=False, conn_seed=1,
                dt=0.1, noise_seed=3, stats_exclude_blocked=False, device=0):
        """Regions with inter-region synapses; regions = [(n, n_exc, k_intra, k_inter), ...],
        source_cdf = uint64 [R, R] cumulative source probabilities scaled to 2^32."""
        nat.require_gpu()
        props = np.ascontiguousarray(_as_np(props, np.float32)[:, :nat.NPROP])
        cdf = np.ascontiguousarray(source_cdf, dtype=np.uint64)
        if cdf.shape != (len(regions), len(regions)):
            raise ValueError("source_cdf must be [regions, regions]")
        arr = (nat.MemberSpec * len(regions))(*[nat.MemberSpec(int(n), int(ne), int(k), int(kx))
                                                for n, ne, k, kx in regions])
        h = ctypes.c_void_p()
        nat.check(nat.lib().dtb_block_create_regions(ctypes.byref(h), arr, len(regions), nat.ptr(cdf),
                                                     int(max_delay), float(w_lo), float(w_hi),
                                                     int(bool(pair_weights)), int(conn_seed), nat.ptr(props),
                                                     float(dt), int(noise_seed), int(bool(stats_exclude_blocked)),
                                                     int(device)))
        return cls(h.value)

    @classmethod
    def region_replicas(cls, props, regions, replicas, source_cdf, max_delay=1, w_lo=0.0, w_hi=1.0,
                        pair_weights=False, conn_seed=1, dt=0.1, noise_seed=3, stats_exclude_blocked=False, device=0):
        """`replicas` copies of one region network back to back, each drawing its
        inter-region sources through the same source_cdf [R, R] inside itself;
        props holds the rows of all replicas."""
        nat.require_gpu()
        props = np.ascontiguousarray(_as_np(props, np.float32)[:, :nat.NPROP])
        cdf = np.ascontiguousarray(source_cdf, dtype=np.uint64)
        if cdf.shape != (len(regions), len(regions)):
            raise ValueError("source_cdf must be [regions, regions]")
        arr = (nat.MemberSpec * len(regions))(*[nat.MemberSpec(int(n), int(ne), int(k), int(kx))
                                                for n, ne, k, kx in regions])
        h = ctypes.c_void_p()
        nat.check(nat.lib().dtb_block_create_region_replicas(ctypes.byref(h), arr, len(regions), int(replicas),
                                                             nat.ptr(cdf), int(max_delay), float(w_lo), float(w_hi),
                                                             int(bool(pair_weights)), int(conn_seed),
                                                             nat.ptr(props), float(dt), int(noise_seed),
                                                             int(bool(stats_exclude_blocked)), int(device)))
        return cls(h.value)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            nat.lib().dtb_block_destroy(self._h)
            self._h = ctypes.c_void_p(None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- info
    def info(self):
        out = np.zeros(9, dtype=np.int64)
        nat.check(nat.lib().dtb_block_info(self._h, nat.ptr(out)))
        return [int(x) for x in out]

    @property
    def steps(self):
        return self.info()[4]

    @property
    def t(self):
        out = np.zeros(1, dtype=np.float64)
        nat.check(nat.lib().dtb_block_time(self._h, nat.ptr(out)))
        return float(out[0])

    # ------------------------------------------------------------ snapshot
    def state_bytes(self):
        out = ctypes.c_int64(0)
        nat.check(nat.lib().dtb_block_state_bytes(self._h, ctypes.byref(out)))
        return out.value

    def save_state(self, out=None):
        """Snapshot of the dynamic state (V, J, arrivals, refractory ends,
        t_last, the spike lists of the delays in flight, step counter, time,
        noise seed, BOLD state) into `out` (uint8 numpy array or torch tensor,
        host or device, at least state_bytes() long); returns it."""
        nbytes = self.state_bytes()
        if out is None:
            out = np.zeros(nbytes, dtype=np.uint8)
        have = out.nbytes if isinstance(out, np.ndarray) else out.numel() * out.element_size()
        nat.check(nat.lib().dtb_block_save_state(self._h, nat.ptr(out), int(have)))
        return out

    def load_state(self, src):
        """Restore a snapshot taken by save_state on a block of the same shape;
        the block then continues exactly as the saved one did."""
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src)
        have = src.nbytes if isinstance(src, np.ndarray) else src.numel() * src.element_size()
        nat.check(nat.lib().dtb_block_load_state(self._h, nat.ptr(src), int(have)))

    # ----------------------------------------------------------------- run
    def run(self, nsteps, noise_mode=nat.NOISE_SCALAR, noise_rate=0.0, uniforms=None, isolated=False, graph=True,
            freqs=None, vmeans=None, imeans=None, sample_spike=None, sample_v=None, events=None, spikes=None,
            bold=None, kernel_ms=None):
        """Advance nsteps; output buffers (numpy or torch, host or device) are filled in place."""
        if uniforms is not None and isinstance(uniforms, np.ndarray):
            uniforms = np.ascontiguousarray(uniforms, dtype=np.float32)
        a = nat.RunArgs(noise_mode=int(noise_mode), noise_rate=float(noise_rate), uniforms=nat.ptr(uniforms),
                        isolated=int(bool(isolated)), use_graph=int(bool(graph)), freqs=nat.ptr(freqs),
                        vmeans=nat.ptr(vmeans), imeans=nat.ptr(imeans), sample_spike=nat.ptr(sample_spike),
                        sample_v=nat.ptr(sample_v), events=nat.ptr(events), spikes=nat.ptr(spikes),
                        bold=nat.ptr(bold), kernel_ms=nat.ptr(kernel_ms))
        nat.check(nat.lib().dtb_block_run(self._h, int(nsteps), ctypes.byref(a)))

    # ---------------------------------------------------------------- BOLD
    BOLD_FIELDS = dict(s=0, q=1, v=2, f_in=3, bold=4, u=5)

    def attach_bold(self, epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02,
                    delta_t=1e-3, steps_per_sample=1, voxel_of_pop=None, voxel_size=None, nvox=None):
        """Balloon-Windkessel per voxel fed by its spike counts (dtb_block_attach_bold)."""
        par = np.array([epsilon, tao_s, tao_f, tao_0, alpha, E_0, V_0, delta_t], dtype=np.float64)
        vop = None if voxel_of_pop is None else np.ascontiguousarray(voxel_of_pop, dtype=np.int32)
        vsz = None if voxel_size is None else np.ascontiguousarray(voxel_size, dtype=np.float64)
        if vop is not None and nvox is None:
            nvox = int(vop.max()) + 1
        nat.check(nat.lib().dtb_block_attach_bold(self._h, nat.ptr(par), int(steps_per_sample),
                                                  int(nvox or self.S), nat.ptr(vop), nat.ptr(vsz)))
        self.bold_sps = int(steps_per_sample)
        self.nvox = int(nvox or self.S)

    def bold_state(self, which):
        out = np.zeros(self.nvox, dtype=np.float64)
        nat.check(nat.lib().dtb_block_bold_state(self._h, self.BOLD_FIELDS[which], 0, nat.ptr(out)))
        return out

    def set_bold_state(self, which, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        if v.shape[0] != self.nvox:
            raise ValueError(f"need {self.nvox} values")
        nat.check(nat.lib().dtb_block_bold_state(self._h, self.BOLD_FIELDS[which], 1, nat.ptr(v)))

    # ---------------------------------------------------- external drive
    def set_external(self, rate_per_step, weight, kmax=16):
        """Poisson(rate_per_step) external spikes per neuron and step, each adding
        `weight` to the excitatory channels (rate 0 turns the drive off)."""
        thr = poisson_thresholds(rate_per_step, kmax) if rate_per_step > 0 else np.zeros(0, np.uint32)
        nat.check(nat.lib().dtb_block_set_external(self._h, int(thr.shape[0]), nat.ptr(thr), float(weight)))
        return thr

    # -------------------------------------------------------- launch shape
    LAUNCH = dict(deliver_grid=0, deliver_lanes=1, deliver_mode=2, neuron_psets=3, stage_segs=4, stage_ept=5,
                  stage_variant=6, accum_occ=7, accum_unroll=8, accum_fixed=9, segment_lists=10)

    def set_launch(self, **kw):
        for k, v in kw.items():
            nat.check(nat.lib().dtb_block_set_launch(self._h, self.LAUNCH[k], int(v)))

    def get_launch(self):
        out = {}
        for k, w in self.LAUNCH.items():
            v = ctypes.c_int32(0)
            nat.check(nat.lib().dtb_block_get_launch(self._h, w, ctypes.byref(v)))
            out[k] = v.value
        return out

    # ----------------------------------------------------------------- get
    def get(self, name, out=None):
        f = nat.FIELD[name]
        n, nnz, S = self.n, self.nnz, self.S
        shapes = dict(V=((n,), np.float32), J=((4, n), np.float32), I_ui=((4, n), np.float32),
                      I_syn=((n,), np.float32), active=((n,), np.uint8), spike=((n,), np.uint8),
                      props=((n, nat.NPROP), np.float32), rowptr=((n + 1,), np.int64), post=((nnz,), np.int32),
                      kind=((nnz,), np.uint8), delay=((nnz,), np.uint8), w0=((nnz,), np.float32),
                      w1=((nnz,), np.float32), t_last=((n,), np.float64), pop_sizes=((S,), np.int64))
        if name == "spike_list":
            nb = ctypes.c_int64(0)
            st = nat.lib().dtb_block_get(self._h, f, None, 0, ctypes.byref(nb))
            if st not in (nat.DTB_OK, nat.DTB_INVALID_PARAMETER):
                nat.check(st)
            buf = np.zeros(max(nb.value // 4, 0), dtype=np.int32)
            if buf.size:
                nat.check(nat.lib().dtb_block_get(self._h, f, nat.ptr(buf), buf.nbytes, None))
            return buf
        shape, dt = shapes[name]
        if out is None:
            out = np.zeros(shape, dtype=dt)
        nbytes = out.nbytes if isinstance(out, np.ndarray) else out.numel() * out.element_size()
        nat.check(nat.lib().dtb_block_get(self._h, f, nat.ptr(out), int(nbytes), None))
        return out

    # ---------------------------------------------------------- properties
    def set_props(self, neuron, col, val):
        neuron = _as_np(neuron, np.uint32)
        col = _as_np(col, np.uint32)
        val = _as_np(val, np.float32)
        nat.check(nat.lib().dtb_block_set_props(self._h, int(neuron.shape[0]), nat.ptr(neuron), nat.ptr(col),
                                                nat.ptr(val)))

    def mul_props_by_pop(self, pop, col, val):
        pop = _as_np(pop, np.uint32)
        col = _as_np(col, np.uint32)
        val = _as_np(val, np.float32)
        nat.check(nat.lib().dtb_block_mul_props_by_pop(self._h, int(pop.shape[0]), nat.ptr(pop), nat.ptr(col),
                                                       nat.ptr(val)))

    def gamma_props_by_pop(self, pop, col, alpha, beta, seed=0):
        pop = _as_np(pop, np.uint32)
        col = _as_np(col, np.uint32)
        alpha = _as_np(alpha, np.float32)
        beta = _as_np(beta, np.float32)
        nat.check(nat.lib().dtb_block_gamma_props_by_pop(self._h, int(pop.shape[0]), nat.ptr(pop), nat.ptr(col),
                                                         nat.ptr(alpha), nat.ptr(beta), int(seed)))

    def replace_props(self, props):
        props = np.ascontiguousarray(_as_np(props, np.float32)[:, :nat.NPROP])
        if props.shape != (self.n, nat.NPROP):
            raise ValueError(f"props must be [{self.n}, {nat.NPROP}]")
        nat.check(nat.lib().dtb_block_replace_props(self._h, nat.ptr(props)))

    def set_weights(self, pre, post, channel, w):
        """w_uij[channel][post][pre] = w for existing synapses (channel 0..3:
        AMPA, NMDA, GABAa, GABAb); a compact block switches to pair weights
        when a call sets the two channels of a pair differently."""
        pre = _as_np(pre, np.uint32).reshape(-1)
        post = _as_np(post, np.uint32).reshape(-1)
        channel = _as_np(channel, np.uint8).reshape(-1)
        w = _as_np(w, np.float32).reshape(-1)
        m = pre.shape[0]
        if not (post.shape[0] == channel.shape[0] == w.shape[0] == m):
            raise ValueError("pre, post, channel and w must have the same length")
        nat.check(nat.lib().dtb_block_set_weights(self._h, int(m), nat.ptr(pre), nat.ptr(post), nat.ptr(channel),
                                                  nat.ptr(w)))
        self.pair = bool(self.info()[6])

    def set_sources(self, sources, steps=(), neurons=(), frozen=None):
        """Forced-spike sources: neurons[k] fires at the step whose 0-based
        index since creation is steps[k]; sources fire at no other step.
        frozen (bool per source, or one bool for all) keeps a source's V and
        turns its noise off (the neurons of an input block)."""
        src = _as_np(sources, np.uint32).reshape(-1)
        fr = None
        if frozen is not None:
            fr = np.ascontiguousarray(np.broadcast_to(np.asarray(frozen, dtype=np.uint8), src.shape))
        st = _as_np(np.asarray(steps), np.uint32).reshape(-1)
        nu = _as_np(np.asarray(neurons), np.uint32).reshape(-1)
        if st.shape != nu.shape:
            raise ValueError("steps and neurons must have the same length")
        nat.check(nat.lib().dtb_block_set_sources(self._h, int(src.shape[0]), nat.ptr(src), nat.ptr(fr),
                                                  int(st.shape[0]), nat.ptr(st), nat.ptr(nu)))

    def set_samples(self, idx):
        idx = _as_np(idx, np.uint32)
        nat.check(nat.lib().dtb_block_set_samples(self._h, int(idx.shape[0]), nat.ptr(idx)))
        self.n_samples = int(idx.shape[0])
