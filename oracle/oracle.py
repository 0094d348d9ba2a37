"""ctypes wrapper of the CPU oracle (oracle/dtb_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports this module.

Each function names the reference file:line it restates; the C source holds
the arithmetic.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_c_i64 = ctypes.c_int64
_c_i32 = ctypes.c_int32
_c_u64 = ctypes.c_uint64
_c_f32 = ctypes.c_float
_c_f64 = ctypes.c_double
_vp = ctypes.c_void_p


def build():
    """Compile the oracle shared libraries with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _lib(real="f64"):
    if real in _LIBS:
        return _LIBS[real]
    path = os.path.join(_HERE, f"libdtb_oracle_{real}.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    sfx = "_" + real
    lib.orc_philox4x32_10.argtypes = [_vp, _vp, _vp]
    lib.orc_noise_uniform.argtypes = [_c_u64, _c_u64, _c_u64]
    lib.orc_noise_uniform.restype = _c_f32
    lib.orc_noise_uniforms.argtypes = [_c_u64, _c_u64, _c_i64, _vp]
    lib.orc_gen_fixed_indegree.argtypes = [_c_i64, _c_i64, _c_i32, _c_i32, _c_f32, _c_f32, _c_i32, _c_u64,
                                           _vp, _vp, _vp, _vp, _vp, _vp]
    lib.orc_gen_members.argtypes = [_c_i32, _vp, _vp, _vp, _c_i32, _c_f32, _c_f32, _c_i32, _c_u64,
                                    _vp, _vp, _vp, _vp, _vp, _vp]
    lib.orc_gen_regions.argtypes = [_c_i32, _vp, _vp, _vp, _vp, _vp, _c_i32, _c_f32, _c_f32, _c_i32, _c_u64,
                                    _vp, _vp, _vp, _vp, _vp, _vp]
    f = getattr(lib, "orc_block_create" + sfx)
    f.argtypes = [_c_i64, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_f64, _c_i32]
    f.restype = _vp
    getattr(lib, "orc_block_destroy" + sfx).argtypes = [_vp]
    getattr(lib, "orc_block_step" + sfx).argtypes = [_vp, _c_i32, _c_f32, _vp, _c_u64, _c_i32, _c_i32, _vp, _vp]
    getattr(lib, "orc_block_num_sub" + sfx).argtypes = [_vp]
    getattr(lib, "orc_block_num_sub" + sfx).restype = _c_i32
    getattr(lib, "orc_block_get" + sfx).argtypes = [_vp, _c_i32, _vp]
    getattr(lib, "orc_block_set" + sfx).argtypes = [_vp, _c_i32, _vp]
    getattr(lib, "orc_block_set_sources" + sfx).argtypes = [_vp, _c_i64, _vp, _vp, _c_i64, _vp, _vp]
    getattr(lib, "orc_block_replace_csr" + sfx).argtypes = [_vp, _c_i64, _vp, _vp, _vp, _vp, _vp]
    getattr(lib, "orc_block_set_external" + sfx).argtypes = [_vp, _c_i32, _vp, _c_f32]
    getattr(lib, "orc_ext_count" + sfx).argtypes = [_c_u64, _c_u64, _c_u64, _c_i32, _vp]
    getattr(lib, "orc_ext_count" + sfx).restype = ctypes.c_uint32
    lib.orc_compact.argtypes = [_vp, _c_i64, _vp]
    lib.orc_compact.restype = _c_i64
    lib.orc_bold_step.argtypes = [_c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    lib.orc_diffusion_enkf.argtypes = [_c_i32, _c_i32, _c_i32, _vp, _c_f64, _vp, _c_f64, _vp]
    lib.orc_omp_threads.restype = _c_i32
    _LIBS[real] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def philox4x32_10(ctr, key):
    """Philox4x32-10 block function (Salmon et al., SC'11)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    _lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def noise_uniforms(seed, step, n):
    """The n noise uniforms of step `step` (1-based), float32 in [0,1)."""
    out = np.empty(n, dtype=np.float32)
    _lib().orc_noise_uniforms(seed, step, n, _p(out))
    return out


def gen_fixed_indegree(n, n_exc, in_degree, max_delay, w_lo, w_hi, pair, seed):
    """Seeded fixed in-degree network as presynaptic-major CSR.

    Layout of cuda/src/configuration.cpp:747-828 (rows = presynaptic
    neurons); receptor pair from the presynaptic type as in
    generation/make_block.py:290-291.  Returns a dict of numpy arrays.
    """
    nnz = n * in_degree
    rowptr = np.zeros(n + 1, dtype=np.int64)
    post = np.zeros(nnz, dtype=np.int32)
    kind = np.zeros(nnz, dtype=np.uint8)
    delay = np.zeros(nnz, dtype=np.uint8)
    w0 = np.zeros(nnz, dtype=np.float32)
    w1 = np.zeros(nnz, dtype=np.float32)
    _lib().orc_gen_fixed_indegree(n, n_exc, in_degree, max_delay, w_lo, w_hi, int(pair), seed,
                                  _p(rowptr), _p(post), _p(kind), _p(delay), _p(w0), _p(w1))
    return dict(rowptr=rowptr, post=post, kind=kind, delay=delay, w0=w0, w1=w1 if pair else None)


def gen_members(members, max_delay, w_lo, w_hi, pair, seed):
    """Uneven members packed in one network; members = [(n, n_exc, in_degree), ...]."""
    ns = np.array([m[0] for m in members], np.int64)
    nexc = np.array([m[1] for m in members], np.int64)
    ks = np.array([m[2] for m in members], np.int32)
    n = int(ns.sum())
    nnz = int((ns * ks).sum())
    rowptr = np.zeros(n + 1, dtype=np.int64)
    post = np.zeros(nnz, dtype=np.int32)
    kind = np.zeros(nnz, dtype=np.uint8)
    delay = np.zeros(nnz, dtype=np.uint8)
    w0 = np.zeros(nnz, dtype=np.float32)
    w1 = np.zeros(nnz, dtype=np.float32)
    _lib().orc_gen_members(len(members), _p(ns), _p(nexc), _p(ks), max_delay, w_lo, w_hi, int(pair), seed,
                           _p(rowptr), _p(post), _p(kind), _p(delay), _p(w0), _p(w1))
    return dict(rowptr=rowptr, post=post, kind=kind, delay=delay, w0=w0, w1=w1 if pair else None)


def gen_regions(regions, cdf, max_delay, w_lo, w_hi, pair, seed):
    """Regions with a source-region table; regions = [(n, n_exc, k_intra, k_inter), ...],
    cdf = uint64 [R, R] cumulative source probabilities scaled to 2^32."""
    ns = np.array([r[0] for r in regions], np.int64)
    nexc = np.array([r[1] for r in regions], np.int64)
    ks = np.array([r[2] for r in regions], np.int32)
    kx = np.array([r[3] for r in regions], np.int32)
    cdf = np.ascontiguousarray(cdf, dtype=np.uint64)
    assert cdf.shape == (len(regions), len(regions))
    n = int(ns.sum())
    nnz = int((ns * (ks + kx)).sum())
    rowptr = np.zeros(n + 1, dtype=np.int64)
    post = np.zeros(nnz, dtype=np.int32)
    kind = np.zeros(nnz, dtype=np.uint8)
    delay = np.zeros(nnz, dtype=np.uint8)
    w0 = np.zeros(nnz, dtype=np.float32)
    w1 = np.zeros(nnz, dtype=np.float32)
    _lib().orc_gen_regions(len(regions), _p(ns), _p(nexc), _p(ks), _p(kx), _p(cdf), max_delay, w_lo, w_hi,
                           int(pair), seed, _p(rowptr), _p(post), _p(kind), _p(delay), _p(w0), _p(w1))
    return dict(rowptr=rowptr, post=post, kind=kind, delay=delay, w0=w0, w1=w1 if pair else None)


def ext_count(seed, step, i, thresholds):
    """External-drive count of neuron i at `step`: #{j : r >= thresholds[j]}."""
    thr = np.ascontiguousarray(thresholds, dtype=np.uint32)
    return int(_lib().orc_ext_count_f64(seed, step, i, thr.shape[0], _p(thr)))


class OracleBlock:
    """CPU restatement of models/block.py:5-177 (block.__init__ / run)."""

    FIELDS = {"V": 0, "J": 1, "I_ui": 2, "I_syn": 3, "t_last": 4, "active": 5, "spike": 6}

    def __init__(self, props, csr, dt, real="f64", pull=False):
        self.real = real
        self.lib = _lib(real)
        self.dtype = np.float64 if real == "f64" else np.float32
        self.sfx = "_" + real
        props = np.ascontiguousarray(props, dtype=np.float32)
        self.n = props.shape[0]
        self._keep = dict(props=props)
        rowptr = np.ascontiguousarray(csr["rowptr"], dtype=np.int64)
        post = np.ascontiguousarray(csr["post"], dtype=np.int32)
        kind = np.ascontiguousarray(csr["kind"], dtype=np.uint8)
        delay = csr.get("delay")
        delay = None if delay is None else np.ascontiguousarray(delay, dtype=np.uint8)
        w0 = np.ascontiguousarray(csr["w0"], dtype=np.float32)
        w1 = csr.get("w1")
        w1 = None if w1 is None else np.ascontiguousarray(w1, dtype=np.float32)
        self.nnz = post.shape[0]
        self.pull = bool(pull)
        self.h = getattr(self.lib, "orc_block_create" + self.sfx)(
            self.n, _p(props), self.nnz, _p(rowptr), _p(post), _p(kind), _p(delay), _p(w0), _p(w1),
            float(dt), int(self.pull))
        self.S = getattr(self.lib, "orc_block_num_sub" + self.sfx)(self.h)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            getattr(self.lib, "orc_block_destroy" + self.sfx)(h)
            self.h = None

    def run(self, noise_rate=0.01, isolated=False, seed=0, uniforms=None, noise_mode=0):
        """One step; returns (sum_activate[S], mean_Vi[S]) like models/block.py:150."""
        s = np.zeros(self.S, dtype=self.dtype)
        m = np.zeros(self.S, dtype=self.dtype)
        u = None if uniforms is None else np.ascontiguousarray(uniforms, dtype=np.float32)
        getattr(self.lib, "orc_block_step" + self.sfx)(self.h, int(noise_mode), float(noise_rate), _p(u), seed,
                                                       int(bool(isolated)), int(self.pull), _p(s), _p(m))
        return s, m

    def set_external(self, thresholds, weight):
        """External Poisson drive, as dtb_block_set_external."""
        thr = np.ascontiguousarray(thresholds, dtype=np.uint32)
        self._keep["ext"] = thr
        getattr(self.lib, "orc_block_set_external" + self.sfx)(self.h, int(thr.shape[0]), _p(thr), float(weight))

    def get(self, name):
        f = self.FIELDS[name]
        if name in ("active", "spike"):
            out = np.zeros(self.n, dtype=np.uint8)
        elif name in ("J", "I_ui"):
            out = np.zeros((4, self.n), dtype=self.dtype)
        else:
            out = np.zeros(self.n, dtype=self.dtype)
        getattr(self.lib, "orc_block_get" + self.sfx)(self.h, f, _p(out))
        return out

    def replace_csr(self, csr):
        """New connectivity (delay 1, push form), state kept: the reference's
        w_uij assignment between steps (models/block.py:172-177)."""
        assert not self.pull
        rowptr = np.ascontiguousarray(csr["rowptr"], dtype=np.int64)
        assert rowptr.shape == (self.n + 1,)
        post = np.ascontiguousarray(csr["post"], dtype=np.int32)
        kind = np.ascontiguousarray(csr["kind"], dtype=np.uint8)
        w0 = np.ascontiguousarray(csr["w0"], dtype=np.float32)
        w1 = csr.get("w1")
        w1 = None if w1 is None else np.ascontiguousarray(w1, dtype=np.float32)
        self.nnz = post.shape[0]
        getattr(self.lib, "orc_block_replace_csr" + self.sfx)(self.h, self.nnz, _p(rowptr), _p(post), _p(kind),
                                                               _p(w0), _p(w1))

    def set_sources(self, sources, steps=(), neurons=(), frozen=None):
        """Forced-spike sources (models/block.py:38-50,104-109): neurons[k]
        fires at the step whose 0-based index is steps[k]; frozen sources
        (input-block neurons, cuda/src/block.cpp:195-221) keep V and draw no
        noise."""
        src = np.ascontiguousarray(sources, dtype=np.uint32)
        fr = None if frozen is None else np.ascontiguousarray(np.broadcast_to(frozen, src.shape), dtype=np.uint8)
        st = np.ascontiguousarray(steps, dtype=np.uint32)
        nu = np.ascontiguousarray(neurons, dtype=np.uint32)
        assert st.shape == nu.shape
        self._keep["src"] = (src, fr, st, nu)
        getattr(self.lib, "orc_block_set_sources" + self.sfx)(self.h, src.shape[0], _p(src), _p(fr), st.shape[0],
                                                               _p(st), _p(nu))

    def set(self, name, values):
        """Overwrite V, J ([4, n]) or t_last between steps; I_ui and I_syn are
        recomputed from V and J (test infrastructure: re-aligning the oracle
        with the GPU state after a threshold tie)."""
        if name not in ("V", "J", "t_last"):
            raise ValueError("only V, J and t_last can be set")
        shape = (4, self.n) if name == "J" else (self.n,)
        v = np.ascontiguousarray(np.asarray(values, dtype=self.dtype).reshape(shape))
        getattr(self.lib, "orc_block_set" + self.sfx)(self.h, self.FIELDS[name], _p(v))

    @property
    def t(self):
        out = np.zeros(1, dtype=self.dtype)
        getattr(self.lib, "orc_block_get" + self.sfx)(self.h, 7, _p(out))
        return float(out[0])


def compact(flags):
    """Ascending positions of set flags (cuda/src/block.cpp:1033-1055)."""
    f = np.ascontiguousarray(flags, dtype=np.uint8)
    out = np.zeros(f.shape[0], dtype=np.int32)
    k = _lib().orc_compact(_p(f), f.shape[0], _p(out))
    return out[:k]


BOLD_DEFAULT = dict(epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02)


class OracleBOLD:
    """CPU restatement of models/bold_model_pytorch.py:6-78 (FP64)."""

    def __init__(self, n, epsilon=200, tao_s=0.8, tao_f=0.4, tao_0=1, alpha=0.2, E_0=0.8, V_0=0.02,
                 delta_t=1e-3, s=None, q=None, v=None, f_in=None):
        self.par = np.array([epsilon, tao_s, tao_f, tao_0, alpha, E_0, V_0, delta_t], dtype=np.float64)
        self.s = np.zeros(n) if s is None else np.array(s, dtype=np.float64)
        self.q = np.zeros(n) if q is None else np.array(q, dtype=np.float64)
        self.v = np.ones(n) if v is None else np.array(v, dtype=np.float64)
        self.f_in = np.ones(n) if f_in is None else np.array(f_in, dtype=np.float64)
        self.n = n

    def run(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(self.n)
        _lib().orc_bold_step(self.n, _p(self.par), _p(self.s), _p(self.q), _p(self.v), _p(self.f_in),
                             _p(u), _p(out))
        return out


def diffusion_enkf(w_hat, bold_sigma, bold_t, solo_rate, xi):
    """data_assimilation/DataAssimilation.py:40-60 with the noise xi given."""
    w = np.ascontiguousarray(w_hat, dtype=np.float64).copy()
    E, V, S = w.shape
    y = np.ascontiguousarray(np.reshape(bold_t, -1), dtype=np.float64)
    x = np.ascontiguousarray(xi, dtype=np.float64)
    _lib().orc_diffusion_enkf(E, V, S, _p(w), float(bold_sigma), _p(y), float(solo_rate), _p(x))
    return w


def omp_threads(real="f32"):
    return _lib(real).orc_omp_threads()
