"""ctypes binding of libdtb.so (the C ABI declared in include/dtb.h).

The product path always runs through this library on the GPU; if the
library is missing or no GPU is visible, calls raise instead of falling back
to any CPU code.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libdtb.so")

DTB_OK = 0
DTB_INVALID_PARAMETER = 1
DTB_UNSET_SAMPLE_FAILURE = 2
DTB_UNSUPPORTED = 3
DTB_RUNTIME_ERROR = 4
DTB_OUT_OF_MEMORY = 5
NPROP = 22

NOISE_SCALAR, NOISE_PER_NEURON, NOISE_UNIFORMS = 0, 1, 2

FIELD = dict(V=0, J=1, I_ui=2, I_syn=3, active=4, spike=5, spike_list=6, props=7, rowptr=8, post=9, kind=10,
             delay=11, w0=12, w1=13, t_last=14, pop_sizes=15)

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_f64 = ctypes.c_double


class DTBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[dtb status {code}] {msg}")
        self.code = code


class BlockDesc(ctypes.Structure):
    _fields_ = [("n", _i64), ("nnz", _i64), ("props", _vp), ("rowptr", _vp), ("post", _vp), ("kind", _vp),
                ("delay", _vp), ("w0", _vp), ("w1", _vp), ("dt", _f64), ("noise_seed", _u64),
                ("stats_exclude_blocked", _i32), ("device", _i32)]


class FixedIndegreeSpec(ctypes.Structure):
    _fields_ = [("n", _i64), ("n_exc", _i64), ("in_degree", _i32), ("max_delay", _i32), ("w_lo", _f32),
                ("w_hi", _f32), ("pair_weights", _i32), ("seed", _u64)]


class MemberSpec(ctypes.Structure):
    _fields_ = [("n", _i64), ("n_exc", _i64), ("in_degree", _i32), ("inter_degree", _i32)]


class RunArgs(ctypes.Structure):
    _fields_ = [("noise_mode", _i32), ("noise_rate", _f32), ("uniforms", _vp), ("isolated", _i32),
                ("use_graph", _i32), ("freqs", _vp), ("vmeans", _vp), ("imeans", _vp), ("sample_spike", _vp),
                ("sample_v", _vp), ("events", _vp), ("spikes", _vp), ("bold", _vp), ("kernel_ms", _vp)]


# argument types of every entry point that returns a status code (tests/test_abi.py
# checks them against the prototypes in include/dtb.h)
SIGNATURES = {
    "dtb_block_create": [_vp, _vp],
    "dtb_block_create_fixed_indegree": [_vp, _vp, _vp, _f64, _u64, _i32, _i32],
    "dtb_block_destroy": [_vp],
    "dtb_block_create_members": [_vp, _vp, _i32, _i32, _f32, _f32, _i32, _u64, _vp, _f64, _u64, _i32, _i32],
    "dtb_block_create_regions": [_vp, _vp, _i32, _vp, _i32, _f32, _f32, _i32, _u64, _vp, _f64, _u64, _i32,
                                 _i32],
    "dtb_block_create_region_replicas": [_vp, _vp, _i32, _i32, _vp, _i32, _f32, _f32, _i32, _u64, _vp, _f64,
                                         _u64, _i32, _i32],
    "dtb_block_run": [_vp, _i64, _vp],
    "dtb_block_get": [_vp, _i32, _vp, _i64, _vp],
    "dtb_block_info": [_vp, _vp],
    "dtb_block_time": [_vp, _vp],
    "dtb_block_state_bytes": [_vp, _vp],
    "dtb_block_save_state": [_vp, _vp, _i64],
    "dtb_block_load_state": [_vp, _vp, _i64],
    "dtb_block_set_props": [_vp, _i64, _vp, _vp, _vp],
    "dtb_block_mul_props_by_pop": [_vp, _i64, _vp, _vp, _vp],
    "dtb_block_gamma_props_by_pop": [_vp, _i64, _vp, _vp, _vp, _vp, _u64],
    "dtb_block_replace_props": [_vp, _vp],
    "dtb_block_set_weights": [_vp, _i64, _vp, _vp, _vp, _vp],
    "dtb_block_set_sources": [_vp, _i64, _vp, _vp, _i64, _vp, _vp],
    "dtb_block_set_samples": [_vp, _i64, _vp],
    "dtb_block_attach_bold": [_vp, _vp, _i32, _i32, _vp, _vp],
    "dtb_block_set_launch": [_vp, _i32, _i32],
    "dtb_block_set_external": [_vp, _i32, _vp, _f32],
    "dtb_block_get_launch": [_vp, _i32, _vp],
    "dtb_block_bold_state": [_vp, _i32, _i32, _vp],
    "dtb_bold_step_f32": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dtb_bold_step_f64": [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dtb_bold_run_f64": [_i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "dtb_enkf_analysis_f64": [_i32, _i32, _i32, _vp, _f64, _vp, _f64, _vp, _vp],
    "dtb_enkf_analysis_f64_noalloc": [_i32, _i32, _i32, _vp, _f64, _vp, _f64, _vp, _vp],
    "dtb_mlp_gemm_bf16_nt": [_i32, _i32, _i32, _vp, _i64, _vp, _i64, _vp, _i32, _vp, _i64, _vp, _i64, _vp],
    "dtb_mlp_relu_backward": [_i64, _vp, _vp, _vp, _vp, _vp],
    "dtb_mlp_bias_grad": [_i32, _i32, _vp, _vp, _vp],
    "dtb_mlp_mse_grad": [_i64, _vp, _vp, _vp, _f32, _vp, _vp, _vp, _vp],
    "dtb_mlp_adam": [_i64, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32, _i32, _vp, _vp],
    "dtb_mlp_transpose_bf16": [_i32, _i32, _vp, _vp, _vp],
    "dtb_mlp_f32_to_bf16": [_i64, _vp, _vp, _vp],
    "dtb_mlp_affine": [_i32, _i32, _i32, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp],
    "dtb_gru_forward": [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp],
    "dtb_gru_backward": [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
}


_lib = None


def lib():
    """Load libdtb.so once; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m digital_twin_brain_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    L.dtb_last_error.restype = ctypes.c_char_p
    L.dtb_abi_version.restype = _i32
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def exported_symbols():
    """Names the header declares, for the load test."""
    return ["dtb_last_error", "dtb_abi_version", "dtb_block_create", "dtb_block_create_fixed_indegree",
            "dtb_block_create_members", "dtb_block_create_regions", "dtb_block_create_region_replicas",
            "dtb_block_destroy", "dtb_block_run", "dtb_block_get", "dtb_block_info", "dtb_block_time",
            "dtb_block_state_bytes", "dtb_block_save_state", "dtb_block_load_state",
            "dtb_block_set_props", "dtb_block_mul_props_by_pop", "dtb_block_gamma_props_by_pop",
            "dtb_block_replace_props", "dtb_block_set_weights", "dtb_block_set_sources", "dtb_block_set_samples", "dtb_block_attach_bold", "dtb_block_bold_state",
            "dtb_block_set_launch", "dtb_block_get_launch", "dtb_block_set_external",
            "dtb_bold_step_f32", "dtb_bold_step_f64", "dtb_bold_run_f64",
            "dtb_enkf_analysis_f64", "dtb_enkf_analysis_f64_noalloc",
            "dtb_mlp_gemm_bf16_nt", "dtb_mlp_relu_backward", "dtb_mlp_bias_grad", "dtb_mlp_mse_grad",
            "dtb_mlp_adam", "dtb_mlp_transpose_bf16", "dtb_mlp_f32_to_bf16", "dtb_mlp_affine",
            "dtb_gru_forward", "dtb_gru_backward"]


def check(status):
    if status != DTB_OK:
        msg = lib().dtb_last_error().decode(errors="replace")
        raise DTBError(status, msg)


def ptr(x):
    """Address of a numpy array or torch tensor (None -> NULL)."""
    if x is None:
        return None
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    raise TypeError(f"cannot take the address of {type(x)}")


def require_gpu():
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("digital_twin_brain_amd needs a ROCm GPU (gfx950): torch.cuda.is_available() is False")
