"""The allocation-free EnKF entry point (dtb_enkf_analysis_f64_noalloc)
against the allocating one (dtb_enkf_analysis_f64), through the C ABI:
identical results on two calls in a row, the second smaller than the
first (the fixed scratch keeps the first call's contents), and a clear
error for a size beyond the fixed scratch."""
import ctypes

import numpy as np
import pytest
import torch

from digital_twin_brain_amd import _native as nat

pytestmark = pytest.mark.gpu


def _analysis(fn, w, y, xi, sigma, rho, stream):
    E, V, S = w.shape
    out = w.clone()
    st = fn(E, V, S, out.data_ptr(), sigma, y.data_ptr(), rho, xi.data_ptr(), stream)
    return st, out


def test_noalloc_enkf_matches_allocating_entry_point():
    L = nat.lib()
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    rng = np.random.default_rng(41)
    for E, V, S in ((256, 90, 10), (64, 30, 6)):
        w = torch.from_numpy(rng.normal(size=(E, V, S)) * 0.1).to(dev)
        y = torch.from_numpy(rng.normal(size=V) * 0.05).to(dev)
        xi = torch.from_numpy(rng.normal(size=(E, V))).to(dev)
        for rho in (0.0, 0.7, 1.0):
            st0, ref = _analysis(L.dtb_enkf_analysis_f64, w, y, xi, 1e-4, rho, stream)
            nat.check(st0)
            st1, got = _analysis(L.dtb_enkf_analysis_f64_noalloc, w, y, xi, 1e-4, rho, stream)
            nat.check(st1)
            torch.cuda.synchronize()
            assert torch.equal(got, ref), f"E={E} V={V} S={S} rho={rho}: results differ"
            assert not torch.equal(got, w)


def test_noalloc_enkf_rejects_sizes_beyond_its_scratch():
    L = nat.lib()
    dev = torch.device("cuda", 0)
    E, V, S = 1024, 90, 10  # 1024 x 90 x 10 > 2^19 doubles of scratch
    w = torch.zeros((E, V, S), dtype=torch.float64, device=dev)
    y = torch.zeros(V, dtype=torch.float64, device=dev)
    xi = torch.zeros((E, V), dtype=torch.float64, device=dev)
    st = L.dtb_enkf_analysis_f64_noalloc(E, V, S, w.data_ptr(), 1e-4, y.data_ptr(), 0.5, xi.data_ptr(), None)
    assert st == nat.DTB_INVALID_PARAMETER
    assert "EnKF" in L.dtb_last_error().decode()
