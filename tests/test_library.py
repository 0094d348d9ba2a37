"""CPU-side checks of the C-ABI library: it loads and exports every entry
point include/dtb.h declares (no compute call, no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from digital_twin_brain_amd import _native as nat

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "dtb.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\**\s+\**(dtb_[a-z0-9_]+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert "dtb_block_run" in syms and "dtb_enkf_analysis_f64" in syms
    assert syms == sorted(nat.exported_symbols())


@pytest.mark.skipif(not os.path.exists(nat.LIB_PATH), reason="libdtb.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(nat.LIB_PATH)
    for s in header_symbols():
        assert hasattr(L, s), s
    lib = nat.lib()
    assert lib.dtb_abi_version() == 1
    assert lib.dtb_last_error() is not None


def test_bucket_magic_division_is_exact():
    """The bucketed delivery maps a target neuron to its bucket as
    (post * m) >> s with s = 26 + ceil(log2 d), m = ceil(2^s / d) (the
    host's choice in ensure_bucketed).  q(x) is monotone, so checking both
    sides of every multiple of d proves q(x) == x // d for all x < 2^26,
    for every bucket size d the host can pick (multiples of 64 below 2^15)."""
    for d in range(64, 1 << 15, 64):
        s = 26
        while (1 << (s - 26)) < d:
            s += 1
        m = ((1 << s) + d - 1) // d
        assert m * (1 << 26) < (1 << 64)
        k = np.arange(1, (1 << 26) // d + 1, dtype=np.uint64)
        x = np.concatenate([k * np.uint64(d) - np.uint64(1), k * np.uint64(d)])
        x = x[x < (1 << 26)]
        q = (x * np.uint64(m)) >> np.uint64(s)
        assert np.array_equal(q, x // np.uint64(d)), d
