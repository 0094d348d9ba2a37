"""CPU-side checks of the C-ABI library: it loads and exports every entry
point include/dtb.h declares (no compute call, no GPU needed)."""
import ctypes
import os
import re

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
