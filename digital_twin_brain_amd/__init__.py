"""MI355X-native Digital Twin Brain network step.

Public surface (drop-ins for the reference's Python interfaces):
- `block`            models/block.py
- `BOLD`             models/bold_model_pytorch.py
- `diffusion_enkf`   data_assimilation/DataAssimilation.py:40-60
- `NativeBlock`      the C-ABI handle (include/dtb.h) used by all of them
"""
from . import _native
from .block import block, csr_from_w
from .bold import BOLD
from .engine import NativeBlock
from .enkf import diffusion_enkf

__all__ = ["block", "csr_from_w", "BOLD", "diffusion_enkf", "NativeBlock", "_native"]
