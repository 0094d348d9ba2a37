"""CPU check of the grouped CSR that persistent delivery reads
(csrc/dtb_group.hpp, the code the library builds it with): compiled with g++
alone and run on random delay-sorted CSRs (tests/group_layout_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "digital_twin_brain_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_grouped_csr_layout(tmp_path):
    exe = str(tmp_path / "group_layout_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Werror", f"-I{CSRC}", "-o", exe,
                    os.path.join(HERE, "group_layout_check.cpp")], check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok 31"
