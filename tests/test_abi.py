"""CPU checks that the ctypes binding (digital_twin_brain_amd/_native.py) matches
the C ABI in include/dtb.h: every status-returning prototype's parameters,
one by one (pointer, integer width, float or double), and the size and field
offsets of every struct the binding passes, as gcc lays them out."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from digital_twin_brain_amd import _native as nat

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
HEADER = os.path.join(ROOT, "include", "dtb.h")

# C type -> (kind, bytes) of a by-value parameter
SCALARS = {"int32_t": ("int", 4), "uint32_t": ("int", 4), "int64_t": ("int", 8), "uint64_t": ("int", 8),
           "float": ("float", 4), "double": ("double", 8)}


def _prototypes():
    txt = re.sub(r"/\*.*?\*/", " ", open(HEADER).read(), flags=re.S)
    protos = {}
    for ret, name, params in re.findall(r"\b(int|int32_t|const char\s*\*)\s*(dtb_\w+)\s*\(([^)]*)\)\s*;", txt):
        params = " ".join(params.split())
        protos[name] = [] if params in ("", "void") else [p.strip() for p in params.split(",")]
    return protos


def _kind(param):
    if "*" in param:
        return ("ptr", ctypes.sizeof(ctypes.c_void_p))
    ctype = " ".join(param.replace("const ", "").split()[:-1])
    return SCALARS[ctype]


def _ctypes_kind(t):
    if t is ctypes.c_void_p:
        return ("ptr", ctypes.sizeof(ctypes.c_void_p))
    if t is ctypes.c_float:
        return ("float", 4)
    if t is ctypes.c_double:
        return ("double", 8)
    return ("int", ctypes.sizeof(t))


def test_every_entry_point_has_a_binding_with_the_header_parameters():
    protos = _prototypes()
    assert sorted(protos) == sorted(nat.exported_symbols())
    status = {k for k in protos if k not in ("dtb_last_error", "dtb_abi_version")}
    assert status == set(nat.SIGNATURES), sorted(status ^ set(nat.SIGNATURES))
    for name in sorted(status):
        want = [_kind(p) for p in protos[name]]
        got = [_ctypes_kind(t) for t in nat.SIGNATURES[name]]
        assert got == want, f"{name}: binding {got} against header {want}"


STRUCTS = {"dtb_block_desc": nat.BlockDesc, "dtb_fixed_indegree_spec": nat.FixedIndegreeSpec,
           "dtb_member_spec": nat.MemberSpec, "dtb_run_args": nat.RunArgs}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_struct_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "dtb.h"', "int main(void) {"]
    for cname, cls in STRUCTS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in cls._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = str(tmp_path / "layout")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o",
                    exe], check=True, capture_output=True, text=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, cls in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for field, _ in cls._fields_:
            assert got[(cname, field)] == getattr(cls, field).offset, f"{cname}.{field}"
