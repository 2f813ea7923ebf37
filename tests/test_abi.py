"""The C-ABI boundary (no GPU): libefl.so loads, exports every entry point
include/efl.h declares, and the ctypes mirror matches the C struct layouts."""
import ctypes as C
import os
import re
import struct
import subprocess

import pytest

from bevy_eulerian_fluid_amd import abi, native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "efl.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|size_t|double)\s+(efl_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    assert "efl_step" in names and "efl_ctx_create" in names
    assert set(names) == set(native.EXPORTS), set(names) ^ set(native.EXPORTS)


def test_library_loads_and_exports_every_symbol():
    L = native.load()
    for name in declared_functions():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", native.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (efl_\w+)", out))
    assert set(declared_functions()) <= exported


def _offload_targets(fatbin):
    """Target triples of every clang offload bundle in a .hip_fatbin section
    (each translation unit contributes one bundle: magic, entry count, then
    per entry offset, size, triple length and triple)."""
    magic = b"__CLANG_OFFLOAD_BUNDLE__"
    out = []
    at = fatbin.find(magic)
    while at >= 0:
        p = at + len(magic)
        (n,) = struct.unpack_from("<Q", fatbin, p)
        p += 8
        triples = []
        for _ in range(n):
            _off, _size, tl = struct.unpack_from("<QQQ", fatbin, p)
            p += 24
            triples.append(fatbin[p:p + tl].decode())
            p += tl
        out.append(triples)
        at = fatbin.find(magic, p)
    return out


def test_library_has_gfx950_code_object(tmp_path):
    """libefl.so carries one device code object per HIP source, each for
    gfx950 and nothing else."""
    objcopy = "/opt/rocm/llvm/bin/llvm-objcopy"
    if not os.path.exists(objcopy):
        pytest.skip("llvm-objcopy unavailable")
    sec = tmp_path / "fatbin"
    subprocess.run([objcopy, "--dump-section", f".hip_fatbin={sec}", native.LIB_PATH, str(tmp_path / "lib.so")],
                   check=True)
    bundles = _offload_targets(sec.read_bytes())
    assert len(bundles) == 4, bundles  # efl_kernels, efl_mg, efl_nn, efl_api
    for triples in bundles:
        dev = [t for t in triples if not t.startswith("host-")]
        assert dev == ["hipv4-amdgcn-amd-amdhsa--gfx950"], triples


def test_ctypes_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    lines = ['#include "efl.h"', "#include <stdio.h>", "int main(void) {"]
    for name in abi.STRUCT_SIZES:
        lines.append(f'  printf("{name} %zu\\n", sizeof({name}));')
    lines.append("  return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n") if l)
    for name, size in abi.STRUCT_SIZES.items():
        assert int(got[name]) == size, name
        assert C.sizeof(getattr(abi, name)) == size, name


def test_enums_match_header():
    text = open(HEADER).read()
    body = text[text.index("EFL_FIELD_U0"):text.index("EFL_FIELD_COUNT")]
    fields = re.findall(r"EFL_FIELD_(\w+)", body)
    assert fields == abi.FIELD_NAMES
    body = text[text.index("EFL_STAGE_INIT_GRID_CENTER"):text.index("EFL_STAGE_COUNT")]
    stages = re.findall(r"EFL_STAGE_(\w+)", body)
    assert stages == abi.STAGE_NAMES


def test_no_device_is_reported_as_error_not_crash():
    """Without a GPU the context constructor fails cleanly with a status."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(native.EflError):
        native.Context(0)
