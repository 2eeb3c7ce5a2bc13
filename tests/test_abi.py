"""The C-ABI library loads and exports every entry point include/olamgpu.h declares
(no compute calls: these run without a GPU)."""

import ctypes
import os
import re

from olam_amd import build, gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "olamgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(olamgpu_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert declared() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(build.LIB), "run __graft_entry__.build() first"
    lib = ctypes.CDLL(build.LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_last_error_without_context():
    lib = gpu.lib()
    assert isinstance(lib.olamgpu_last_error(None), bytes)
