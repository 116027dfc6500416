"""The C ABI library loads and exports every symbol include/bem_hip.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(REPO, "include", "bem_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bem_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_expected_entry_points(bem):
    syms = declared_symbols()
    assert len(syms) >= 16
    assert set(syms) == set(bem._lib.EXPORTS)


def test_library_exports_every_declared_symbol(bem):
    lib = bem._lib.lib()
    for name in declared_symbols():
        assert hasattr(lib, name), name
    assert lib.bem_version() == 1


def test_null_arguments_are_rejected_without_gpu(bem):
    lib = bem._lib.lib()
    assert lib.bem_create(None, 0) == bem._lib.BEM_ERR_INVALID
    assert b"null" in lib.bem_last_error()
    assert lib.bem_destroy(None) == 0
    assert lib.bem_set_mesh(None, None, None, 0, 0, 5, None) == bem._lib.BEM_ERR_INVALID


def test_built_for_gfx950():
    so = os.path.join(REPO, "taichi-bem_amd", "libbem_hip.so")
    data = open(so, "rb").read()
    assert b"gfx950" in data
