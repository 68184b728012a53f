"""C ABI: libgsr.so loads and exports every entry point include/gsr.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from pointrix_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "gsr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gsr_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "gsr_alpha_blending_bwd" in syms and "gsr_sort_count" in syms
    assert len(syms) >= 16


def test_library_exports_all_declared():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_version_and_workspace_sizes():
    lib = _lib.load()
    assert lib.gsr_version() >= 1
    a = lib.gsr_sort_ws_gauss_bytes(1000)
    b = lib.gsr_sort_ws_gauss_bytes(100000)
    assert 0 < a < b
    assert lib.gsr_sort_ws_inst_bytes(0) > 0


def test_struct_layouts_match_bindings():
    """The ctypes mirrors of the table structs have the library's sizes."""
    from pointrix_amd import densify, ops, optim
    lib = _lib.load()
    assert lib.gsr_adam_entry_bytes() == ctypes.sizeof(optim._Entry)
    assert lib.gsr_plane_grads_bytes() == ctypes.sizeof(ops._PlaneGrads)
    assert lib.gsr_row_copy_bytes() == ctypes.sizeof(densify._RowCopy)
