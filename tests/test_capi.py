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
    assert lib.gsr_l1_terms_bytes() == ctypes.sizeof(ops._L1Terms)


def _header_prototypes():
    src = open(os.path.join(ROOT, "include", "gsr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"#[^\n]*", "", src)
    out = {}
    for m in re.finditer(r"(int|size_t|char\s*\*)\s*(gsr_[a-z0-9_]+)\s*\(([^)]*)\)\s*;", src):
        out[m.group(2)] = (m.group(1), [p.strip() for p in m.group(3).split(",")])
    return out


def _ctype_of(param):
    if param in ("void", ""):
        return None
    if "*" in param:
        return ctypes.c_void_p
    base = param.rsplit(None, 1)[0].replace("const ", "").strip()
    return {"int64_t": ctypes.c_int64, "int": ctypes.c_int, "float": ctypes.c_float,
            "double": ctypes.c_double, "size_t": ctypes.c_size_t}[base]


def test_binding_argument_types_match_header():
    """Every ctypes signature has the header's parameter count and types, in
    order (a missing or shifted argument would pass garbage to a kernel)."""
    protos = _header_prototypes()
    assert set(protos) == set(_lib.SIGNATURES)
    for name, (_, params) in protos.items():
        want = [t for t in (_ctype_of(p) for p in params) if t is not None]
        got = list(_lib.SIGNATURES[name][1])
        assert got == want, (name, [t.__name__ for t in got], [t.__name__ for t in want])
