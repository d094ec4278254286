"""CPU checks of the C ABI library: it loads and exports every symbol that
include/ochre.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from ochre_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ochre.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(ochre_[a-z_]+)\s*\(", src)
    return sorted(set(names))


def test_header_declares_the_bound_symbols():
    decl = declared_functions()
    assert set(_lib.EXPORTS) <= set(decl)
    assert "ochre_rasterizer_fill" in decl and "ochre_rasterize_plain" in decl


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_binds_and_reports_build_info():
    lib = _lib.load()
    info = lib.ochre_build_info()
    assert b"gfx950" in info


def test_context_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        return
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.ochre_context_create(0, ctypes.byref(h)) == -6
    assert lib.ochre_rasterizer_new() is None
