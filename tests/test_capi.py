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


def declared_signatures():
    """{name: [parameter declarations]} from include/ochre.h."""
    src = open(os.path.join(ROOT, "include", "ochre.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    sigs = {}
    for m in re.finditer(r"\b(ochre_[a-z_]+)\s*\(([^;{]*?)\)\s*;", src):
        params = [p.strip() for p in m.group(2).split(",") if p.strip() and p.strip() != "void"]
        sigs[m.group(1)] = params
    return sigs


def test_every_pointer_taking_entry_point_has_declared_argtypes():
    """ctypes passes an undeclared Python int argument as a 32-bit C int: a
    device pointer handed to an entry point without argtypes is truncated and
    the kernel reading it faults.  Every function of the header whose
    parameters include a pointer, a 64-bit integer, a size or a float must
    have argtypes of the right length once the library is loaded."""
    lib = _lib.load()
    wide = re.compile(r"\*|int64_t|size_t|float|double")
    missing = []
    for name, params in declared_signatures().items():
        if not any(wide.search(p) for p in params):
            continue
        at = getattr(lib, name).argtypes
        if at is None or len(at) != len(params):
            missing.append((name, len(params), None if at is None else len(at)))
    assert not missing, missing
