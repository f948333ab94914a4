"""The C-ABI library loads and exports every entry point include/afd.h declares (no GPU needed)."""
import ctypes
import os
import re

from astrofd_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "afd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(afd_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "afd_create" in names and "afd_step" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = L.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"libafd.so lacks {missing}"


def test_binding_covers_header():
    assert set(declared_symbols()) <= set(L.exported_symbols())


def test_version_and_param_count_without_device():
    lib = L.load()
    assert b"gfx950" in lib.afd_version()
    n = ctypes.c_int32()
    assert lib.afd_mlp_param_count(ctypes.byref(n)) == 0
    assert n.value == 4674


def test_invalid_params_rejected_before_touching_a_device():
    lib = L.load()
    p = L.Params()
    p.physics = 7
    h = ctypes.c_void_p()
    rc = lib.afd_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"physics" in lib.afd_last_error()
