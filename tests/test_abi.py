"""The C ABI library loads and exports every symbol include/halma.h declares (CPU only:
no compute calls)."""
import ctypes
import os
import re

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "halma.h")
LIB = os.path.join(ROOT, "pyhalma_amd", "libhalma.so")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(halma_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "halma_pot_sum_f32" in syms and "halma_last_error" in syms
    assert len(syms) >= 8


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build with python -m pyhalma_amd.build"
    lib = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pyhalma_amd import _lib

    assert set(declared_symbols()) <= set(_lib.SIGNATURES), (
        set(declared_symbols()) - set(_lib.SIGNATURES))


def test_error_string_without_gpu():
    lib = ctypes.CDLL(LIB)
    lib.halma_last_error.restype = ctypes.c_char_p
    lib.halma_tune_set.argtypes = [ctypes.c_int, ctypes.c_int]
    assert lib.halma_tune_set(1, 3) == -1  # tests per lane must be 1, 2 or 4
    assert b"tests per lane" in lib.halma_last_error()
    assert lib.halma_version() == 10000


def test_symmetric_limit_key_without_gpu():
    from pyhalma_amd import _lib

    lib = _lib.load()
    key = _lib.TUNE_SYM_MAX
    old = lib.halma_tune_get(key)
    assert old == 65536
    assert lib.halma_tune_set(key, 65537) == -1
    assert b"symmetric-kernel halo limit" in lib.halma_last_error()
    assert lib.halma_tune_set(key, 1000) == 0 and lib.halma_tune_get(key) == 1000
    assert lib.halma_tune_set(key, old) == 0
