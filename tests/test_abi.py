"""The C-ABI library loads and exports every entry point include/aeolus_hip.h declares (CPU only)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "aeolus_hip.h")).read()
    return sorted(set(re.findall(r"\b(aeo_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    from aeolus_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libaeolus_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_no_context_is_an_error_not_a_crash():
    from aeolus_hip import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libaeolus_hip.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    lib.aeo_sync.restype = ctypes.c_int
    lib.aeo_last_error.restype = ctypes.c_char_p
    # before aeo_init every call reports AEO_ERR_STATE (no GPU touched)
    assert lib.aeo_sync() == -4
    assert b"aeo_init" in lib.aeo_last_error()


def test_product_never_imports_oracle():
    pkg = os.path.join(ROOT, "aeolus.jl_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", ".jl")):
                txt = open(os.path.join(dirpath, f)).read()
                for pat in ("import oracle", "from oracle", "liboracle", "orc_"):
                    assert pat not in txt, (os.path.join(dirpath, f), pat)
