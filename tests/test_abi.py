"""The C-ABI library loads and exports every entry point include/cymeca.h
declares; argument validation needs no GPU."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "cymeca.h")).read()
    return sorted(set(re.findall(r"\b(cymg_[a-z_0-9]+)\s*\(", text)))


def test_header_lists_all_python_bindings(pkg):
    assert sorted(pkg.EXPORTED) == declared_symbols()


def test_library_exports_declared_symbols(pkg):
    L = C.CDLL(pkg.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_invalid_arguments_without_gpu(pkg):
    L = pkg.lib()
    h = C.c_void_p()
    assert L.cymg_create(4, 0, C.byref(h)) == pkg.ERR_INVALID
    assert L.cymg_create(3, 0, None) == pkg.ERR_INVALID
    assert L.cymg_destroy(None) == pkg.ERR_INVALID
    assert L.cymg_clear(None) == pkg.ERR_INVALID
    assert L.cymg_last_error(None) == b"null context"
