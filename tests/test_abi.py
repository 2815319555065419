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


def _header_enum(prefix):
    """Name -> value of the header's enumerators that start with `prefix`."""
    text = open(os.path.join(ROOT, "include", "cymeca.h")).read()
    out = {}
    for body in re.findall(r"enum\s*\{([^}]*)\}", text):
        for name, value in re.findall(r"(\w+)\s*=\s*(-?\d+)", body):
            if name.startswith(prefix):
                out[name] = int(value)
    return out


def test_option_numbers_match_the_header(pkg):
    """Every CYMG_OPT_* of the header has the package constant of the same
    number (OPT_*), and the package defines no other option."""
    opts = _header_enum("CYMG_OPT_")
    assert "CYMG_OPT_BLOCK_STAGE" in opts and "CYMG_OPT_FP32_APPLY" in opts
    for name, value in opts.items():
        assert getattr(pkg, name[len("CYMG_"):]) == value, name
    mine = {n for n in dir(pkg) if n.startswith("OPT_")}
    assert mine == {n[len("CYMG_"):] for n in opts}
