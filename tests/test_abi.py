"""The C-ABI library loads and exports every entry point include/ne_runtime.h declares (no GPU calls)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ne_runtime.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(ne_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    assert "ne_advance_generation" in syms and "ne_create" in syms and len(syms) >= 20


def test_library_exports_every_declared_symbol(pkg):
    lib = ctypes.CDLL(pkg.runtime.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(pkg.runtime.EXPORTED_SYMBOLS) == declared_symbols()


def test_status_strings_without_a_device(pkg):
    lib = pkg.load_library()
    assert lib.ne_status_string(0) == b"ok"
    assert lib.ne_status_string(14) == b"slab full"
    assert lib.ne_status_string(16) == b"slab repack failed"
    assert lib.ne_status_string(999) == b"unknown status"
