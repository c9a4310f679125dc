"""The C-ABI library: it loads, and it exports every function include/rampack_amd.h declares (no compute calls,
so these run without a GPU)."""
import ctypes
import os
import re

import rampack_amd as ra

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "rampack_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rp_[a-z_]+)\s*\(", text)))


def test_header_declarations_are_listed():
    assert declared_functions() == sorted(ra.EXPORTED_SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = ra.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_library_has_gfx950_code_object():
    data = open(ra.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_no_cpu_fallback_symbols():
    """The product library carries no host implementation of the overlap test: the only pair-decision path is the
    device kernel (the oracle lives in a separate library that the product never loads)."""
    lib = ra.load_library()
    for name in ("orc_pair_decisions", "ref_pair_decisions"):
        assert not hasattr(lib, name)


def test_last_error_is_a_string():
    lib = ra.load_library()
    assert isinstance(lib.rp_last_error(), bytes)


def test_structure_layouts_match_header():
    assert ctypes.sizeof(ra.ShapeDesc) == 48
    assert ctypes.sizeof(ra.MoveDesc) == 24
    assert ctypes.sizeof(ra.BoxDesc) == 16
    assert ctypes.sizeof(ra.Counters) == 10 * 8 + 3 * 8
    assert ctypes.sizeof(ra.Timing) == 12 * 8
