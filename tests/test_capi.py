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


def test_structure_layouts_match_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives the header's struct."""
    mirrors = {"rp_shape_desc": ra.ShapeDesc, "rp_move_desc": ra.MoveDesc, "rp_box_desc": ra.BoxDesc,
               "rp_counters": ra.Counters, "rp_timing": ra.Timing, "rp_guard_stats": ra.GuardStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "rampack_amd.h"', "int main(void) {"]
    for cname, py in mirrors.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    seen = 0
    for line in filter(None, out):
        cname, field, value = line.split()
        py = mirrors[cname]
        got = ctypes.sizeof(py) if field == "size" else getattr(py, field).offset
        assert got == int(value), (cname, field)
        seen += 1
    assert seen == sum(len(py._fields_) + 1 for py in mirrors.values())


def test_run_replicas_argument_checks():
    """The list checks compare pointers before any context is read: a context listed twice or a null entry is
    refused, and an empty list is a no-op."""
    lib = ra.load_library()
    fake = (ctypes.c_void_p * 3)(0x1000, 0x2000, 0x1000)
    assert lib.rp_run_replicas(fake, 0, 1, 0, None) == ra.RP_OK
    assert lib.rp_run_replicas(fake, 3, 1, 0, None) == ra.RP_ERR_PRECONDITION
    assert b"replicas 0 and 2 are the same context" in lib.rp_last_error()
    nulls = (ctypes.c_void_p * 2)(0x1000, None)
    assert lib.rp_run_replicas(nulls, 2, 1, 0, None) == ra.RP_ERR_PRECONDITION
    assert b"replica 1: null context" in lib.rp_last_error()
