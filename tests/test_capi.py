"""The C-ABI library loads and exports every symbol include/viltrum_amd.h declares
(no GPU calls)."""
import ctypes as C
import os
import re

import viltrum_amd as V

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "viltrum_amd.h")).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(vmx_\w+)\s*\(", src, flags=re.M)))


def test_header_declares_the_python_symbol_list():
    assert declared_symbols() == sorted(V.SYMBOLS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(V.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_layouts_match_header():
    # sizes of the C structs as seen by ctypes (x86-64 SysV layout)
    assert C.sizeof(V.Problem) == 6 * 4 + 8 + 12 * 8 + 3 * 8 + 3 * 8
    assert C.sizeof(V.Stats) == 8 * 8 + 5 * 8
    assert C.sizeof(V.TuneReport) == 5 * 4 + 4 + 2 * 8  # five int32, padding to the doubles
