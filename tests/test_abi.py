"""CPU checks of the drop-in boundary: the C-ABI library loads (no GPU
needed to load it) and exports every entry point include/openholo_amd.h
declares; the ctypes binding covers the same set."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ohx_[a-z0-9_]+)\s*\(", text)))


def test_header_exports_match_binding():
    from openholo_amd import capi
    declared = header_symbols(os.path.join(ROOT, "include", "openholo_amd.h"))
    assert declared, "no entry points parsed"
    assert sorted(capi.EXPORTS) == declared


def test_library_loads_and_exports_every_symbol():
    from openholo_amd import capi
    lib = capi.lib()
    for name in header_symbols(os.path.join(ROOT, "include", "openholo_amd.h")):
        assert hasattr(lib, name), name


def test_library_is_gfx950_code_object():
    path = os.path.join(ROOT, "openholo_amd", "libopenholo_amd.so")
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


def test_status_strings_without_gpu():
    from openholo_amd import capi
    lib = capi.lib()
    assert lib.ohx_status_string(0) == b"ok"
    assert lib.ohx_status_string(-4) == b"admission denied"


def test_null_context_rejected():
    from openholo_amd import capi
    lib = capi.lib()
    p = capi.params((16, 16), (8e-6, 8e-6), 532e-9)
    assert lib.ohx_pc_diffract(None, None, 0, None, ctypes.byref(p), 0, 2) == -1
    assert lib.ohx_field_zero(None, None, 0) == -1
