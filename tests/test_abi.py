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


def test_host_alloc_contract():
    """ohx_host_alloc: page-locked, zero-filled memory where a device is
    present; without one OHX_ERR_DEVICE and a null pointer, the status the
    host classes take as the cue to use heap planes.  Null and zero-size
    arguments behave as documented either way."""
    from openholo_amd import capi
    lib = capi.lib()
    p = ctypes.c_void_p(1)
    assert lib.ohx_host_alloc(0, ctypes.byref(p)) == capi.OK and p.value is None
    assert lib.ohx_host_alloc(64, None) == -1          # OHX_ERR_INVALID_ARG
    assert lib.ohx_host_free(None) == capi.OK
    try:
        buf = capi.HostBuffer(1 << 20)
    except capi.OhxError as e:
        assert e.status == -3                          # OHX_ERR_DEVICE: no device here
        q = ctypes.c_void_p(1)
        assert lib.ohx_host_alloc(1 << 20, ctypes.byref(q)) == -3 and q.value is None
        return
    try:
        a = buf.array("u1", 1 << 20)
        assert not a.any()
        a[:] = 7
        assert int(buf.array("u1", 1 << 20).sum()) == 7 << 20
        del a
    finally:
        buf.free()
