"""ctypes front end of the CPU oracle (oracle/oph_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
to check the result, and by ``bench.py``'s CPU-baseline leg as the timed CPU
reference restatement.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "bin", "liboph_oracle.so")

VERTEX_DTYPE = np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])
assert VERTEX_DTYPE.itemsize == 56

RS, FRESNEL = 0, 1
ENCODE_PHASE, ENCODE_AMPLITUDE, ENCODE_REAL, ENCODE_IMAGINARY = 0, 1, 2, 3


class Geom(ctypes.Structure):
    _fields_ = [
        ("pn_x", ctypes.c_int32), ("pn_y", ctypes.c_int32),
        ("off_x", ctypes.c_int32), ("off_y", ctypes.c_int32),
        ("pp_x", ctypes.c_double), ("pp_y", ctypes.c_double),
        ("scale", ctypes.c_double * 3), ("distance", ctypes.c_double),
    ]


def make_geom(pn, pp, scale=(1.0, 1.0, 1.0), distance=0.0, offset=(0, 0)) -> Geom:
    g = Geom()
    g.pn_x, g.pn_y = int(pn[0]), int(pn[1])
    g.off_x, g.off_y = int(offset[0]), int(offset[1])
    g.pp_x, g.pp_y = float(pp[0]), float(pp[1])
    for i in range(3):
        g.scale[i] = float(scale[i])
    g.distance = float(distance)
    return g


def build() -> None:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_gen_restate.argtypes = [ctypes.POINTER(Geom), ctypes.c_void_p, ctypes.c_int64,
                                      dp, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(dp), ctypes.c_int]
        L.orc_field_exact.argtypes = [ctypes.POINTER(Geom), ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
        L.orc_mask.argtypes = [ctypes.POINTER(Geom), dp, ctypes.c_double, ctypes.c_int,
                               ctypes.c_int, ctypes.c_int]
        L.orc_mask.restype = ctypes.c_int
        L.orc_box.argtypes = [ctypes.POINTER(Geom), dp, ctypes.c_double, ctypes.c_int,
                              ctypes.POINTER(ctypes.c_int64)]
        L.orc_encode.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int]
        L.orc_normalize.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int64,
                                    ctypes.POINTER(ctypes.c_void_p)]
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def gen_restate(geom: Geom, vertices: np.ndarray, lambdas, diff_flag: int,
                n_threads: int = 0) -> list[np.ndarray]:
    """Reference CPU path (genCghPointCloudCPU).  Returns one complex128
    (pn_y, pn_x) field per wavelength."""
    v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    lam = np.ascontiguousarray(lambdas, dtype=np.float64)
    fields = [np.zeros((geom.pn_y, geom.pn_x), np.complex128) for _ in lam]
    arr = (ctypes.POINTER(ctypes.c_double) * len(fields))(*[_dptr(f.view(np.float64)) for f in fields])
    lib().orc_gen_restate(ctypes.byref(geom), v.ctypes.data, len(v), _dptr(lam), len(lam),
                          int(diff_flag), arr, int(n_threads))
    return fields


def field_exact(geom: Geom, vertices: np.ndarray, lam: float, channel: int, diff_flag: int,
                pix: np.ndarray, use_phase: bool = False, n_threads: int = 0) -> np.ndarray:
    """Independent extended-precision oracle at pixel indices pix (row*pn_x+col)."""
    v = np.ascontiguousarray(vertices, dtype=VERTEX_DTYPE)
    p = np.ascontiguousarray(pix, dtype=np.int64)
    out = np.zeros(len(p), np.complex128)
    lib().orc_field_exact(ctypes.byref(geom), v.ctypes.data, len(v), float(lam), int(channel),
                          int(diff_flag), int(bool(use_phase)), p.ctypes.data, len(p),
                          out.ctypes.data, int(n_threads))
    return out


def mask(geom: Geom, pos_scaled, lam: float, diff_flag: int, col: int, row: int) -> int:
    p = np.ascontiguousarray(pos_scaled, dtype=np.float64)
    return lib().orc_mask(ctypes.byref(geom), _dptr(p), float(lam), int(diff_flag), int(col), int(row))


def box(geom: Geom, pos_scaled, lam: float, diff_flag: int) -> tuple[int, int, int, int]:
    p = np.ascontiguousarray(pos_scaled, dtype=np.float64)
    b = (ctypes.c_int64 * 4)()
    lib().orc_box(ctypes.byref(geom), _dptr(p), float(lam), int(diff_flag), b)
    return tuple(int(x) for x in b)


def encode(field: np.ndarray, flag: int = ENCODE_PHASE) -> np.ndarray:
    f = np.ascontiguousarray(field, dtype=np.complex128)
    out = np.zeros(f.shape, np.float64)
    lib().orc_encode(f.ctypes.data, f.size, out.ctypes.data, int(flag))
    return out


def normalize(encs: list[np.ndarray]) -> list[np.ndarray]:
    es = [np.ascontiguousarray(e, dtype=np.float64) for e in encs]
    outs = [np.zeros(e.shape, np.uint8) for e in es]
    ein = (ctypes.c_void_p * len(es))(*[e.ctypes.data for e in es])
    eout = (ctypes.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    lib().orc_normalize(ein, len(es), es[0].size, eout)
    return outs
