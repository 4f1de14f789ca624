"""HIP events on an arbitrary stream (ctypes on libamdhip64), used to time
kernels on the stream they are launched on."""
from __future__ import annotations

import ctypes
import ctypes.util
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        for cand in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64.so not found")
        vp = ctypes.c_void_p
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        _hip.hipEventRecord.argtypes = [vp, vp]
        _hip.hipEventSynchronize.argtypes = [vp]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        _hip.hipEventDestroy.argtypes = [vp]
    return _hip


class Event:
    def __init__(self):
        h = ctypes.c_void_p()
        st = hip().hipEventCreate(ctypes.byref(h))
        if st != 0:
            raise RuntimeError(f"hipEventCreate failed: {st}")
        self.h = h

    def record(self, stream):
        st = hip().hipEventRecord(self.h, ctypes.c_void_p(stream))
        if st != 0:
            raise RuntimeError(f"hipEventRecord failed: {st}")

    def synchronize(self):
        hip().hipEventSynchronize(self.h)

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        st = hip().hipEventElapsedTime(ctypes.byref(ms), self.h, later.h)
        if st != 0:
            raise RuntimeError(f"hipEventElapsedTime failed: {st}")
        return float(ms.value)

    def __del__(self):
        try:
            if self.h:
                hip().hipEventDestroy(self.h)
        except Exception:
            pass
