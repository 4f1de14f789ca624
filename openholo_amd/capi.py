"""ctypes binding of the C ABI in include/openholo_amd.h.

This is the reference-side binding a Python caller would add (the reference
itself has no Python binding for this path): plain C calls on
``libopenholo_amd.so``.  Loading fails loudly when the library has not been
built -- there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libopenholo_amd.so")

OK = 0
DIFF_RS, DIFF_FRESNEL = 0, 1
MODE_GPU, MODE_FLOAT, MODE_FASTMATH = 1, 2, 4
MODE_DOUBLE = 0
MODE_MFMA_BF16X1, MODE_MFMA_BF16X3, MODE_MFMA_FP16X3 = 0x10, 0x20, 0x40
ENCODE_PHASE, ENCODE_AMPLITUDE, ENCODE_REAL, ENCODE_IMAGINARY = 0, 1, 2, 3

# every symbol include/openholo_amd.h declares (checked by the CPU test suite)
EXPORTS = [
    "ohx_ctx_create", "ohx_ctx_destroy", "ohx_ctx_stream", "ohx_ctx_device", "ohx_status_string",
    "ohx_last_error", "ohx_mem_info", "ohx_malloc", "ohx_free", "ohx_host_alloc", "ohx_host_free", "ohx_memcpy_h2d", "ohx_memcpy_d2h",
    "ohx_memcpy_d2d", "ohx_memcpy_d2h_async_after", "ohx_copy_synchronize", "ohx_memcpy_d2h_after", "ohx_synchronize", "ohx_field_zero", "ohx_pc_diffract", "ohx_pc_diffract_noalloc", "ohx_field_accumulate",
    "ohx_pc_probe_f64", "ohx_sincos_rev", "ohx_wrp_generate", "ohx_encode", "ohx_encode_ssb", "ohx_normalize_u8", "ohx_encode_normalize_u8", "ohx_last_launch_stats",
    "ohx_server_create", "ohx_server_submit", "ohx_server_wait", "ohx_server_cancel",
    "ohx_server_counters_get", "ohx_server_destroy",
]
ERR_ADMISSION, ERR_NOT_FOUND, ERR_TIMEOUT, ERR_CANCELLED = -4, -5, -6, -8


class PcParams(ctypes.Structure):
    _fields_ = [
        ("pn_x", ctypes.c_int32), ("pn_y", ctypes.c_int32),
        ("offset_x", ctypes.c_int32), ("offset_y", ctypes.c_int32),
        ("pp_x", ctypes.c_double), ("pp_y", ctypes.c_double),
        ("scale", ctypes.c_double * 3), ("distance", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("channel", ctypes.c_int32), ("use_phase", ctypes.c_int32),
    ]


class WrpParams(ctypes.Structure):
    _fields_ = [
        ("pn_x", ctypes.c_int32), ("pn_y", ctypes.c_int32),
        ("pp_x", ctypes.c_double), ("pp_y", ctypes.c_double),
        ("scale", ctypes.c_double * 3), ("lambda_", ctypes.c_double),
        ("channel", ctypes.c_int32), ("random_phase", ctypes.c_int32),
        ("wrp_location", ctypes.c_double), ("propagation_distance", ctypes.c_double),
    ]


def wrp_params(pn, pp, lam, wrp_location, distance, channel=0, scale=(1.0, 1.0, 1.0)) -> WrpParams:
    p = WrpParams()
    p.pn_x, p.pn_y = int(pn[0]), int(pn[1])
    p.pp_x, p.pp_y = float(pp[0]), float(pp[1])
    for i in range(3):
        p.scale[i] = float(scale[i])
    p.lambda_ = float(lam)
    p.channel = int(channel)
    p.random_phase = 0
    p.wrp_location = float(wrp_location)
    p.propagation_distance = float(distance)
    return p


class LaunchStats(ctypes.Structure):
    _fields_ = [("pairs_nominal", ctypes.c_int64), ("workgroups", ctypes.c_int64),
                ("kernel_kind", ctypes.c_int32), ("gemm_points", ctypes.c_int32),
                ("pairs_in_box", ctypes.c_int64)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("max_active", ctypes.c_int32), ("use_graph", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("mem_budget", ctypes.c_uint64), ("slice_pairs", ctypes.c_double),
                ("slice_pairs_mfma", ctypes.c_double)]


class Job(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("n_points", ctypes.c_int64), ("params", PcParams),
                ("diff_flag", ctypes.c_uint32), ("mode", ctypes.c_uint32), ("out_field", ctypes.c_void_p),
                ("out_u8", ctypes.c_void_p)]


class JobStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32), ("submit_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("end_ms", ctypes.c_double),
                ("device_bytes", ctypes.c_uint64)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("jobs_done", ctypes.c_uint64), ("iterations", ctypes.c_uint64), ("graph_replays", ctypes.c_uint64),
                ("reserved_bytes", ctypes.c_uint64), ("peak_reserved_bytes", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64)]


def params(pn, pp, lam, channel=0, scale=(1.0, 1.0, 1.0), distance=0.0, offset=(0, 0),
           use_phase=False) -> PcParams:
    p = PcParams()
    p.pn_x, p.pn_y = int(pn[0]), int(pn[1])
    p.offset_x, p.offset_y = int(offset[0]), int(offset[1])
    p.pp_x, p.pp_y = float(pp[0]), float(pp[1])
    for i in range(3):
        p.scale[i] = float(scale[i])
    p.distance = float(distance)
    p.lambda_ = float(lam)
    p.channel = int(channel)
    p.use_phase = int(bool(use_phase))
    return p


class OhxError(RuntimeError):
    def __init__(self, status, what, detail=""):
        super().__init__(f"{what} failed: status {status} ({detail})")
        self.status = status


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OhxError(-7, "load libopenholo_amd.so",
                           f"{LIB_PATH} not built; run __graft_entry__.build() or make -C openholo_amd")
        L = ctypes.CDLL(LIB_PATH)
        vp, i64, u32, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint32, ctypes.c_int
        sz = ctypes.c_size_t
        sig = {
            "ohx_ctx_create": ([i32, ctypes.POINTER(vp)], i32),
            "ohx_ctx_destroy": ([vp], i32),
            "ohx_ctx_stream": ([vp], vp),
            "ohx_ctx_device": ([vp], i32),
            "ohx_status_string": ([i32], ctypes.c_char_p),
            "ohx_last_error": ([vp], ctypes.c_char_p),
            "ohx_mem_info": ([vp, ctypes.POINTER(sz), ctypes.POINTER(sz)], i32),
            "ohx_malloc": ([vp, sz, ctypes.POINTER(vp)], i32),
            "ohx_free": ([vp, vp], i32),
            "ohx_host_alloc": ([sz, ctypes.POINTER(vp)], i32),
            "ohx_host_free": ([vp], i32),
            "ohx_memcpy_h2d": ([vp, vp, vp, sz], i32),
            "ohx_memcpy_d2h": ([vp, vp, vp, sz], i32),
            "ohx_memcpy_d2d": ([vp, vp, vp, sz], i32),
            "ohx_memcpy_d2h_async_after": ([vp, vp, vp, sz], i32),
            "ohx_copy_synchronize": ([vp], i32),
            "ohx_memcpy_d2h_after": ([vp, vp, vp, sz], i32),
            "ohx_synchronize": ([vp], i32),
            "ohx_field_zero": ([vp, vp, i64], i32),
            "ohx_pc_diffract": ([vp, vp, i64, vp, ctypes.POINTER(PcParams), u32, u32], i32),
            "ohx_pc_diffract_noalloc": ([vp, vp, i64, vp, ctypes.POINTER(PcParams), u32, u32], i32),
            "ohx_field_accumulate": ([vp, vp, vp, i64], i32),
            "ohx_pc_probe_f64": ([vp, vp, i64, ctypes.POINTER(PcParams), u32, vp, i64, vp], i32),
            "ohx_sincos_rev": ([vp, vp, i64, vp, vp], i32),
            "ohx_wrp_generate": ([vp, vp, i64, ctypes.POINTER(WrpParams), vp, vp], i32),
            "ohx_encode": ([vp, vp, i64, vp, u32], i32),
            "ohx_encode_ssb": ([vp, vp, i32, i32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, vp], i32),
            "ohx_normalize_u8": ([vp, ctypes.POINTER(vp), i32, i64, ctypes.POINTER(vp)], i32),
            "ohx_encode_normalize_u8": ([vp, ctypes.POINTER(vp), i32, i64, u32, ctypes.POINTER(vp),
                                         ctypes.POINTER(vp)], i32),
            "ohx_last_launch_stats": ([vp, ctypes.POINTER(LaunchStats)], i32),
            "ohx_server_create": ([ctypes.POINTER(ServerOpts), ctypes.POINTER(vp)], i32),
            "ohx_server_submit": ([vp, ctypes.POINTER(Job), ctypes.POINTER(ctypes.c_uint64)], i32),
            "ohx_server_wait": ([vp, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(JobStats)], i32),
            "ohx_server_cancel": ([vp, ctypes.c_uint64], i32),
            "ohx_server_counters_get": ([vp, ctypes.POINTER(ServerCounters)], i32),
            "ohx_server_destroy": ([vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


class HostBuffer:
    """Page-locked, zero-filled host memory (ohx_host_alloc), the memory the
    drop-in's complex_H planes live in; `array` views it as numpy.  A view does
    not keep the buffer alive: drop every view before `free`."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        st = lib().ohx_host_alloc(self.nbytes, ctypes.byref(p))
        if st != OK:
            raise OhxError(st, "ohx_host_alloc", f"{self.nbytes} bytes")
        self.ptr = p.value

    def array(self, dtype, count: int) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count * dtype.itemsize > self.nbytes:
            raise ValueError("view larger than the buffer")
        raw = (ctypes.c_char * (count * dtype.itemsize)).from_address(self.ptr)
        return np.frombuffer(raw, dtype=dtype, count=count)

    def free(self):
        if self.ptr:
            st = lib().ohx_host_free(self.ptr)
            self.ptr = None
            if st != OK:
                raise OhxError(st, "ohx_host_free", "")


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx, self.nbytes = ctx, int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(lib().ohx_malloc(ctx.h, self.nbytes, ctypes.byref(p)), "ohx_malloc")
        self.ptr = p.value

    def free(self):
        if self.ptr:
            self.ctx._check(lib().ohx_free(self.ctx.h, self.ptr), "ohx_free")
            self.ptr = None

    def upload(self, arr: np.ndarray):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        self.ctx._check(lib().ohx_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), "h2d")

    def download(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._check(lib().ohx_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def download_after(self, out: np.ndarray):
        """ohx_memcpy_d2h_after into `out` (a contiguous array): waits for the work
        enqueued so far on the context's stream, not for work enqueued later."""
        assert out.flags["C_CONTIGUOUS"] and out.nbytes <= self.nbytes
        self.ctx._check(lib().ohx_memcpy_d2h_after(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes),
                        "d2h_after")
        return out

    def download_async_after(self, out: np.ndarray):
        """ohx_memcpy_d2h_async_after into `out`: enqueued after the work enqueued
        so far on the context's stream; `out` is valid after Context.copy_sync()
        and must stay alive until then."""
        assert out.flags["C_CONTIGUOUS"] and out.nbytes <= self.nbytes
        self.ctx._check(lib().ohx_memcpy_d2h_async_after(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes),
                        "d2h_async_after")
        return out


class Context:
    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        st = lib().ohx_ctx_create(int(device), ctypes.byref(h))
        if st != OK:
            raise OhxError(st, "ohx_ctx_create", lib().ohx_status_string(st).decode())
        self.h = h.value

    def _check(self, st, what):
        if st != OK:
            raise OhxError(st, what, lib().ohx_last_error(self.h).decode())

    def close(self):
        if self.h:
            lib().ohx_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return lib().ohx_ctx_stream(self.h)

    def alloc(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        a = np.ascontiguousarray(arr)
        b = DeviceBuffer(self, max(a.nbytes, 1))
        if a.nbytes:
            b.upload(a)
        return b

    def sync(self):
        self._check(lib().ohx_synchronize(self.h), "ohx_synchronize")

    def copy_sync(self):
        """ohx_copy_synchronize: wait for every read-back enqueued so far."""
        self._check(lib().ohx_copy_synchronize(self.h), "ohx_copy_synchronize")

    def field_zero(self, field: DeviceBuffer, n_px: int):
        self._check(lib().ohx_field_zero(self.h, field.ptr, int(n_px)), "ohx_field_zero")

    def diffract(self, d_vertices: DeviceBuffer, n_points: int, field: DeviceBuffer, p: PcParams,
                 diff: int, mode: int):
        self._check(lib().ohx_pc_diffract(self.h, d_vertices.ptr, int(n_points), field.ptr,
                                          ctypes.byref(p), int(diff), int(mode)), "ohx_pc_diffract")

    def accumulate(self, dst: DeviceBuffer, src: DeviceBuffer, n_px: int):
        self._check(lib().ohx_field_accumulate(self.h, dst.ptr, src.ptr, int(n_px)), "ohx_field_accumulate")

    def probe_f64(self, d_vertices: DeviceBuffer, n_points: int, p: PcParams, diff: int,
                  pix: np.ndarray) -> np.ndarray:
        pix = np.ascontiguousarray(pix, np.int64)
        dpix = self.to_device(pix)
        out = self.alloc(16 * max(len(pix), 1))
        try:
            self._check(lib().ohx_pc_probe_f64(self.h, d_vertices.ptr, int(n_points), ctypes.byref(p),
                                               int(diff), dpix.ptr, len(pix), out.ptr), "ohx_pc_probe_f64")
            return out.download(np.complex128, len(pix))
        finally:
            dpix.free()
            out.free()

    def wrp_generate(self, d_vertices: DeviceBuffer, n_points: int, p: WrpParams, field: DeviceBuffer,
                     wrp: DeviceBuffer | None = None):
        self._check(lib().ohx_wrp_generate(self.h, d_vertices.ptr, int(n_points), ctypes.byref(p), field.ptr,
                                           wrp.ptr if wrp is not None else None), "ohx_wrp_generate")

    def encode(self, field: DeviceBuffer, n_px: int, enc: DeviceBuffer, flag: int):
        self._check(lib().ohx_encode(self.h, field.ptr, int(n_px), enc.ptr, int(flag)), "ohx_encode")

    def encode_ssb(self, field: DeviceBuffer, pn, pp, enc: DeviceBuffer, spectrum_shift=(0.0, 0.5)):
        """ophPointCloud::encodeHologram of one channel (band_limit is unused by
        the reference and not taken)."""
        self._check(lib().ohx_encode_ssb(self.h, field.ptr, int(pn[0]), int(pn[1]), float(pp[0]), float(pp[1]),
                                         float(spectrum_shift[0]), float(spectrum_shift[1]), enc.ptr),
                    "ohx_encode_ssb")

    def normalize_u8(self, encs, n_px: int, outs):
        a = (ctypes.c_void_p * len(encs))(*[e.ptr for e in encs])
        b = (ctypes.c_void_p * len(outs))(*[o.ptr for o in outs])
        self._check(lib().ohx_normalize_u8(self.h, a, len(encs), int(n_px), b), "ohx_normalize_u8")

    def encode_normalize_u8(self, fields, n_px: int, flag: int, encs, outs):
        """ohx_encode_normalize_u8: encode every channel and normalise to 8 bits."""
        f = (ctypes.c_void_p * len(fields))(*[x.ptr for x in fields])
        e = (ctypes.c_void_p * len(encs))(*[x.ptr for x in encs])
        o = (ctypes.c_void_p * len(outs))(*[x.ptr for x in outs])
        self._check(lib().ohx_encode_normalize_u8(self.h, f, len(fields), int(n_px), int(flag), e, o),
                    "ohx_encode_normalize_u8")

    def last_stats(self) -> LaunchStats:
        s = LaunchStats()
        self._check(lib().ohx_last_launch_stats(self.h, ctypes.byref(s)), "ohx_last_launch_stats")
        return s

    def mem_info(self):
        f, t = ctypes.c_size_t(), ctypes.c_size_t()
        self._check(lib().ohx_mem_info(self.h, ctypes.byref(f), ctypes.byref(t)), "ohx_mem_info")
        return f.value, t.value


class Server:
    """Continuous-batching hologram server (one channel per job)."""

    def __init__(self, device=0, max_active=32, use_graph=True, mem_budget=0, slice_pairs=0.0,
                 slice_pairs_mfma=0.0):
        o = ServerOpts(int(device), int(max_active), int(bool(use_graph)), 0, int(mem_budget),
                       float(slice_pairs), float(slice_pairs_mfma))
        h = ctypes.c_void_p()
        st = lib().ohx_server_create(ctypes.byref(o), ctypes.byref(h))
        if st != OK:
            raise OhxError(st, "ohx_server_create", lib().ohx_status_string(st).decode())
        self.h = h.value
        self._keep = {}

    def submit(self, vertices: np.ndarray, p: PcParams, diff: int, mode: int, want_field=True, want_u8=False) -> int:
        v = np.ascontiguousarray(vertices)
        npx = p.pn_x * p.pn_y
        f = np.empty(npx, np.complex128) if want_field else None
        u = np.empty(npx, np.uint8) if want_u8 else None
        j = Job(v.ctypes.data if len(v) else None, len(v), p, int(diff), int(mode),
                f.ctypes.data if f is not None else None, u.ctypes.data if u is not None else None)
        jid = ctypes.c_uint64()
        st = lib().ohx_server_submit(self.h, ctypes.byref(j), ctypes.byref(jid))
        if st != OK:
            raise OhxError(st, "ohx_server_submit", lib().ohx_status_string(st).decode())
        self._keep[jid.value] = (f, u)
        return jid.value

    def wait(self, jid: int, timeout_s: float = -1.0):
        """-> (status, JobStats, field or None, u8 or None)"""
        stt = JobStats()
        st = lib().ohx_server_wait(self.h, ctypes.c_uint64(jid), float(timeout_s), ctypes.byref(stt))
        if st == ERR_TIMEOUT or st == ERR_NOT_FOUND:
            return st, stt, None, None
        f, u = self._keep.pop(jid, (None, None))
        return st, stt, f, u

    def cancel(self, jid: int) -> int:
        return lib().ohx_server_cancel(self.h, ctypes.c_uint64(jid))

    def counters(self) -> ServerCounters:
        c = ServerCounters()
        lib().ohx_server_counters_get(self.h, ctypes.byref(c))
        return c

    def close(self):
        if self.h:
            lib().ohx_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def generate(ctx: Context, vertices: np.ndarray, pn, pp, lambdas, diff: int, mode: int,
             scale=(1.0, 1.0, 1.0), distance=0.0, offset=(0, 0), use_phase=False) -> list[np.ndarray]:
    """Convenience: upload vertices, generate each channel, download complex128 fields."""
    n = len(vertices)
    dv = ctx.to_device(vertices)
    npx = int(pn[0]) * int(pn[1])
    fld = ctx.alloc(16 * npx)
    out = []
    try:
        for ch, lam in enumerate(lambdas):
            ctx.field_zero(fld, npx)
            p = params(pn, pp, lam, channel=ch, scale=scale, distance=distance, offset=offset,
                       use_phase=use_phase)
            ctx.diffract(dv, n, fld, p, diff, mode)
            out.append(fld.download(np.complex128, npx).reshape(int(pn[1]), int(pn[0])))
    finally:
        dv.free()
        fld.free()
    return out


def generate_wrp(ctx: Context, vertices: np.ndarray, pn, pp, lam, wrp_location, distance, channel=0,
                 scale=(1.0, 1.0, 1.0)):
    """Convenience: one WRP channel; returns (WRP plane, hologram field) as
    complex128 (pn_y, pn_x) arrays."""
    n = len(vertices)
    dv = ctx.to_device(vertices)
    npx = int(pn[0]) * int(pn[1])
    fld, wrp = ctx.alloc(16 * npx), ctx.alloc(16 * npx)
    try:
        ctx.wrp_generate(dv, n, wrp_params(pn, pp, lam, wrp_location, distance, channel, scale), fld, wrp)
        shape = (int(pn[1]), int(pn[0]))
        return (wrp.download(np.complex128, npx).reshape(shape), fld.download(np.complex128, npx).reshape(shape))
    finally:
        dv.free()
        fld.free()
        wrp.free()
