"""The allocation-free entry point ohx_pc_diffract_noalloc against the
existing entry point ohx_pc_diffract, through the C ABI.

Both run the same direct kernels; the allocation-free one keeps every tile's
points in one workgroup while ohx_pc_diffract may split them into partial
planes, so the two agree to FP64 rounding of the per-pixel sums: relative L2
<= 1e-12 (the FP32 direct kernels' own bound against the oracle is 1e-5),
identical sets of written pixels, identical in-box pair counts.  Two calls in
a row, the second with fewer points on a smaller hologram (the static scratch
keeps the first call's contents).  Also: no device memory is allocated by the
call, and sizes beyond OHX_NOALLOC_MAX_POINTS and the matrix-core modes are
refused with a status code.
"""
import ctypes

import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

FP32 = capi.MODE_GPU | capi.MODE_FLOAT
FP64 = capi.MODE_GPU
MAX_POINTS = 524288


def _run(ctx, fn, dv, n, pn, pp, lam, diff, mode):
    npx = pn[0] * pn[1]
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        p = capi.params(pn, pp, lam)
        ctx._check(fn(ctx.h, dv.ptr, int(n), fld.ptr, ctypes.byref(p), int(diff), int(mode)), fn.__name__)
        st = ctx.last_stats()
        return fld.download(np.complex128, npx), st.pairs_in_box
    finally:
        fld.free()


@pytest.mark.parametrize("mode", [FP32, FP64])
@pytest.mark.parametrize("diff", [0, 1])
def test_noalloc_matches_existing_entry_point(diff, mode):
    lib = capi.lib()
    calls = [  # (workload, points, hologram): the second call is smaller
        (synth.WORKLOADS[2], 40_000, (3840, 2160)),
        (synth.WORKLOADS[1], 3_000, (1920, 1080)),
    ]
    with capi.Context(0) as ctx:
        for w, n, pn in calls:
            v = synth.points(w, n)
            dv = ctx.to_device(v)
            try:
                got, box_got = _run(ctx, lib.ohx_pc_diffract_noalloc, dv, n, pn, w.pp, w.lambdas[0], diff, mode)
                want, box_want = _run(ctx, lib.ohx_pc_diffract, dv, n, pn, w.pp, w.lambdas[0], diff, mode)
            finally:
                dv.free()
            rel = float(np.linalg.norm(got - want) / np.linalg.norm(want))
            print(f"diff={diff} mode={mode} n={n} {pn}: rel-L2 {rel:.2e}")
            assert np.array_equal(got != 0, want != 0)
            assert rel <= 1e-12
            assert box_got == box_want > 0


def test_noalloc_allocates_nothing_and_refuses_bad_sizes():
    lib = capi.lib()
    w = synth.WORKLOADS[2]
    with capi.Context(0) as ctx:
        v = synth.points(w, 300_000)
        dv = ctx.to_device(v)
        fld = ctx.alloc(16 * w.n_pixels)
        try:
            p = capi.params(w.pn, w.pp, w.lambdas[0])
            # warm every kernel variant the entry point launches
            for diff in (0, 1):
                for mode in (FP32, FP64):
                    ctx._check(lib.ohx_pc_diffract_noalloc(ctx.h, dv.ptr, 100, fld.ptr, ctypes.byref(p), diff, mode),
                               "warm")
            ctx.sync()
            free0, _ = ctx.mem_info()
            ctx._check(lib.ohx_pc_diffract_noalloc(ctx.h, dv.ptr, len(v), fld.ptr, ctypes.byref(p), 0, FP32),
                       "noalloc 300k")
            ctx.sync()
            free1, _ = ctx.mem_info()
            # 300k points of per-point constants would be 19 MB if allocated
            assert free0 - free1 < (2 << 20), (free0, free1)
            st = lib.ohx_pc_diffract_noalloc(ctx.h, dv.ptr, MAX_POINTS + 1, fld.ptr, ctypes.byref(p), 0, FP32)
            assert st == -1
            assert b"exceeds the maximum" in lib.ohx_last_error(ctx.h)
            st = lib.ohx_pc_diffract_noalloc(ctx.h, dv.ptr, 1000, fld.ptr, ctypes.byref(p), 1,
                                             FP32 | capi.MODE_MFMA_BF16X3)
            assert st == -7
        finally:
            dv.free()
            fld.free()
