"""Pin of the FP32 sine/cosine of revolutions that the FP32 kernels use
(openholo_amd/csrc/ohx_internal.h sin_rev / cos_rev, the hardware
v_sin_f32 / v_cos_f32), evaluated through ohx_sincos_rev.

The direct kernels hand these functions unreduced phases below 65 cycles and
the matrix-core operand generation below 21 cycles, relying on the hardware
to reduce revolutions itself.  Checked here on such arguments against
sin(2 pi fract(x)) in FP64 of the same FP32 argument:
  * max |error| <= 5e-7 and RMS error <= 1e-7 for |x| <= 65 (measured on
    MI355X: 1.24e-7 and 3.5e-8; a phase error of 5e-7 rad is 20x below the
    FP32 direct kernels' 1e-5 field bound);
  * sin/cos(x) against sin/cos(fract(x)) on the device: <= 4e-7 apart
    (measured 2.1e-7: the unreduced argument costs at most 2 ulp of the
    result).
"""
import numpy as np
import pytest

from openholo_amd import capi

pytestmark = pytest.mark.gpu


def _eval(ctx, x):
    x = np.ascontiguousarray(x, np.float32)
    dx, ds, dc = ctx.to_device(x), ctx.alloc(4 * len(x)), ctx.alloc(4 * len(x))
    try:
        ctx._check(capi.lib().ohx_sincos_rev(ctx.h, dx.ptr, len(x), ds.ptr, dc.ptr), "ohx_sincos_rev")
        return ds.download(np.float32, len(x)), dc.download(np.float32, len(x))
    finally:
        for b in (dx, ds, dc):
            b.free()


def test_sincos_unreduced_revolutions():
    rng = np.random.default_rng(65)
    lim = 65.0
    x = np.concatenate([
        rng.uniform(-lim, lim, 1 << 20),
        np.arange(-65, 66, dtype=np.float64),                       # whole revolutions
        np.arange(-65, 66) + 0.25, np.arange(-65, 66) + 0.5, np.arange(-65, 66) + 0.75,
        np.nextafter(np.arange(-64, 65, dtype=np.float32), np.float32(np.inf)),
        np.nextafter(np.arange(-64, 65, dtype=np.float32), np.float32(-np.inf)),
        rng.uniform(-1e-3, 1e-3, 4096),
    ]).astype(np.float32)
    with capi.Context(0) as ctx:
        s, c = _eval(ctx, x)
        xd = x.astype(np.float64)
        fr = xd - np.floor(xd)                                       # exact for FP32 arguments
        es = s.astype(np.float64) - np.sin(2 * np.pi * fr)
        ec = c.astype(np.float64) - np.cos(2 * np.pi * fr)
        err = np.maximum(np.abs(es), np.abs(ec))
        rms = float(np.sqrt(np.mean(es ** 2 + ec ** 2) / 2))
        print(f"|x| <= {lim}: max |err| {err.max():.3e}, rms {rms:.3e}")
        assert err.max() <= 5e-7
        assert rms <= 1e-7
        # the same arguments reduced on the host first
        s1, c1 = _eval(ctx, fr.astype(np.float32))
        d = np.maximum(np.abs(s1.astype(np.float64) - s), np.abs(c1.astype(np.float64) - c))
        print(f"unreduced vs reduced: max |diff| {d.max():.3e}, {int((d > 0).sum())} of {len(x)} differ")
        assert d.max() <= 4e-7
