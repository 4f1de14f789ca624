"""The matrix-core paths carry no constant offset from their accumulation.

The FP32 accumulation of the MFMA instructions rounds with a bias toward
-inf.  Summed over the K loop of every split-K slice, that bias is a
constant offset of the field, the same in every slice, which grows with the
point count and dominated the error of long launches (4M points on 512 x 256
pixels: mean error 0.95 of the rms error, rms 6e-5 of the incoherent sum for
RS).  Odd slices accumulate the negated products and are subtracted in the
reduction, which cancels it.  Checked against the FP32 direct kernel on
every pixel, with the error measured against the incoherent sum
sqrt(sum |term|^2): the mean error (the offset) at most 0.1 of the rms error,
and the rms error below 2e-5.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

FP32 = capi.MODE_GPU | capi.MODE_FLOAT
CASES = [
    (capi.DIFF_FRESNEL, "bf16x3", FP32 | capi.MODE_MFMA_BF16X3),
    (capi.DIFF_FRESNEL, "fp16x3", FP32 | capi.MODE_MFMA_FP16X3),
    (capi.DIFF_RS, "bf16x3", FP32 | capi.MODE_MFMA_BF16X3),
]


@pytest.mark.parametrize("diff,name,mode", CASES)
def test_no_constant_offset(diff, name, mode):
    pn, pp, lam, n = (512, 256), (8e-6, 8e-6), 532e-9, 1 << 22
    npx = pn[0] * pn[1]
    print(f"diff {diff} {name}: {n} points, {pn}", flush=True)
    v = synth.random_cloud(777 + diff, n, 2e-3, (0.45, 0.55))
    inc = float(np.sqrt(np.sum((v["color"][:, 0] / (lam * v["pos"][:, 2])) ** 2)))
    with capi.Context(0) as ctx:
        dv = ctx.to_device(v)
        fld = ctx.alloc(16 * npx)
        try:
            p = capi.params(pn, pp, lam)
            out = []
            for m in (FP32, mode):
                ctx.field_zero(fld, npx)
                ctx.diffract(dv, n, fld, p, diff, m)
                ctx.sync()
                out.append(fld.download(np.complex128, npx))
        finally:
            dv.free()
            fld.free()
    e = out[1] - out[0]
    rms = float(np.sqrt(np.mean(np.abs(e) ** 2))) / inc
    mean = complex(np.mean(e)) / inc
    print(f"diff {diff} {name}: error rms {rms:.2e}, mean ({mean.real:+.2e}, {mean.imag:+.2e}) "
          "of the incoherent sum", flush=True)
    assert abs(mean) <= 0.1 * rms
    assert rms < 2e-5
