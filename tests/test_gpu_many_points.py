"""More points in one call than one matrix-core launch takes.

The matrix-core kernels address their per-point records with 32-bit byte
offsets, so one launch takes at most 2^24 points; a call with more runs as
several launches into the same field.  Here 2^24 + 4133 points on a 64 x 64
hologram (each call two launches per path), against the library's FP64 probe
on 512 pixels and the FP32 direct kernel on every pixel.  Tolerances: the
BF16 3-term bound of the other GEMM tests, 1e-4 rel-L2; the direct kernel
1e-5.

The points spread over +-2 mm, so that the field decorrelates across the
hologram and its rel-L2 averages many independent sums.  (Points within
+-0.2 mm of the axis give a nearly uniform field, one random-walk sum of
2^24 random phases whose magnitude can fall far below its rms, which
inflates every relative error by the same factor.)  The errors are also
printed against the incoherent sum sqrt(sum |term|^2), which does not
fluctuate that way.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

FP32 = capi.MODE_GPU | capi.MODE_FLOAT
BF16X3 = FP32 | capi.MODE_MFMA_BF16X3
N = (1 << 24) + 4133


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("diff", [capi.DIFF_FRESNEL, capi.DIFF_RS])
def test_more_points_than_one_launch(diff):
    pn, pp, lam = (64, 64), (8e-6, 8e-6), 532e-9
    npx = pn[0] * pn[1]
    print(f"diff {diff}: {N} points, {pn}", flush=True)
    v = synth.random_cloud(20261203 + diff, N, 2.0e-3, (0.45, 0.55))
    # |term| = a / (lambda z) for Fresnel, a z / (lambda r^2) ~ that for RS here
    inc = float(np.sqrt(np.sum((v["color"][:, 0] / (lam * v["pos"][:, 2])) ** 2)))
    pix = np.random.default_rng(5).choice(npx, 512, replace=False)
    with capi.Context(0) as ctx:
        dv = ctx.to_device(v)
        fld = ctx.alloc(16 * npx)
        try:
            p = capi.params(pn, pp, lam)
            ref = ctx.probe_f64(dv, N, p, diff, pix)
            out = {}
            for name, mode in (("direct", FP32), ("bf16x3", BF16X3)):
                ctx.field_zero(fld, npx)
                ctx.diffract(dv, N, fld, p, diff, mode)
                ctx.sync()
                out[name] = fld.download(np.complex128, npx)
                st = ctx.last_stats()
                if name == "bf16x3" and diff == capi.DIFF_RS:
                    assert st.gemm_points == N   # every point on the matrix cores
        finally:
            dv.free()
            fld.free()
    e_direct = rel_l2(out["direct"][pix], ref)
    e_gemm = rel_l2(out["bf16x3"][pix], ref)
    e_all = rel_l2(out["bf16x3"], out["direct"])
    n_gemm = float(np.sqrt(np.mean(np.abs(out["bf16x3"][pix] - ref) ** 2))) / inc
    print(f"diff {diff}: direct {e_direct:.2e}, bf16x3 {e_gemm:.2e} vs FP64 probe "
          f"({n_gemm:.2e} of the incoherent sum; |field| rms / incoherent sum "
          f"{float(np.sqrt(np.mean(np.abs(ref) ** 2))) / inc:.2f}); bf16x3 vs direct on all pixels {e_all:.2e}",
          flush=True)
    assert e_direct < 1e-5
    assert e_gemm < 1e-4
    assert e_all < 1e-4
