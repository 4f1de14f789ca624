"""GPU parity of the matrix-core Fresnel path H = -j V U^T (through the C ABI).

Field relative-L2 tolerances against the FP64 oracle (stated per variant):
  BF16 1-term : 1e-2   (one BF16 rounding per factor, ~2^-9)
  BF16 3-term : 1e-4   (hi*hi + hi*lo + lo*hi, ~2^-17 per factor)
  FP16 3-term : 2e-5
The mask (set of written pixels) is exact: U and V are exactly zero outside
each point's box, as the reference's loops never touch those pixels.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

VARIANTS = {
    "bf16x1": (capi.MODE_GPU | capi.MODE_FLOAT | capi.MODE_MFMA_BF16X1, 1e-2),
    "bf16x3": (capi.MODE_GPU | capi.MODE_FLOAT | capi.MODE_MFMA_BF16X3, 1e-4),
    "fp16x3": (capi.MODE_GPU | capi.MODE_FLOAT | capi.MODE_MFMA_FP16X3, 2e-5),
}
FP32 = capi.MODE_GPU | capi.MODE_FLOAT


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", list(VARIANTS))
def test_gemm_vs_exact_small(ctx, orc, name):
    mode, tol = VARIANTS[name]
    pn, pp, lam = (300, 170), (8e-6, 8e-6), 532e-9       # not tile multiples
    v = synth.random_cloud(31, 1999, 1.5e-3, (0.15, 0.3))  # not a stage multiple
    g = orc.make_geom(pn, pp)
    got = capi.generate(ctx, v, pn, pp, [lam], 1, mode)[0].ravel()
    pix = np.random.default_rng(2).choice(pn[0] * pn[1], 3000, replace=False)
    want = orc.field_exact(g, v, lam, 0, 1, pix)
    err = rel_l2(got[pix], want)
    print(f"{name}: rel-L2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("name", list(VARIANTS))
def test_gemm_partial_masks_exact(ctx, orc, name):
    """Tiny z: boxes a few pixels wide, so most tiles see partial masks."""
    mode, tol = VARIANTS[name]
    pn, pp, lam = (150, 90), (10e-6, 10e-6), 633e-9
    g = orc.make_geom(pn, pp, offset=(2, -3))
    v = synth.random_cloud(41, 40, 500e-6, (1.5e-3, 4e-3))
    for i in (0, 7, 19):
        want = orc.gen_restate(g, v[i:i + 1], [lam], 1)[0]
        got = capi.generate(ctx, v[i:i + 1], pn, pp, [lam], 1, mode, offset=(2, -3))[0]
        assert np.array_equal(want != 0, got != 0), (name, i)
    want = orc.gen_restate(g, v, [lam], 1)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], 1, mode, offset=(2, -3))[0]
    assert rel_l2(got, want) < tol


@pytest.mark.parametrize("name", list(VARIANTS))
def test_gemm_complex_amplitude_full_size(ctx, orc, name):
    """3840x2160 at 3.74 um, 532 nm, complex amplitudes (the 1M-point
    workload), first 10,000 points; 3,000 sampled pixels vs the FP64 oracle."""
    mode, tol = VARIANTS[name]
    w = synth.WORKLOADS[3]
    v = synth.points(w, 10_000)
    g = orc.make_geom(w.pn, w.pp)
    got = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], 1, mode, use_phase=True)[0].ravel()
    pix = np.random.default_rng(4).choice(w.n_pixels, 3000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, 1, pix, use_phase=True)
    err = rel_l2(got[pix], want)
    print(f"{name}: rel-L2 {err:.3e}")
    assert err < tol


def test_gemm_matches_direct_and_is_deterministic(ctx):
    w = synth.WORKLOADS[2]
    v = synth.points(w, 5000)
    mode, tol = VARIANTS["fp16x3"]
    a = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[2]], 1, mode)[0]
    b = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[2]], 1, mode)[0]
    assert np.array_equal(a, b)
    d = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[2]], 1, FP32)[0]
    assert rel_l2(a, d) < 2e-5


def test_gemm_rs_fp16x3_runs_direct(ctx):
    """RS has no FP16 matrix-core form (its amplitudes exceed the FP16 range):
    with MODE_MFMA_FP16X3 it runs the FP32 direct sum, bit for bit."""
    pn, pp = (96, 64), (8e-6, 8e-6)
    v = synth.random_cloud(5, 50, 200e-6, (0.1, 0.2))
    a = capi.generate(ctx, v, pn, pp, [532e-9], 0, VARIANTS["fp16x3"][0])[0]
    b = capi.generate(ctx, v, pn, pp, [532e-9], 0, FP32)[0]
    assert np.array_equal(a, b)


def test_gemm_chunked_point_range(ctx):
    """5M points at 3840x2160: the split-K scratch would exceed its 4 GiB cap,
    so the launch runs in two point chunks, each reduced into the field.  The
    result against the library's FP64 probe stays within the BF16x3 bound."""
    mode, tol = VARIANTS["bf16x3"]
    w = synth.WORKLOADS[3]
    n = 5_000_000
    v = synth.random_cloud(77, n, 8e-3, (0.4, 0.6))
    p = capi.params(w.pn, w.pp, w.lambdas[0])
    dv = ctx.to_device(v)
    npx = w.n_pixels
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, n, fld, p, 1, mode)
        st = ctx.last_stats()
        pix = np.random.default_rng(5).choice(npx, 128, replace=False)
        got = fld.download(np.complex128, npx)[pix]
        want = ctx.probe_f64(dv, n, p, 1, pix)
    finally:
        dv.free()
        fld.free()
    # two chunks of 2.5M points, each 255 tiles of 128 x 256 pixels x 39 splits
    assert st.kernel_kind == 2 and st.workgroups == 2 * 255 * 39
    err = rel_l2(got, want)
    print(f"chunked bf16x3: rel-L2 {err:.3e}")
    assert err < tol
