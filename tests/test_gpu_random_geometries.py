"""GPU parity over seeded random geometries: every kernel of ohx_pc_diffract
against the oracle, for draws of resolution (not tile multiples), pixel pitch
(x and y apart), wavelength, pixel offset, point count, depth range (near
points whose boxes are a few pixels wide, mid-range, far points whose boxes
cover the hologram) and complex amplitudes.

Real amplitudes: the whole field against the FP64 restatement of
genCghPointCloudCPU (the written pixels must be the restatement's; field
rel-L2 within each kernel's bound).  Complex amplitudes (Vertex::phase used,
which the reference's CPU path ignores): 2000 pixels against the
extended-precision oracle.  Bounds (field rel-L2):
  FP64 direct 1e-9, FP32 direct 1e-5, BF16 3-term 1e-4, FP16 3-term 2e-5
  (Fresnel; RS with FP16x3 runs the FP32 direct kernel), BF16 1-term 1e-2.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

F32 = capi.MODE_GPU | capi.MODE_FLOAT
MODES = {
    "fp64": (capi.MODE_GPU, 1e-9),
    "fp32": (F32, 1e-5),
    "bf16x3": (F32 | capi.MODE_MFMA_BF16X3, 1e-4),
    "fp16x3": (F32 | capi.MODE_MFMA_FP16X3, 2e-5),
    "bf16x1": (F32 | capi.MODE_MFMA_BF16X1, 1e-2),
}
DEPTHS = {"near": (2e-3, 8e-3), "mid": (0.02, 0.08), "far": (0.15, 0.6)}


def case(seed):
    rng = np.random.default_rng(7100 + seed)
    pn = (int(rng.integers(60, 700)), int(rng.integers(40, 480)))
    pp = (float(rng.uniform(3e-6, 12e-6)), float(rng.uniform(3e-6, 12e-6)))
    lam = float(rng.uniform(440e-9, min(660e-9, 1.9 * min(pp))))
    off = (int(rng.integers(-6, 7)), int(rng.integers(-6, 7)))
    n = int(rng.integers(1, 300))
    depth = list(DEPTHS)[seed % 3]
    half = float(rng.uniform(0.3, 1.2)) * max(pn[0] * pp[0], pn[1] * pp[1]) / 2
    v = synth.random_cloud(7200 + seed, n, half, DEPTHS[depth], complex_amplitude=seed % 4 == 3)
    return pn, pp, lam, off, v, depth, seed % 4 == 3


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("diff", [0, 1])
def test_random_geometry_all_kernels(ctx, orc, seed, diff):
    pn, pp, lam, off, v, depth, cplx = case(seed)
    g = orc.make_geom(pn, pp, offset=off)
    npx = pn[0] * pn[1]
    if cplx:
        pix = np.random.default_rng(7300 + seed).choice(npx, min(2000, npx), replace=False)
        want = orc.field_exact(g, v, lam, 0, diff, pix, use_phase=True)
    else:
        want = orc.gen_restate(g, v, [lam], diff)[0].ravel()
    if not np.any(want != 0):
        pytest.skip("no point's box meets the hologram in this draw")
    for name, (mode, tol) in MODES.items():
        got = capi.generate(ctx, v, pn, pp, [lam], diff, mode, offset=off, use_phase=cplx)[0].ravel()
        if cplx:
            got = got[pix]
        err = rel_l2(got, want)
        print(f"seed {seed} diff {diff} {depth} {pn} pp {pp[0]:.2e}/{pp[1]:.2e} lam {lam:.3e} n {len(v)} "
              f"cplx {cplx} {name}: rel-L2 {err:.2e}")
        assert np.array_equal(got != 0, want != 0), (seed, diff, name)
        assert err < tol, (seed, diff, name, err)
