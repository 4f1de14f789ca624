"""GPU parity of the WRP generator (ohx_wrp_generate) against the numpy
oracle (oracle/wrp_oracle.py), through the C ABI.

Tolerances (relative L2): WRP plane 1e-12 (FP64 terms, a different
summation order), hologram field 1e-10 (FP64 FFTs of different libraries);
the set of nonzero WRP pixels is exact.
"""
import numpy as np
import pytest

from openholo_amd import capi, synth
from oracle import wrp_oracle as wo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def cloud(seed, n, xy, z):
    v = synth.random_cloud(seed, n, xy, z)
    rng = np.random.default_rng(seed)
    v["color"][:, 1] = rng.uniform(0.1, 1.0, n)
    v["color"][:, 2] = rng.uniform(0.1, 1.0, n)
    return v


@pytest.mark.parametrize("case", [
    # pn, pp, lambda, channel, scale, dz past zmax, propagation distance
    ((256, 192), (8e-6, 8e-6), 532e-9, 0, (1.0, 1.0, 1.0), 1.0e-3, 0.05),
    ((200, 150), (6.4e-6, 8e-6), 633e-9, 1, (2.0, 1.5, 1.0), -0.8e-3, 0.02),   # dz < 0, odd sizes
    ((128, 96), (10e-6, 10e-6), 473e-9, 2, (1.0, 1.0, 3.0), 2.5e-3, -0.03),    # back-propagation
    ((160, 120), (8e-6, 8e-6), 532e-9, 0, (1.0, 1.0, 1.0), 4.0e-3, 0.03),      # w = 33: table in global memory
])
def test_wrp_matches_oracle(ctx, case):
    pn, pp, lam, ch, scale, dz, dist = case
    v = cloud(11 + ch, 300, 0.9e-3, (0.0, 2e-3))
    zmax = max(wo.MIN_DOUBLE, float((v["pos"][:, 2] * scale[2]).max()))
    loc = zmax + dz
    wrp, field = capi.generate_wrp(ctx, v, pn, pp, lam, loc, dist, channel=ch, scale=scale)
    st = ctx.last_stats()
    w = wo.window_half_width(lam, dz, pp[0])
    assert st.kernel_kind == 3 and st.pairs_nominal == len(v) * (2 * w) ** 2
    want_wrp, want_field = wo.generate(v, pn, pp, scale, lam, ch, loc, dist)
    assert np.count_nonzero(want_wrp) > 0.2 * want_wrp.size
    assert np.array_equal(wrp != 0, want_wrp != 0)
    assert rel(wrp, want_wrp) < 1e-12
    assert rel(field, want_field) < 1e-10


def test_wrp_is_deterministic_and_clips_far_points(ctx):
    pn, pp, lam = (160, 120), (8e-6, 8e-6), 532e-9
    v = cloud(21, 500, 1.2e-3, (0.0, 1e-3))          # many windows off the plane
    zmax = float(v["pos"][:, 2].max())
    a = capi.generate_wrp(ctx, v, pn, pp, lam, zmax + 1.2e-3, 0.04)
    b = capi.generate_wrp(ctx, v, pn, pp, lam, zmax + 1.2e-3, 0.04)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    want = wo.wrp_plane(v, pn, pp, (1.0, 1.0, 1.0), lam, 0, zmax + 1.2e-3)
    assert rel(a[0], want) < 1e-12


def test_wrp_dense_cluster(ctx):
    """3000 points within +-15 pixels: every tile near the cluster keeps more
    points than one LDS pass holds (1024), so the gather runs several passes."""
    pn, pp, lam = (96, 80), (8e-6, 8e-6), 532e-9
    v = cloud(51, 3000, 0.12e-3, (0.0, 1e-3))
    zmax = float(v["pos"][:, 2].max())
    loc = zmax + 1.0e-3
    wrp, field = capi.generate_wrp(ctx, v, pn, pp, lam, loc, 0.02)
    want_wrp, want_field = wo.generate(v, pn, pp, (1.0, 1.0, 1.0), lam, 0, loc, 0.02)
    assert rel(wrp, want_wrp) < 1e-12
    assert rel(field, want_field) < 1e-10


def test_wrp_zero_window_gives_zero_field(ctx):
    v = cloud(31, 50, 0.3e-3, (0.0, 1e-3))
    zmax = float(v["pos"][:, 2].max())
    wrp, field = capi.generate_wrp(ctx, v, (64, 48), (8e-6, 8e-6), 532e-9, zmax + 1e-9, 0.01)
    assert not wrp.any() and not field.any()


def test_wrp_argument_errors(ctx):
    v = cloud(41, 10, 0.3e-3, (0.0, 1e-3))
    dv = ctx.to_device(v)
    fld = ctx.alloc(16 * 64 * 48)
    try:
        p = capi.wrp_params((64, 48), (8e-6, 8e-6), 532e-9, 2e-3, 0.01)
        p.random_phase = 1
        with pytest.raises(capi.OhxError):
            ctx.wrp_generate(dv, len(v), p, fld)
        p = capi.wrp_params((64, 48), (8e-6, 8e-6), 532e-9, 2e-3, 0.01, channel=3)
        with pytest.raises(capi.OhxError):
            ctx.wrp_generate(dv, len(v), p, fld)
    finally:
        dv.free()
        fld.free()
