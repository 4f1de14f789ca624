"""GPU parity of the direct point-cloud sums, through the C ABI.

Tolerances (field relative L2 unless stated):
  FP64 direct (MODE_DOUBLE)  vs restatement / exact oracle : 1e-9 (the
      reference's own FP64 evaluation of k r ~ 6e6 rad carries ~1e-9 rad)
  FP32 direct (MODE_FLOAT)   vs exact FP64 oracle          : 1e-5
  mask (set of contributing pixels)                        : bit-exact
  8-bit phase image vs the reference restatement           : <= 1 LSB on >= 99.9 %
"""
import json
import os

import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "single_point.json")
FP32 = capi.MODE_GPU | capi.MODE_FLOAT
FP64 = capi.MODE_GPU


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _one(x, y, z, a, phase=0.0):
    v = np.zeros(1, synth.VERTEX_DTYPE)
    v["pos"][0] = (x, y, z)
    v["color"][0] = (a, a, a)
    v["phase"][0] = phase
    return v


# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [FP64, FP32])
def test_golden_single_point(ctx, mode):
    data = json.load(open(GOLDEN))
    for c in data["cases"]:
        v = _one(*c["p"], c["a"])
        f = capi.generate(ctx, v, c["pn"], c["pp"], [c["lam"]], c["mode"], mode, offset=c["off"])[0]
        tol = 2e-9 if mode == FP64 else 1e-5
        for col, row, re, im in c["values"]:
            want = complex(float(re), float(im))
            assert abs(f[row, col] - want) <= tol * abs(want), (c["name"], mode, col, row, f[row, col], want)


@pytest.mark.parametrize("diff", [0, 1])
@pytest.mark.parametrize("mode", [FP64, FP32])
def test_mask_exact_tiny_grid(ctx, orc, diff, mode):
    """Box edges and the RS range test on a 100x37 grid (not a tile multiple)
    with offsets: the set of written pixels equals the reference's."""
    pn, pp, off = (100, 37), (10e-6, 8e-6), (3, -2)
    g = orc.make_geom(pn, pp, offset=off, scale=(1.0, 1.0, 1.0), distance=0.5e-3)
    v = synth.random_cloud(101 + diff, 60, 350e-6, (1.0e-3, 3.5e-3))
    lam = 633e-9
    want = orc.gen_restate(g, v, [lam], diff)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], diff, mode, distance=0.5e-3, offset=off)[0]
    # per point masks: compare single-point fields' supports
    for i in range(0, 60, 7):
        w1 = orc.gen_restate(g, v[i:i + 1], [lam], diff)[0]
        g1 = capi.generate(ctx, v[i:i + 1], pn, pp, [lam], diff, mode, distance=0.5e-3, offset=off)[0]
        assert np.array_equal(w1 != 0, g1 != 0), (i, np.argwhere((w1 != 0) != (g1 != 0))[:5])
    tol = 1e-9 if mode == FP64 else 1e-5
    assert rel_l2(got, want) < tol


@pytest.mark.parametrize("diff", [0, 1])
def test_fp64_matches_restatement(ctx, orc, diff):
    pn, pp = (160, 96), (8e-6, 8e-6)
    g = orc.make_geom(pn, pp, scale=(2.0, 2.0, 1.0), distance=0.02)
    v = synth.random_cloud(7 + diff, 300, 200e-6, (0.0, 0.01))
    lam = 473e-9
    want = orc.gen_restate(g, v, [lam], diff)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], diff, FP64, scale=(2.0, 2.0, 1.0), distance=0.02)[0]
    assert np.array_equal(got != 0, want != 0)
    assert rel_l2(got, want) < 1e-9


@pytest.mark.parametrize("diff", [0, 1])
def test_fp32_realistic_geometry_vs_exact(ctx, orc, diff):
    """1920x1080 at 8 um, lambda 532 nm, z ~ 0.5 m (the 10k-point RS workload's
    geometry) with a 4,000-point subset: FP32 field vs the extended-precision
    oracle on 3000 sampled pixels."""
    w = synth.WORKLOADS[1]
    v = synth.points(w, 4000)
    g = orc.make_geom(w.pn, w.pp)
    got = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], diff, FP32)[0].ravel()
    pix = np.random.default_rng(3).choice(w.n_pixels, 3000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, diff, pix)
    assert rel_l2(got[pix], want) < 1e-5


@pytest.mark.parametrize("diff", [0, 1])
def test_fp64_realistic_geometry_vs_exact(ctx, orc, diff):
    """FP64 mode at the 10k-point RS workload's geometry (1920x1080 at 8 um,
    532 nm, z ~ 0.5 m), 3,000 points: every (point, tile) pair takes the
    separable form (column factor x row factor, and for RS the short series
    correction).  Field vs the extended-precision oracle on 3000 pixels."""
    w = synth.WORKLOADS[1]
    v = synth.points(w, 3000)
    g = orc.make_geom(w.pn, w.pp)
    got = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], diff, FP64)[0].ravel()
    pix = np.random.default_rng(13).choice(w.n_pixels, 3000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, diff, pix)
    assert rel_l2(got[pix], want) < 1e-11


@pytest.mark.parametrize("diff", [0, 1])
def test_fp64_near_and_far_points_in_one_launch(ctx, orc, diff):
    """Near points (z 2-6 mm: partial masks; for RS the series is not accurate
    there, so their pairs take the per-pixel exact list) and far points
    (z ~ 0.3 m: separable form) interleaved in one cloud, so both lists share
    tiles.  Supports equal the reference restatement's; values vs the
    extended-precision oracle."""
    pn, pp, lam = (300, 170), (6e-6, 6e-6), 532e-9
    near = synth.random_cloud(41 + diff, 40, 400e-6, (2e-3, 6e-3))
    far = synth.random_cloud(43 + diff, 40, 400e-6, (0.30, 0.31))
    v = np.concatenate([near[:20], far[:20], near[20:], far[20:]])
    g = orc.make_geom(pn, pp)
    want = orc.gen_restate(g, v, [lam], diff)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], diff, FP64)[0]
    assert np.array_equal(got != 0, want != 0)
    assert rel_l2(got, want) < 2e-9          # the restatement's own k r rounding
    pix = np.random.default_rng(17).choice(pn[0] * pn[1], 3000, replace=False)
    exact = orc.field_exact(g, v, lam, 0, diff, pix)
    assert rel_l2(got.ravel()[pix], exact) < 1e-11


@pytest.mark.parametrize("diff", [0, 1])
def test_fp32_near_and_far_points_in_one_launch(ctx, orc, diff):
    """FP32 mode with near points (z 2-6 mm: their pairs need FP64 terms,
    which the FP32 kernel leaves to its follow-up kernel) and far points
    (separable FP32 form) interleaved in one cloud.  Supports equal the
    reference restatement's; values vs the extended-precision oracle."""
    pn, pp, lam = (300, 170), (6e-6, 6e-6), 532e-9
    near = synth.random_cloud(51 + diff, 40, 400e-6, (2e-3, 6e-3))
    far = synth.random_cloud(53 + diff, 40, 400e-6, (0.30, 0.31))
    v = np.concatenate([near[:20], far[:20], near[20:], far[20:]])
    g = orc.make_geom(pn, pp)
    want = orc.gen_restate(g, v, [lam], diff)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], diff, FP32)[0]
    assert np.array_equal(got != 0, want != 0)
    pix = np.random.default_rng(19).choice(pn[0] * pn[1], 3000, replace=False)
    exact = orc.field_exact(g, v, lam, 0, diff, pix)
    assert rel_l2(got.ravel()[pix], exact) < 1e-5
    if diff == 0:
        # RS near points alone: the series bound fails on every tile, so every
        # pair is an FP64 term of the follow-up kernel
        got_near = capi.generate(ctx, near, pn, pp, [lam], diff, FP32)[0]
        exact_near = orc.field_exact(g, near, lam, 0, diff, pix)
        assert rel_l2(got_near.ravel()[pix], exact_near) < 1e-11


@pytest.mark.parametrize("mode", [FP32, FP64])
@pytest.mark.parametrize("diff", [0, 1])
def test_point_splits(ctx, orc, diff, mode):
    """1920x1200 is 1140 tiles of 64x32 (FP32) or 1125 of 128x16 (FP64):
    with 2, 3 or 4 resident workgroups per CU
    (512-1024 at once) the last round would run part empty, so the launch
    splits each tile's points into parts (partial planes added in split
    order).  The split launch matches the extended-precision oracle and is
    deterministic."""
    w = synth.WORKLOADS[1]
    pn = (1920, 1200)
    v = synth.points(w, 3000)
    g = orc.make_geom(pn, w.pp)
    got = capi.generate(ctx, v, pn, w.pp, [w.lambdas[0]], diff, mode)[0].ravel()
    st = ctx.last_stats()
    tiles = (1920 // 64) * (1200 // 32 + 1) if mode == FP32 else (1920 // 128) * (1200 // 16)
    assert st.workgroups % tiles == 0 and st.workgroups > tiles      # really split
    again = capi.generate(ctx, v, pn, w.pp, [w.lambdas[0]], diff, mode)[0].ravel()
    assert np.array_equal(got, again)
    pix = np.random.default_rng(23).choice(pn[0] * pn[1], 2000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, diff, pix)
    assert rel_l2(got[pix], want) < (1e-5 if mode == FP32 else 1e-11)


@pytest.mark.parametrize("diff", [0, 1])
def test_probe_matches_exact_oracle(ctx, orc, diff):
    w = synth.WORKLOADS[2]
    v = synth.points(w, 500)
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(5).choice(w.n_pixels, 400, replace=False)
    dv = ctx.to_device(v)
    try:
        got = ctx.probe_f64(dv, len(v), capi.params(w.pn, w.pp, w.lambdas[1], channel=1), diff, pix)
    finally:
        dv.free()
    want = orc.field_exact(g, v, w.lambdas[1], 1, diff, pix)
    assert rel_l2(got, want) < 1e-12


def test_complex_amplitude_extension(ctx, orc):
    """use_phase: amplitude color*exp(j phase) (the 1M-point Fresnel
    workload's complex amplitudes), first 5,000 points."""
    w = synth.WORKLOADS[3]
    v = synth.points(w, 5000)
    g = orc.make_geom(w.pn, w.pp)
    got = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], 1, FP32, use_phase=True)[0].ravel()
    pix = np.random.default_rng(9).choice(w.n_pixels, 2000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, 1, pix, use_phase=True)
    assert rel_l2(got[pix], want) < 1e-5


@pytest.mark.parametrize("diff", [0, 1])
def test_box_pixel_count(ctx, orc, diff):
    """ohx_last_launch_stats().pairs_in_box equals the sum over points of the
    reference box areas (ophGen.cpp:384-407,467-490), on partial masks; and
    for Fresnel, whose mask is the box, the count of written pixels of a
    one-point field."""
    pn, pp, off, dist = (100, 37), (10e-6, 8e-6), (3, -2), 0.5e-3
    g = orc.make_geom(pn, pp, offset=off, scale=(1.0, 1.0, 1.0), distance=dist)
    v = synth.random_cloud(301 + diff, 45, 350e-6, (1.0e-3, 3.5e-3))
    lam = 633e-9
    want = 0
    for q in v:
        b = orc.box(g, q["pos"], lam, diff)     # the oracle adds the geometry distance
        want += (b[1] - b[0]) * (b[3] - b[2])
    assert 0 < want < len(v) * pn[0] * pn[1]          # partial masks
    capi.generate(ctx, v, pn, pp, [lam], diff, FP32, distance=dist, offset=off)
    st = ctx.last_stats()
    assert st.pairs_nominal == len(v) * pn[0] * pn[1]
    assert st.pairs_in_box == want
    if diff == 1:
        f1 = capi.generate(ctx, v[:1], pn, pp, [lam], diff, FP32, distance=dist, offset=off)[0]
        assert ctx.last_stats().pairs_in_box == int((f1 != 0).sum())


def test_partial_masks_large_geometry(ctx, orc):
    """7680x4320 at 3.74 um, lambda 473 nm, z down to 0.3 m: single-point
    masks are partial, so tiles straddling a mask edge take the per-pixel
    mask path.  Each point's support is bit-exact; values within FP32."""
    w = synth.WORKLOADS[5]
    g = orc.make_geom(w.pn, w.pp)
    lam = w.lambdas[2]
    v = np.concatenate([synth.points(w, 3), _one(11.0e-3, -9.5e-3, 0.30, 0.8),
                        _one(-10.2e-3, 7.7e-3, 0.33, 0.5)])
    partial = 0
    for diff in (0, 1):
        for i in range(len(v)):
            want = orc.gen_restate(g, v[i:i + 1], [lam], diff)[0]
            got = capi.generate(ctx, v[i:i + 1], w.pn, w.pp, [lam], diff, FP32)[0]
            nz_w = want != 0
            partial += int(nz_w.sum() < nz_w.size)
            assert np.array_equal(nz_w, got != 0), (diff, i)
            assert rel_l2(got, want) < 1e-5, (diff, i)
    assert partial >= 4


# ---------------------------------------------------------------------------
def test_encode_and_normalize_match_reference(ctx, orc):
    rng = np.random.default_rng(17)
    npx = 70_001
    fields = [rng.normal(size=npx) + 1j * rng.normal(size=npx) for _ in range(3)]
    fields[0][:4] = [1, 1j, -1, -1j]
    for flag in (capi.ENCODE_PHASE, capi.ENCODE_AMPLITUDE, capi.ENCODE_REAL, capi.ENCODE_IMAGINARY):
        dfs = [ctx.to_device(f.astype(np.complex128)) for f in fields]
        encs = [ctx.alloc(8 * npx) for _ in fields]
        u8s = [ctx.alloc(npx) for _ in fields]
        try:
            for df, de in zip(dfs, encs):
                ctx.encode(df, npx, de, flag)
            got_e = [de.download(np.float64, npx) for de in encs]
            ctx.normalize_u8(encs, npx, u8s)
            got_u = [du.download(np.uint8, npx) for du in u8s]
        finally:
            for b in dfs + encs + u8s:
                b.free()
        want_e = [orc.encode(f, flag) for f in fields]
        want_u = orc.normalize(want_e)
        for ge, we in zip(got_e, want_e):
            assert np.allclose(ge, we, rtol=4e-16, atol=4e-16)
        for gu, wu in zip(got_u, want_u):
            assert np.abs(gu.astype(int) - wu.astype(int)).max() <= 1
            assert (gu == wu).mean() > 0.9999


@pytest.mark.parametrize("flag", [0, 1, 2, 3])
@pytest.mark.parametrize("npx,nch", [(1, 1), (7, 2), (70_001, 3), (1 << 20, 3), (3_000_001, 1)])
def test_fused_encode_normalize_equals_two_calls(ctx, orc, flag, npx, nch):
    """ohx_encode_normalize_u8 (encode pass that also takes the min/max) writes
    exactly what ohx_encode per channel followed by ohx_normalize_u8 writes:
    encoded planes and 8-bit planes bit for bit, NaN pixels skipped by the
    min/max as the reference's comparisons skip them; the 8-bit planes also
    equal the oracle's normaliser on the same encoded planes."""
    rng = np.random.default_rng(npx + 10 * flag)
    fields = [(rng.normal(size=npx) + 1j * rng.normal(size=npx)) * (ch + 1) for ch in range(nch)]
    if npx > 10:
        fields[0][3] = complex(float("nan"), 1.0)
    dfs = [ctx.to_device(f.astype(np.complex128)) for f in fields]
    e1 = [ctx.alloc(8 * npx) for _ in fields]
    u1 = [ctx.alloc(npx) for _ in fields]
    e2 = [ctx.alloc(8 * npx) for _ in fields]
    u2 = [ctx.alloc(npx) for _ in fields]
    try:
        for df, de in zip(dfs, e1):
            ctx.encode(df, npx, de, flag)
        ctx.normalize_u8(e1, npx, u1)
        ctx.encode_normalize_u8(dfs, npx, flag, e2, u2)
        for a, b in zip(e1, e2):
            x, y = a.download(np.float64, npx), b.download(np.float64, npx)
            assert np.array_equal(x.view(np.uint64), y.view(np.uint64))
        got = [b.download(np.uint8, npx) for b in u2]
        for a, g in zip(u1, got):
            assert np.array_equal(a.download(np.uint8, npx), g)
        encs = [b.download(np.float64, npx) for b in e2]
        want = orc.normalize(encs)
        for g, w_, e in zip(got, want, encs):   # the oracle's uint8 cast of NaN is undefined
            ok = np.isfinite(e)
            assert np.array_equal(g[ok], w_[ok])
    finally:
        for b in dfs + e1 + u1 + e2 + u2:
            b.free()


class _View:
    """A device pointer offset into a DeviceBuffer (for alignment cases)."""
    def __init__(self, buf, off):
        self.ptr = buf.ptr + off


@pytest.mark.parametrize("npx", [1, 3, 5, 1023, 70_001])
def test_normalize_unaligned_and_odd_sizes(ctx, orc, npx):
    """The vectorised min/max and u8 kernels against the oracle's normaliser on
    identical encoded inputs (bit-exact), for buffers that are 16-byte aligned
    and for encoded buffers 8 bytes / outputs 1 byte off alignment."""
    rng = np.random.default_rng(npx)
    encs = [rng.normal(size=npx) * 3 + 1 for _ in range(2)]
    want = orc.normalize(encs)
    for e_off, u_off in ((0, 0), (8, 1), (0, 3)):
        bufs = [ctx.alloc(8 * npx + 16) for _ in encs] + [ctx.alloc(npx + 16) for _ in encs]
        try:
            de = [_View(b, e_off) for b in bufs[:2]]
            du = [_View(b, u_off) for b in bufs[2:]]
            for d, e in zip(de, encs):
                capi.lib().ohx_memcpy_h2d(ctx.h, d.ptr, e.ctypes.data, 8 * npx)
            ctx.normalize_u8(de, npx, du)
            ctx.sync()
            for d, w in zip(du, want):
                got = np.empty(npx, np.uint8)
                capi.lib().ohx_memcpy_d2h(ctx.h, got.ctypes.data, d.ptr, npx)
                assert np.array_equal(got, w), (npx, e_off, u_off)
        finally:
            for b in bufs:
                b.free()


def test_end_to_end_u8_phase_image(ctx, orc):
    """generateHologram(RS) -> encoding(PHASE) -> normalize() against the
    reference restatement at the 10k-point RS workload's geometry (first 500
    points): <= 1 LSB on >= 99.9 % of pixels."""
    w = synth.WORKLOADS[1]
    v = synth.points(w, 500)
    g = orc.make_geom(w.pn, w.pp)
    ref_f = orc.gen_restate(g, v, [w.lambdas[0]], 0)[0]
    ref_u = orc.normalize([orc.encode(ref_f, orc.ENCODE_PHASE)])[0]
    npx = w.n_pixels
    dv = ctx.to_device(v)
    df, de, du = ctx.alloc(16 * npx), ctx.alloc(8 * npx), ctx.alloc(npx)
    try:
        ctx.field_zero(df, npx)
        ctx.diffract(dv, len(v), df, capi.params(w.pn, w.pp, w.lambdas[0]), 0, FP32)
        ctx.encode(df, npx, de, capi.ENCODE_PHASE)
        ctx.normalize_u8([de], npx, [du])
        got_u = du.download(np.uint8, npx).reshape(ref_u.shape)
    finally:
        for b in (dv, df, de, du):
            b.free()
    d = np.abs(got_u.astype(int) - ref_u.astype(int))
    d = np.minimum(d, 255 - d)   # phase wraps: 0 and 255 are neighbours
    assert (d <= 1).mean() >= 0.999
    assert (got_u == ref_u).mean() >= 0.99


# ---------------------------------------------------------------------------
def test_superposition_and_determinism_full_size(ctx):
    """3840x2160, 3 channels, the 100k-point workload's first 3000 points:
    field(A u B) == field(A) + field(B) to FP32 accuracy, and reruns are
    bit-identical."""
    w = synth.WORKLOADS[2]
    v = synth.points(w, 3000)
    for diff in (0, 1):
        ab = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], diff, FP32)[0]
        ab2 = capi.generate(ctx, v, w.pn, w.pp, [w.lambdas[0]], diff, FP32)[0]
        assert np.array_equal(ab, ab2)
        a = capi.generate(ctx, v[:1300], w.pn, w.pp, [w.lambdas[0]], diff, FP32)[0]
        b = capi.generate(ctx, v[1300:], w.pn, w.pp, [w.lambdas[0]], diff, FP32)[0]
        assert rel_l2(a + b, ab) < 2e-6


def test_accumulate_and_zero(ctx):
    n = 12_345
    a = (np.arange(n) + 1j * np.arange(n)[::-1]).astype(np.complex128)
    b = (0.5 * np.arange(n) - 2j).astype(np.complex128)
    da, db = ctx.to_device(a), ctx.to_device(b)
    try:
        ctx.accumulate(da, db, n)
        assert np.array_equal(da.download(np.complex128, n), a + b)
        ctx.field_zero(da, n)
        assert not da.download(np.complex128, n).any()
    finally:
        da.free()
        db.free()


def test_edge_cases(ctx):
    pn, pp = (64, 32), (8e-6, 8e-6)
    # no points: no-op; point far outside: empty mask
    far = _one(1.0, 1.0, 0.01, 1.0)
    for diff in (0, 1):
        f = capi.generate(ctx, far, pn, pp, [532e-9], diff, FP32)[0]
        assert not f.any()
        f = capi.generate(ctx, far[:0], pn, pp, [532e-9], diff, FP32)[0]
        assert not f.any()
    # invalid arguments are status codes, not crashes
    dv = ctx.to_device(far)
    fld = ctx.alloc(16 * 64 * 32)
    try:
        with pytest.raises(capi.OhxError) as e:
            ctx.diffract(dv, 1, fld, capi.params(pn, pp, 532e-9), 7, FP32)
        assert e.value.status == -1
        with pytest.raises(capi.OhxError):
            ctx.diffract(dv, 1, fld, capi.params(pn, pp, 532e-9, channel=3), 0, FP32)
        with pytest.raises(capi.OhxError):
            ctx.diffract(dv, 1, fld, capi.params(pn, (1e-7, 1e-7), 532e-9), 0, FP32)
        for bad in (capi.params(pn, (float("inf"), 8e-6), 532e-9), capi.params(pn, pp, float("nan")),
                    capi.params(pn, pp, 532e-9, distance=float("inf")),
                    capi.params(pn, pp, 532e-9, scale=(1.0, float("nan"), 1.0))):
            with pytest.raises(capi.OhxError) as e:
                ctx.diffract(dv, 1, fld, bad, 1, FP32)
            assert e.value.status == -1
    finally:
        dv.free()
        fld.free()


def _stream_busy(ctx):
    """hipStreamQuery on the context's stream: True while work is pending."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamQuery.argtypes = [ctypes.c_void_p]
    hip.hipStreamQuery.restype = ctypes.c_int
    st = hip.hipStreamQuery(capi.lib().ohx_ctx_stream(ctx.h))
    assert st in (0, 600), st          # hipSuccess or hipErrorNotReady
    return st == 600


def test_d2h_after_overlaps_next_launch(ctx):
    """ohx_memcpy_d2h_async_after orders the copy of field 0 after field 0's
    launch and before nothing later: a long launch enqueued after the copy
    (1e5 points into a 3840x2160 field, ~0.15 s) is still running when
    ohx_copy_synchronize returns the copied field.  Values equal plain
    synchronous read-backs, into pageable and page-locked memory."""
    pn, pp = (512, 256), (8e-6, 8e-6)
    npx = pn[0] * pn[1]
    v = synth.random_cloud(77, 3000, 1.0e-3, (0.2, 0.3))
    w = synth.WORKLOADS[2]
    vl = synth.points(w, 100_000)
    dv, dvl = ctx.to_device(v), ctx.to_device(vl)
    f0, f1 = ctx.alloc(16 * npx), ctx.alloc(16 * w.n_pixels)
    host = capi.HostBuffer(16 * npx)
    try:
        p0 = capi.params(pn, pp, 532e-9)
        pl = capi.params(w.pn, w.pp, w.lambdas[0])
        ctx.field_zero(f1, w.n_pixels)
        ctx.diffract(dvl, len(vl), f1, pl, 0, FP32)          # warm: scratch grown before timing
        ctx.sync()
        want0 = capi.generate(ctx, v, pn, pp, [532e-9], 0, FP32)[0].ravel()
        for pinned in (False, True):
            out = host.array(np.complex128, npx) if pinned else np.empty(npx, np.complex128)
            out[:] = 0
            ctx.field_zero(f0, npx)
            ctx.diffract(dv, len(v), f0, p0, 0, FP32)
            f0.download_async_after(out)
            ctx.field_zero(f1, w.n_pixels)
            ctx.diffract(dvl, len(vl), f1, pl, 0, FP32)      # enqueued after the copy
            ctx.copy_sync()
            if pinned:   # page-locked: the copy did not wait for the long launch
                assert _stream_busy(ctx)
            ctx.sync()
            assert np.array_equal(out, want0), pinned
            del out
        # the blocking form copies what was enqueued before the call
        early = np.empty(npx, np.complex128)
        ctx.field_zero(f0, npx)
        ctx.diffract(dv, len(v), f0, p0, 0, FP32)
        f0.download_after(early)
        assert np.array_equal(early, want0)
    finally:
        host.free()
        for b in (dv, dvl, f0, f1):
            b.free()


def test_fused_encode_normalize_one_signed_range(ctx):
    """Encoded values all negative (real part of a field with negative real
    parts) or all NaN: the min/max keys start from the reference normaliser's
    DBL_MAX / DBL_MIN in both the fused and the two-call path, so the 8-bit
    planes agree bit for bit there too."""
    npx = 100_000
    rng = np.random.default_rng(9)
    cases = [-(np.abs(rng.normal(size=npx)) + 0.1) + 1j * rng.normal(size=npx),
             np.full(npx, complex(float("nan"), float("nan")))]
    for f in cases:
        df = ctx.to_device(f.astype(np.complex128))
        e1, u1, e2, u2 = ctx.alloc(8 * npx), ctx.alloc(npx), ctx.alloc(8 * npx), ctx.alloc(npx)
        try:
            ctx.encode(df, npx, e1, capi.ENCODE_REAL)
            ctx.normalize_u8([e1], npx, [u1])
            ctx.encode_normalize_u8([df], npx, capi.ENCODE_REAL, [e2], [u2])
            assert np.array_equal(u1.download(np.uint8, npx), u2.download(np.uint8, npx))
        finally:
            for b in (df, e1, u1, e2, u2):
                b.free()
