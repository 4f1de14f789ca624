"""GPU parity of the RS matrix-core path (rank-2 tile expansion on MFMA, the
remaining points on the FP32 direct kernel), through the C ABI.

Field relative-L2 tolerances against the FP64 oracle, as for the Fresnel
matrix-core variants:
  BF16 1-term : 1e-2
  BF16 3-term : 1e-4
ohx_last_launch_stats().gemm_points names how many points ran on the matrix
cores; the tests check it against the eligibility each geometry implies
(points whose mask is their box and whose tile expansion holds).
"""
import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

FP32 = capi.MODE_GPU | capi.MODE_FLOAT
RS_MFMA = {
    "bf16x1": (FP32 | capi.MODE_MFMA_BF16X1, 1e-2),
    "bf16x3": (FP32 | capi.MODE_MFMA_BF16X3, 1e-4),
}


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(ctx, v, pn, pp, lam, mode, channel=0, pix=None):
    """One RS channel through ohx_pc_diffract: (field or field[pix], stats)."""
    npx = pn[0] * pn[1]
    dv = ctx.to_device(v)
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, len(v), fld, capi.params(pn, pp, lam, channel=channel), 0, mode)
        st = ctx.last_stats()
        f = fld.download(np.complex128, npx)
        return (f if pix is None else f[pix]), st
    finally:
        dv.free()
        fld.free()


@pytest.mark.parametrize("name", list(RS_MFMA))
def test_rs_gemm_2160p_rgb_vs_oracle(ctx, orc, name):
    """100k-point workload geometry (3840x2160 at 3.74 um, x, y in +-8 mm, z in
    [0.45, 0.55] m), first 10,000 points, each of 638/532/473 nm: every point's
    mask is the whole hologram, so all of them run on the matrix cores; field
    vs the oracle on 3,000 pixels."""
    mode, tol = RS_MFMA[name]
    w = synth.WORKLOADS[2]
    v = synth.points(w, 10_000)
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(500).choice(w.n_pixels, 3000, replace=False)
    for ch, lam in enumerate(w.lambdas):
        want = orc.field_exact(g, v, lam, ch, 0, pix)
        got, st = run(ctx, v, w.pn, w.pp, lam, mode, ch, pix)
        err = rel_l2(got, want)
        print(f"{name} ch={ch}: rel-L2 {err:.3e}, {st.gemm_points} of {len(v)} points on matrix cores")
        assert st.gemm_points == len(v) and st.kernel_kind == 2
        assert err < tol, (ch, err)


@pytest.mark.parametrize("name", list(RS_MFMA))
@pytest.mark.parametrize("n", [1, 9, 1999])
def test_rs_gemm_mixed_eligibility_vs_oracle(ctx, orc, name, n):
    """300x170 at 8 um (not tile multiples), z in [0.15, 0.3] m: the nearest
    points' tile expansion is out of bounds on some of their tiles, which the
    direct kernel takes while the masked matrix-core instance takes the
    others; n = 1 and 9 leave partial 8-point stages."""
    mode, tol = RS_MFMA[name]
    pn, pp, lam = (300, 170), (8e-6, 8e-6), 532e-9
    v = synth.random_cloud(31, 1999, 1.5e-3, (0.15, 0.3))[:n]
    g = orc.make_geom(pn, pp)
    pix = np.random.default_rng(2).choice(pn[0] * pn[1], 3000, replace=False)
    want = orc.field_exact(g, v, lam, 0, 0, pix)
    got, st = run(ctx, v, pn, pp, lam, mode, pix=pix)
    err = rel_l2(got, want)
    print(f"{name} n={n}: rel-L2 {err:.3e}, {st.gemm_points} on matrix cores")
    assert err < tol
    assert 0 < st.gemm_points <= n


@pytest.mark.parametrize("name", list(RS_MFMA))
def test_rs_gemm_partial_masks_go_direct(ctx, orc, name):
    """Points a few mm from the hologram have boxes a few pixels wide: they run
    on the direct kernel (exact masks), the far points on the matrix cores.
    Whole field vs the reference restatement."""
    mode, tol = RS_MFMA[name]
    pn, pp, lam = (150, 90), (10e-6, 10e-6), 633e-9
    g = orc.make_geom(pn, pp, offset=(2, -3))
    near = synth.random_cloud(41, 40, 500e-6, (1.5e-3, 4e-3))
    far = synth.random_cloud(42, 300, 500e-6, (0.2, 0.3))
    v = np.concatenate([near[:20], far, near[20:]])
    want = orc.gen_restate(g, v, [lam], 0)[0]
    got = capi.generate(ctx, v, pn, pp, [lam], 0, mode, offset=(2, -3))[0]
    st = ctx.last_stats()
    err = rel_l2(got, want)
    print(f"{name}: rel-L2 {err:.3e}, {st.gemm_points} on matrix cores")
    assert st.gemm_points == len(far)
    assert err < tol
    # the near points alone: no point on the matrix cores, bit-identical to FP32
    a = capi.generate(ctx, near, pn, pp, [lam], 0, mode, offset=(2, -3))[0]
    assert ctx.last_stats().gemm_points == 0
    b = capi.generate(ctx, near, pn, pp, [lam], 0, FP32, offset=(2, -3))[0]
    assert np.array_equal(a, b)


def test_rs_gemm_deterministic_and_close_to_direct(ctx):
    """Same input twice: bit-identical fields; and within 3e-5 of the FP32
    direct kernel on the whole 1920x1080 field (10k-point RS workload)."""
    w = synth.WORKLOADS[1]
    v = synth.points(w, 4000)
    mode = RS_MFMA["bf16x3"][0]
    a, sa = run(ctx, v, w.pn, w.pp, w.lambdas[0], mode)
    b, _ = run(ctx, v, w.pn, w.pp, w.lambdas[0], mode)
    assert np.array_equal(a, b)
    d, _ = run(ctx, v, w.pn, w.pp, w.lambdas[0], FP32)
    err = rel_l2(a, d)
    print(f"vs FP32 direct: rel-L2 {err:.3e}, {sa.gemm_points} of {len(v)} on matrix cores")
    assert sa.gemm_points > 0
    assert err < 3e-5


def test_rs_gemm_complex_amplitude_vs_oracle(ctx, orc):
    """Complex amplitudes (Vertex::phase used): the 1M-point workload's
    geometry and amplitude distribution, first 5,000 points, RS on the matrix
    cores vs the oracle with the phase term on 3,000 pixels."""
    mode, tol = RS_MFMA["bf16x3"]
    w = synth.WORKLOADS[3]
    v = synth.points(w, 5_000)
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(510).choice(w.n_pixels, 3000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, 0, pix, use_phase=True)
    npx = w.n_pixels
    dv = ctx.to_device(v)
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, len(v), fld, capi.params(w.pn, w.pp, w.lambdas[0], use_phase=True), 0, mode)
        st = ctx.last_stats()
        got = fld.download(np.complex128, npx)[pix]
    finally:
        dv.free()
        fld.free()
    err = rel_l2(got, want)
    print(f"complex amplitudes: rel-L2 {err:.3e}, {st.gemm_points} on matrix cores")
    assert st.gemm_points == len(v)
    assert err < tol


def test_rs_gemm_chunked_point_range(ctx):
    """5M points at 3840x2160: the RS split-K planes would exceed their 4 GiB
    cap, so the matrix-core part runs in two point chunks, each reduced into
    the field.  Against the library's FP64 probe within the BF16x3 bound."""
    mode, tol = RS_MFMA["bf16x3"]
    w = synth.WORKLOADS[2]
    n = 5_000_000
    v = synth.random_cloud(78, n, 8e-3, (0.45, 0.55))
    p = capi.params(w.pn, w.pp, w.lambdas[1])
    dv = ctx.to_device(v)
    npx = w.n_pixels
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, n, fld, p, 0, mode)
        st = ctx.last_stats()
        pix = np.random.default_rng(6).choice(npx, 128, replace=False)
        got = fld.download(np.complex128, npx)[pix]
        want = ctx.probe_f64(dv, n, p, 0, pix)
    finally:
        dv.free()
        fld.free()
    assert st.kernel_kind == 2 and st.gemm_points == n
    assert st.workgroups == 2 * 510 * 39   # two chunks of 2.5M points, 39 splits each
    err = rel_l2(got, want)
    print(f"chunked RS bf16x3: rel-L2 {err:.3e}")
    assert err < tol


def test_rs_gemm_masked_points_exact_mask(ctx, orc):
    """7680x4320 at 3.74 um, 473 nm: points of the 2M-point workload whose box
    ends inside the hologram run on the masked matrix-core instance (on the
    tiles where the reference's range test holds with margin and the
    expansion bound holds; the direct kernel takes the others).  Three such
    points, each alone: the nonzero pixels are exactly those of the FP32
    direct kernel, which applies the reference's per-pixel mask, and the
    values match the oracle on pixels inside and on both sides of every box
    edge and on pixels of the box that the range test excludes."""
    mode, tol = RS_MFMA["bf16x3"]
    w = synth.WORKLOADS[5]
    ch = 2
    lam = w.lambdas[ch]
    v = synth.points(w, 3000)
    g = orc.make_geom(w.pn, w.pp)
    pnx, pny = w.pn
    rng = np.random.default_rng(520)
    found = cut = 0
    for i in range(len(v)):
        x0, x1, y0, y1 = orc.box(g, v[i]["pos"], lam, 0)
        if (x0, x1, y0, y1) == (0, pnx, 0, pny) or x1 <= x0 or y1 <= y0:
            continue
        f, st = run(ctx, v[i:i + 1], w.pn, w.pp, lam, mode, ch)
        if st.gemm_points != 1:
            continue   # the expansion holds on none of this point's tiles: direct kernel
        d, _ = run(ctx, v[i:i + 1], w.pn, w.pp, lam, FP32, ch)
        assert np.array_equal(f != 0, d != 0), i
        f = f.reshape(pny, pnx)
        inbox_zero = np.flatnonzero((f[y0:y1, x0:x1] == 0).ravel())   # excluded by the range test
        cut += len(inbox_zero) > 0
        rows = [y0, y1 - 1] + ([y0 - 1] if y0 > 0 else []) + ([y1] if y1 < pny else [])
        cols = [x0, x1 - 1] + ([x0 - 1] if x0 > 0 else []) + ([x1] if x1 < pnx else [])
        pix = [int(r) * pnx + int(c) for r in rows for c in rng.integers(max(x0 - 1, 0), min(x1 + 1, pnx), 40)]
        pix += [int(r) * pnx + int(c) for c in cols for r in rng.integers(max(y0 - 1, 0), min(y1 + 1, pny), 40)]
        pix += [int(r) * pnx + int(c) for r, c in zip(rng.integers(y0, y1, 200), rng.integers(x0, x1, 200))]
        for k in rng.choice(inbox_zero, min(50, len(inbox_zero)), replace=False) if len(inbox_zero) else []:
            pix.append((y0 + int(k) // (x1 - x0)) * pnx + x0 + int(k) % (x1 - x0))
        pix = np.unique(np.array(pix, np.int64))
        want = orc.field_exact(g, v[i:i + 1], lam, ch, 0, pix)
        got = f.ravel()[pix]
        assert np.array_equal(got != 0, want != 0), i
        err = rel_l2(got, want)
        print(f"point {i}: box {x0}-{x1} x {y0}-{y1}, {len(inbox_zero)} box pixels outside the mask, "
              f"rel-L2 {err:.3e} on {len(pix)} pixels")
        assert err < tol, (i, err)
        found += 1
        if found == 3:
            break
    assert found == 3


def _tile_bound_ratio(pp, lam, dxc, dyc, z, theta):
    """Largest of the three expansion-bound ratios (left side / limit) of the
    RS matrix-core path on a 128 x 128 tile whose reference pixel is dxc, dyc
    from the point (the quantities rs_bound_ok compares)."""
    um = vm = 64 * pp
    nx, ny = um * (2 * dxc + um), vm * (2 * dyc + vm)
    zz, sn, il = z * z, nx + ny, 1.0 / lam
    return max(sn ** 3 * il / (16 * 2e-7 * zz * zz * z), sn * sn / (1e-7 * zz * zz),
               0.125 * 4 * np.pi * il * nx * ny / (theta * zz * z))


def test_rs_gemm_split_points_share_tiles_with_direct(ctx, orc):
    """7680x4320 at 3.74 um, 473 nm: points near z = 0.3 m far off axis (boxes
    smaller than the hologram) whose expansion bound fails on their farthest
    tiles but holds on the nearest.
    The masked matrix-core instance takes the tiles where it holds and the
    direct kernel the others.  Each point alone: counted as a matrix-core
    point, every 128 x 128 tile within 1e-4 of the FP32 direct kernel (a tile
    dropped or added twice would be off by its whole value), and the field
    within the BF16x3 bound of the oracle on pixels of both kinds of tile."""
    mode, tol = RS_MFMA["bf16x3"]
    w = synth.WORKLOADS[5]
    ch, lam, pp = 2, w.lambdas[2], w.pp[0]
    pnx, pny = w.pn
    g = orc.make_geom(w.pn, w.pp)
    v = synth.points(w, 20000)
    xs = -pnx * pp / 2 + (np.arange(pnx // 128) * 128 + 63) * pp        # tile reference pixels
    ys = -pny * pp / 2 + (pny - (np.arange((pny + 127) // 128) * 128 + 64)) * pp
    rng = np.random.default_rng(530)
    found = 0
    for i in np.argsort(v["pos"][:, 2]):
        px, py, z = (float(c) for c in v["pos"][i])
        x0, x1, y0, y1 = orc.box(g, v[i]["pos"], lam, 0)
        if x1 <= x0 or y1 <= y0:
            continue
        dx, dy = np.abs(xs - px), np.abs(ys - py)
        if not (_tile_bound_ratio(pp, lam, dx.max(), dy.max(), z, 8e-3) > 1.5 and
                _tile_bound_ratio(pp, lam, dx.min(), dy.min(), z, 8e-3) < 0.5):
            continue
        a, st = run(ctx, v[i:i + 1], w.pn, w.pp, lam, mode, ch)
        if st.gemm_points != 1:
            continue   # the range test has no margin on this box: direct kernel only
        b, _ = run(ctx, v[i:i + 1], w.pn, w.pp, lam, FP32, ch)
        ta = a.reshape(pny, pnx)[:pny // 128 * 128].reshape(pny // 128, 128, pnx // 128, 128)
        tb = b.reshape(pny, pnx)[:pny // 128 * 128].reshape(pny // 128, 128, pnx // 128, 128)
        nb = np.linalg.norm(tb, axis=(1, 3))
        assert np.array_equal(np.linalg.norm(ta, axis=(1, 3)) > 0, nb > 0), i   # same tiles touched
        err = np.linalg.norm(ta - tb, axis=(1, 3))[nb > 0] / nb[nb > 0]
        ratios = np.array([[_tile_bound_ratio(pp, lam, abs(xc - px), abs(yc - py), z, 8e-3) for xc in xs]
                           for yc in ys[:pny // 128]])
        print(f"point {i}: z {z:.4f}, tiles failing the bound {np.mean(ratios > 1):.3f}, "
              f"largest per-tile rel-L2 vs FP32 direct {err.max():.2e}")
        assert err.max() < 1e-4, i
        pix = rng.choice(pnx * pny, 3000, replace=False)
        want = orc.field_exact(g, v[i:i + 1], lam, ch, 0, pix)
        assert rel_l2(a[pix], want) < tol, i
        found += 1
        if found == 2:
            break
    assert found == 2
