"""GPU parity at the benchmark workloads' full geometries, against the
extended-precision oracle (oracle/oph_oracle.c orc_field_exact) on sampled
pixels, through the C ABI and the drop-in CLI.

Workloads (openholo_amd/synth.py):
  * 100k points, 3840x2160 at 3.74 um, RGB 638/532/473 nm, RS and Fresnel:
    FP32 direct kernels, first 10,000 points, all three channels.
  * 1M points with complex amplitudes, 3840x2160, 532 nm: the matrix-core
    variants and the FP32 direct kernel, first 20,000 points.
  * 2M points, 7680x4320, RGB, x, y in +-12 mm, z in [0.3, 0.7] m (partial
    masks): RS FP32 direct and Fresnel FP16 3-term, first 2,000 points, on
    random pixels plus pixels on the mask-box edges of the points whose box
    ends inside the hologram.
  * the reference example's Dice point cloud and configuration
    (Example/dataset/Dice/PointCloud: 7,493 points, 1920x1080 at 4.3 um,
    532 nm, z ~ 0.1 m, partial RS masks) through ophpc_cli.

Tolerances (field relative L2 vs the oracle), as stated for each kernel:
  FP32 direct 1e-5; BF16 1-term 1e-2; BF16 3-term 1e-4; FP16 3-term 2e-5;
  FP64 direct 1e-9.  Encoded buffers 4e-16 absolute against the oracle's
  encoder on the same field; 8-bit output exact against the oracle's
  normaliser on the same encoded buffers.
"""
import os
import subprocess

import numpy as np
import pytest

from openholo_amd import capi, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "openholo_amd", "bin", "ophpc_cli")
FP32 = capi.MODE_GPU | capi.MODE_FLOAT
FP64 = capi.MODE_GPU
MFMA = {
    "bf16x1": (FP32 | capi.MODE_MFMA_BF16X1, 1e-2),
    "bf16x3": (FP32 | capi.MODE_MFMA_BF16X3, 1e-4),
    "fp16x3": (FP32 | capi.MODE_MFMA_FP16X3, 2e-5),
}
VERTEX = np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def field_at(ctx, v, pn, pp, lam, channel, diff, mode, pix, use_phase=False):
    """One channel generated through ohx_pc_diffract; the field at pix."""
    npx = pn[0] * pn[1]
    dv = ctx.to_device(v)
    fld = ctx.alloc(16 * npx)
    try:
        ctx.field_zero(fld, npx)
        ctx.diffract(dv, len(v), fld, capi.params(pn, pp, lam, channel=channel, use_phase=use_phase), diff, mode)
        return fld.download(np.complex128, npx)[pix]
    finally:
        dv.free()
        fld.free()


@pytest.mark.parametrize("diff", [0, 1])
def test_rgb_2160p_fp32_direct_vs_oracle(ctx, orc, diff):
    """100k-point workload geometry, 10,000 points, each channel at full
    3840x2160: FP32 direct field vs the oracle on 3,000 pixels; and the
    library's FP64 probe (the reference behind the benchmark's rel-L2
    figures) vs the oracle at the same pixels."""
    w = synth.WORKLOADS[2]
    v = synth.points(w, 10_000)
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(200 + diff).choice(w.n_pixels, 3000, replace=False)
    dv = ctx.to_device(v)
    try:
        for ch, lam in enumerate(w.lambdas):
            want = orc.field_exact(g, v, lam, ch, diff, pix)
            got = field_at(ctx, v, w.pn, w.pp, lam, ch, diff, FP32, pix)
            err = rel_l2(got, want)
            print(f"diff={diff} ch={ch}: FP32 direct rel-L2 {err:.3e}")
            assert err < 1e-5, (ch, err)
            probe = ctx.probe_f64(dv, len(v), capi.params(w.pn, w.pp, lam, channel=ch), diff, pix)
            assert rel_l2(probe, want) < 1e-12, ch
    finally:
        dv.free()


def test_1m_2160p_matrix_core_and_direct_vs_oracle(ctx, orc):
    """1M-point complex-amplitude workload geometry, first 20,000 points:
    every matrix-core variant and the FP32 direct kernel vs the oracle on
    3,000 pixels."""
    w = synth.WORKLOADS[3]
    v = synth.points(w, 20_000)
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(300).choice(w.n_pixels, 3000, replace=False)
    want = orc.field_exact(g, v, w.lambdas[0], 0, 1, pix, use_phase=True)
    errs = {}
    for name, (mode, tol) in list(MFMA.items()) + [("direct", (FP32, 1e-5))]:
        got = field_at(ctx, v, w.pn, w.pp, w.lambdas[0], 0, 1, mode, pix, use_phase=True)
        errs[name] = rel_l2(got, want)
        print(f"{name}: rel-L2 {errs[name]:.3e}")
    for name, (mode, tol) in list(MFMA.items()) + [("direct", (FP32, 1e-5))]:
        assert errs[name] < tol, (name, errs[name])


def _edge_pixels(orc, g, v, lam, diff, pn, rng, per_point=24, max_points=60):
    """Pixels just inside and just outside the mask-box edges of points whose
    box ends inside the hologram."""
    out = []
    used = 0
    for q in v:
        x0, x1, y0, y1 = orc.box(g, q["pos"], lam, diff)
        if x1 <= x0 or y1 <= y0:
            continue
        cols = ([x0 - 1, x0] if x0 > 0 else []) + ([x1 - 1, x1] if x1 < pn[0] else [])
        rows = ([y0 - 1, y0] if y0 > 0 else []) + ([y1 - 1, y1] if y1 < pn[1] else [])
        if not cols and not rows:
            continue
        for _ in range(per_point // 2):
            if cols:
                out.append(int(rng.integers(y0, y1)) * pn[0] + cols[int(rng.integers(len(cols)))])
            if rows:
                out.append(rows[int(rng.integers(len(rows)))] * pn[0] + int(rng.integers(x0, x1)))
        used += 1
        if used >= max_points:
            break
    return np.array(out, np.int64), used


@pytest.mark.parametrize("diff", [0, 1])
def test_4320p_partial_masks_vs_oracle(ctx, orc, diff):
    """2M-point workload geometry (7680x4320, x, y in +-12 mm, z in
    [0.3, 0.7] m), first 2,000 points, all three channels: RS through the
    FP32 direct kernel, Fresnel through the FP16 3-term matrix-core kernel,
    as that workload runs them.  Pixels: 2,500 random plus pixels on the
    box edges of points whose mask ends inside the hologram (the 473 nm box
    of a point at z = 0.3 m is +-19 mm wide against a 28.7 mm hologram)."""
    w = synth.WORKLOADS[5]
    v = synth.points(w, 2_000)
    g = orc.make_geom(w.pn, w.pp)
    mode, tol = (FP32, 1e-5) if diff == 0 else MFMA["fp16x3"]
    rng = np.random.default_rng(500 + diff)
    for ch, lam in enumerate(w.lambdas):
        edge, n_partial = _edge_pixels(orc, g, v, lam, diff, w.pn, rng)
        if ch == 2:
            assert n_partial >= 10 and len(edge) >= 200, (n_partial, len(edge))
        pix = np.unique(np.concatenate([rng.choice(w.n_pixels, 2500, replace=False), edge]))
        want = orc.field_exact(g, v, lam, ch, diff, pix)
        got = field_at(ctx, v, w.pn, w.pp, lam, ch, diff, mode, pix)
        err = rel_l2(got, want)
        err_edge = rel_l2(got[np.isin(pix, edge)], want[np.isin(pix, edge)]) if len(edge) else 0.0
        print(f"diff={diff} ch={ch}: rel-L2 {err:.3e} (box-edge pixels {err_edge:.3e}, "
              f"{n_partial} points with partial masks)")
        assert err < tol, (ch, err)
        if len(edge):
            assert err_edge < tol, (ch, err_edge)


def test_4320p_rs_matrix_core_vs_oracle(ctx, orc):
    """2M-point workload geometry, first 2,000 points, all three channels, RS
    with MODE_MFMA_BF16X3: points whose tile expansion holds (most of them;
    those whose box ends inside the hologram on the masked instance) run on
    the matrix cores, the rest (z near 0.3 m far off axis) on the FP32 direct
    kernel.  Same pixels as the direct-kernel case (random plus box edges)."""
    w = synth.WORKLOADS[5]
    v = synth.points(w, 2_000)
    g = orc.make_geom(w.pn, w.pp)
    mode, tol = MFMA["bf16x3"]
    rng = np.random.default_rng(500)
    for ch, lam in enumerate(w.lambdas):
        edge, n_partial = _edge_pixels(orc, g, v, lam, 0, w.pn, rng)
        pix = np.unique(np.concatenate([rng.choice(w.n_pixels, 2500, replace=False), edge]))
        want = orc.field_exact(g, v, lam, ch, 0, pix)
        npx = w.n_pixels
        dv = ctx.to_device(v)
        fld = ctx.alloc(16 * npx)
        try:
            ctx.field_zero(fld, npx)
            ctx.diffract(dv, len(v), fld, capi.params(w.pn, w.pp, lam, channel=ch), 0, mode)
            st = ctx.last_stats()
            got = fld.download(np.complex128, npx)[pix]
        finally:
            dv.free()
            fld.free()
        err = rel_l2(got, want)
        sel = np.isin(pix, edge)
        err_edge = rel_l2(got[sel], want[sel]) if sel.any() else 0.0
        print(f"ch={ch}: rel-L2 {err:.3e} (box-edge pixels {err_edge:.3e}), "
              f"{st.gemm_points} of {len(v)} points on matrix cores")
        assert 0.9 * len(v) < st.gemm_points <= len(v)
        assert err < tol, (ch, err)
        if sel.any():
            assert err_edge < tol, (ch, err_edge)


@pytest.mark.parametrize("diff", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_dice_example_end_to_end(tmp_path, orc, diff, prec):
    """The reference example's own input (Example/dataset/Dice/PointCloud:
    dice_100000.ply with Dice_PointCloud.xml) through ophpc_cli, FP32
    (-m 0) and FP64 (-m 1), RS and Fresnel: field vs the oracle on 3,000
    pixels, encoded buffers vs the oracle's encoder, 8-bit output vs the
    oracle's normaliser.  At z ~ 0.1 m the RS box (+-6.3 mm) and the
    Fresnel box (+-6.2 mm) end inside the 8.3 x 4.6 mm hologram for part
    of the points, so partial masks are exercised."""
    d = tmp_path
    r = subprocess.run([CLI, "-a", "0", "-c", os.path.join(GOLD, "Dice_PointCloud.xml"),
                        "-i", os.path.join(GOLD, "dice_100000.ply"), "-e", "0", "-m", str(prec), "-p", "1",
                        "-f", str(diff), "-o", str(d / "R"), "--dump-points", str(d / "p.bin"),
                        "--dump-field", str(d / "f.bin"), "--dump-enc", str(d / "e.bin"),
                        "--dump-u8", str(d / "u.bin")], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    pn, pp, lam = (1920, 1080), (4.3e-6, 4.3e-6), 532e-9
    npx = pn[0] * pn[1]
    v = np.fromfile(d / "p.bin", VERTEX)
    assert len(v) == 7493
    f = np.fromfile(d / "f.bin", np.complex128)
    e = np.fromfile(d / "e.bin", np.float64)
    u = np.fromfile(d / "u.bin", np.uint8)
    g = orc.make_geom(pn, pp, scale=(0.01, 0.01, 0.01), distance=0.1)
    partial = 0
    for q in v[::50]:
        x0, x1, y0, y1 = orc.box(g, q["pos"] * 0.01, lam, diff)
        partial += int((x1 - x0) * (y1 - y0) < npx)
    assert partial > 0
    pix = np.random.default_rng(600 + 2 * diff + prec).choice(npx, 3000, replace=False)
    want = orc.field_exact(g, v, lam, 0, diff, pix)
    err = rel_l2(f[pix], want)
    print(f"dice diff={diff} prec={prec}: rel-L2 {err:.3e}, {partial} of {len(v[::50])} sampled points partial")
    assert err < (1e-5 if prec == 0 else 1e-9)
    assert np.allclose(e, orc.encode(f, orc.ENCODE_PHASE), rtol=0, atol=4e-16)
    assert np.array_equal(u, orc.normalize([e])[0])


@pytest.mark.parametrize("wid,diff", [(1, 0), (3, 1), (5, 0), (5, 1)])
def test_fp64_probe_vs_oracle_at_workload_geometry(ctx, orc, wid, diff):
    """The library's FP64 probe (ohx_pc_probe_f64, the reference behind the
    benchmark's rel-L2 figures; it shares k_pc_prep with the kernels) against
    the extended-precision oracle at the 10k-point 1080p, the 1M-point
    complex-amplitude and the 2M-point 4320p workload geometries (first
    3,000 points, last channel, 1,000 pixels; partial masks at 4320p):
    relative L2 <= 1e-12."""
    w = synth.WORKLOADS[wid]
    v = synth.points(w, 3000)
    ch = len(w.lambdas) - 1
    g = orc.make_geom(w.pn, w.pp)
    pix = np.random.default_rng(700 + wid).choice(w.n_pixels, 1000, replace=False)
    dv = ctx.to_device(v)
    try:
        got = ctx.probe_f64(dv, len(v), capi.params(w.pn, w.pp, w.lambdas[ch], channel=ch,
                                                     use_phase=w.complex_amplitude), diff, pix)
    finally:
        dv.free()
    want = orc.field_exact(g, v, w.lambdas[ch], ch, diff, pix, use_phase=w.complex_amplitude)
    err = rel_l2(got, want)
    print(f"workload {wid} diff={diff}: probe rel-L2 {err:.2e}")
    assert err < 1e-12
