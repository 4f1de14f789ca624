"""The C++ drop-in end to end on the GPU: the reference CLI's flags and the
reference example's call sequence, on the reference's own input files
(tests/golden), checked against the CPU oracle.

Tolerances: field rel-L2 vs the FP64 oracle 1e-5 (FP32 direct, -m 0) and
1e-9 (FP64 direct, -m 1) on sampled pixels; encoded buffers vs the oracle's
encoder applied to the same field 1e-15; 8-bit output vs the oracle's
normaliser applied to the same encoded buffers: exact.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openholo_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")
VERTEX = np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])
CFG_RGB = os.path.join(GOLD, "Generation_PointCloud_RGB.xml")
PLY = os.path.join(GOLD, "pointcloud_1470.ply")


def run(*args):
    r = subprocess.run(list(args), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, f"{args[0]} failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}"
    return r


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("diff", [0, 1])
@pytest.mark.parametrize("prec", [0, 1])
def test_cli_rgb_end_to_end(tmp_path, orc, diff, prec):
    d = tmp_path
    run(os.path.join(BIN, "ophpc_cli"), "-a", "0", "-c", CFG_RGB, "-i", PLY, "-e", "0", "-m", str(prec),
        "-p", "1", "-f", str(diff), "-o", str(d / "Result"), "--dump-points", str(d / "p.bin"),
        "--dump-field", str(d / "f.bin"), "--dump-enc", str(d / "e.bin"), "--dump-u8", str(d / "u.bin"))
    pn, pp, lams = (1920, 1080), (10e-6, 10e-6), [638e-9, 520e-9, 450e-9]
    npx = pn[0] * pn[1]
    v = np.fromfile(d / "p.bin", VERTEX)
    f = np.fromfile(d / "f.bin", np.complex128).reshape(3, npx)
    e = np.fromfile(d / "e.bin", np.float64).reshape(3, npx)
    u = np.fromfile(d / "u.bin", np.uint8).reshape(3, npx)
    g = orc.make_geom(pn, pp, scale=(0.01, 0.01, 0.01), distance=0.5)
    pix = np.random.default_rng(8).choice(npx, 1500, replace=False)
    tol = 1e-5 if prec == 0 else 1e-9
    for ch, lam in enumerate(lams):
        want = orc.field_exact(g, v, lam, ch, diff, pix)
        assert rel_l2(f[ch][pix], want) < tol, ch
        assert np.allclose(e[ch], orc.encode(f[ch], orc.ENCODE_PHASE), rtol=0, atol=4e-16)
    want_u = orc.normalize(list(e))
    for ch in range(3):
        assert np.array_equal(u[ch], want_u[ch])
    # merged 24-bit BMP, rotated 180 degrees (IMG_Merge=1, IMG_Rotation=1)
    bmp = open(d / "Result" / f"PointCloud_{'R-S' if diff == 0 else 'Fresnel'}_Phase.bmp", "rb").read()
    assert bmp[:2] == b"BM"
    size, _, off = struct.unpack("<IIi", bmp[2:14])
    w, h, planes, bpp = struct.unpack("<iiHH", bmp[18:30])
    assert (size, w, h, planes, bpp) == (len(bmp), 1920, 1080, 1, 24)
    pixels = np.frombuffer(bmp[off:], np.uint8).reshape(1080, 1920, 3)
    bgr = np.stack([u[2], u[1], u[0]], axis=-1).reshape(1080, 1920, 3)
    assert np.array_equal(pixels, bgr[::-1, ::-1])


def test_example_program_ohc_and_bmp(tmp_path):
    d = tmp_path
    cfg = os.path.join(GOLD, "Dice_PointCloud.xml")
    run(os.path.join(BIN, "examPointCloud"), cfg, PLY, str(d / "out"), "rs", "1")
    run(os.path.join(BIN, "ophpc_cli"), "-c", cfg, "-i", PLY, "-p", "1", "-f", "0", "-o", str(d / "R"),
        "--dump-field", str(d / "f.bin"), "--dump-u8", str(d / "u.bin"))
    npx = 1920 * 1080
    field = np.fromfile(d / "f.bin", np.complex128)
    ohc = open(d / "out.ohc", "rb").read()
    assert ohc[:2] == b"OH"
    fsize, vmaj, vmin, _, _, data_off = struct.unpack("<QBBIII", ohc[2:24])
    hsize, nx, ny, ppx, ppy, punit, nw = struct.unpack("<IIIddBI", ohc[24:57])
    ctype, carr, wunit, dtype, store, code, pcode = struct.unpack("<7B", ohc[57:64])
    assert (fsize, vmaj, vmin) == (len(ohc), 3, 0)
    assert (nx, ny, ppx, ppy, punit, nw) == (1920, 1080, 4.3e-6, 4.3e-6, 1, 1)
    assert (ctype, carr, wunit, dtype, store, code, pcode) == (2, 2, 1, 10, 1, 2, 0)
    assert hsize == 65 + 8 and data_off == 24 + hsize
    assert struct.unpack("<d", ohc[89:97])[0] == 532e-9
    planes = np.frombuffer(ohc[data_off:], np.float64).reshape(2, npx)
    assert np.array_equal(planes[0], field.real) and np.array_equal(planes[1], field.imag)
    # and back through Openholo::loadAsOhc
    run(os.path.join(BIN, "ophpc_cli"), "--load-ohc", str(d / "out.ohc"), "--dump-field", str(d / "back.bin"))
    assert np.array_equal(np.fromfile(d / "back.bin", np.complex128), field)
    # 8-bit palette BMP of the single channel
    bmp = open(d / "out.bmp", "rb").read()
    off = struct.unpack("<i", bmp[10:14])[0]
    assert struct.unpack("<H", bmp[28:30])[0] == 8 and off == 14 + 40 + 1024
    assert np.array_equal(np.frombuffer(bmp[off:], np.uint8), np.fromfile(d / "u.bin", np.uint8))


@pytest.mark.parametrize("diff", [0, 1])
def test_cli_mfma(tmp_path, orc, diff):
    """The drop-in with --mfma bf16x3 on the reference's RGB example, Fresnel
    (-f 1) and RS (-f 0): every channel's field vs the oracle within the
    BF16 3-term bound."""
    d = tmp_path
    run(os.path.join(BIN, "ophpc_cli"), "-c", CFG_RGB, "-i", PLY, "-m", "0", "-p", "1", "-f", str(diff),
        "--mfma", "bf16x3", "-o", str(d / "R"), "--dump-points", str(d / "p.bin"), "--dump-field", str(d / "f.bin"))
    v = np.fromfile(d / "p.bin", VERTEX)
    f = np.fromfile(d / "f.bin", np.complex128).reshape(3, -1)
    g = orc.make_geom((1920, 1080), (10e-6, 10e-6), scale=(0.01, 0.01, 0.01), distance=0.5)
    pix = np.random.default_rng(9).choice(1920 * 1080, 1500, replace=False)
    for ch, lam in enumerate([638e-9, 520e-9, 450e-9]):
        err = rel_l2(f[ch][pix], orc.field_exact(g, v, lam, ch, diff, pix))
        print(f"diff={diff} ch={ch}: rel-L2 {err:.3e}")
        assert err < 1e-4


def test_cli_wrp_end_to_end(tmp_path, orc):
    """ex_WRP's call sequence (-a 4) on the reference's WRP configuration and
    point cloud: each channel's hologram and the last channel's WRP against
    the numpy oracle (oracle/wrp_oracle.py), then the encoder, normaliser and
    merged 24-bit BMP."""
    from oracle import wrp_oracle as wo
    d = tmp_path
    cfg = os.path.join(GOLD, "Generation_WRP_RGB.xml")
    run(os.path.join(BIN, "ophpc_cli"), "-a", "4", "-c", cfg, "-i", PLY, "-e", "0", "-m", "1", "-p", "1",
        "-o", str(d / "Result"), "--dump-points", str(d / "p.bin"), "--dump-field", str(d / "f.bin"),
        "--dump-enc", str(d / "e.bin"), "--dump-u8", str(d / "u.bin"), "--dump-wrp", str(d / "w.bin"))
    pn, pp, lams = (1920, 1080), (1e-5, 1e-5), [638e-9, 520e-9, 450e-9]
    scale, loc, dist = (0.0015, 0.0015, 0.0003), 0.003, 0.5
    npx = pn[0] * pn[1]
    v = np.fromfile(d / "p.bin", VERTEX)
    f = np.fromfile(d / "f.bin", np.complex128).reshape(3, pn[1], pn[0])
    e = np.fromfile(d / "e.bin", np.float64).reshape(3, npx)
    u = np.fromfile(d / "u.bin", np.uint8).reshape(3, npx)
    wrp = np.fromfile(d / "w.bin", np.complex128).reshape(pn[1], pn[0])
    for ch, lam in enumerate(lams):
        want_wrp, want = wo.generate(v, pn, pp, scale, lam, ch, loc, dist)
        assert rel_l2(f[ch], want) < 1e-10, ch
        if ch == 2:
            assert rel_l2(wrp, want_wrp) < 1e-12
        assert np.allclose(e[ch], orc.encode(f[ch].ravel(), orc.ENCODE_PHASE), rtol=0, atol=4e-16)
    want_u = orc.normalize(list(e))
    for ch in range(3):
        assert np.array_equal(u[ch], want_u[ch])
    bmp = open(d / "Result" / "WRP_Phase.bmp", "rb").read()
    size, _, off = struct.unpack("<IIi", bmp[2:14])
    w, h, planes, bpp = struct.unpack("<iiHH", bmp[18:30])
    assert (size, w, h, planes, bpp) == (len(bmp), 1920, 1080, 1, 24)
    pixels = np.frombuffer(bmp[off:], np.uint8).reshape(1080, 1920, 3)
    assert np.array_equal(pixels, np.stack([u[2], u[1], u[0]], axis=-1).reshape(1080, 1920, 3))


def _bmp_rows(path, w, h, ch):
    bmp = open(path, "rb").read()
    assert bmp[:2] == b"BM"
    off = struct.unpack("<i", bmp[10:14])[0]
    bw, bh, _, bpp = struct.unpack("<iiHH", bmp[18:30])
    assert (bw, bh, bpp) == (w, h, 8 * ch)
    stride = (w * ch + 3) & ~3
    raw = np.frombuffer(bmp[off:], np.uint8)
    assert raw.size == stride * h
    rows = raw.reshape(h, stride)
    assert not rows[:, w * ch:].any()                 # row padding is zero
    return rows[:, :w * ch].reshape(h, w, ch)


@pytest.mark.parametrize("merge,flip,rotate", [(0, 1, 0), (1, 2, 1)])
def test_cli_odd_width_bmp(tmp_path, merge, flip, rotate):
    """1001 x 601 hologram (rows padded to 4 bytes): per-channel 8-bit files
    (IMG_Merge 0, '<i>_' prefixes) or one merged 24-bit file, flipped
    (IMG_Flip 1 vertical, 2 horizontal) and rotated by 180 degrees."""
    d = tmp_path
    text = open(CFG_RGB).read()
    for tag, val in [("SLM_PixelNumX", "1001"), ("SLM_PixelNumY", "601"), ("IMG_Merge", str(merge)),
                     ("IMG_Flip", str(flip)), ("IMG_Rotation", str(rotate))]:
        a = text.index("<%s>" % tag) + len(tag) + 2
        text = text[:a] + val + text[text.index("</%s>" % tag):]
    cfg = d / "odd.xml"
    cfg.write_text(text)
    run(os.path.join(BIN, "ophpc_cli"), "-c", str(cfg), "-i", PLY, "-m", "0", "-p", "1", "-f", "1",
        "-o", str(d / "R"), "--dump-u8", str(d / "u.bin"))
    w, h = 1001, 601
    u = np.fromfile(d / "u.bin", np.uint8).reshape(3, h, w)
    mode = flip ^ (3 if rotate else 0)
    def tf(img):
        if mode & 1:
            img = img[::-1]
        if mode & 2:
            img = img[:, ::-1]
        return img
    if merge:
        got = _bmp_rows(d / "R" / "PointCloud_Fresnel_Phase.bmp", w, h, 3)
        assert np.array_equal(got, tf(np.stack([u[2], u[1], u[0]], axis=-1)))
    else:
        for ch in range(3):
            got = _bmp_rows(d / "R" / f"{ch}_PointCloud_Fresnel_Phase.bmp", w, h, 1)
            assert np.array_equal(got[..., 0], tf(u[ch]))


@pytest.mark.parametrize("diff", [0, 1])
def test_cli_resize_between_holograms(tmp_path, diff):
    """setResolution between two generateHologram calls in one process: the
    drop-in frees and re-allocates its page-locked complex_H planes and its
    device buffers.  Each field equals a one-shot generation at its own size
    through the C ABI (same kernels, so to 1e-13), and the 8-bit output has
    the second size."""
    from openholo_amd import capi
    d = tmp_path
    run(os.path.join(BIN, "ophpc_cli"), "-a", "0", "-c", CFG_RGB, "-i", PLY, "-e", "0", "-m", "0", "-p", "1",
        "-f", str(diff), "-o", str(d / "Result"), "--dump-points", str(d / "p.bin"),
        "--dump-field", str(d / "f1.bin"), "--resize", "700,500", "--dump-field2", str(d / "f2.bin"),
        "--dump-u8", str(d / "u.bin"))
    v = np.fromfile(d / "p.bin", VERTEX)
    pp, lams = (10e-6, 10e-6), [638e-9, 520e-9, 450e-9]
    ctx = capi.Context(0)
    try:
        for name, pn in (("f1.bin", (1920, 1080)), ("f2.bin", (700, 500))):
            f = np.fromfile(d / name, np.complex128).reshape(3, pn[1], pn[0])
            want = capi.generate(ctx, v, pn, pp, lams, diff, capi.MODE_GPU | capi.MODE_FLOAT,
                                 scale=(0.01, 0.01, 0.01), distance=0.5)
            for ch in range(3):
                assert np.abs(f[ch]).max() > 0, (name, ch)
                assert rel_l2(f[ch], want[ch]) < 1e-13, (name, ch)
    finally:
        ctx.close()
    assert os.path.getsize(d / "u.bin") == 3 * 700 * 500


def test_cli_edit_of_complex_field_reaches_encoding(tmp_path, orc):
    """A caller that edits complex_H through getComplexField() between
    generateHologram and encoding gets the edited field encoded, as in the
    reference, which encodes from the host field (ophGen.cpp:816-849): the
    CLI conjugates every pixel, and the phase encoding equals the oracle's
    encoder applied to the conjugated field."""
    d = tmp_path
    run(os.path.join(BIN, "ophpc_cli"), "-a", "0", "-c", CFG_RGB, "-i", PLY, "-e", "0", "-m", "0", "-p", "1",
        "-f", "1", "-o", str(d / "Result"), "--dump-field", str(d / "f.bin"), "--conj-field",
        "--dump-enc", str(d / "e.bin"), "--dump-u8", str(d / "u.bin"))
    npx = 1920 * 1080
    f = np.fromfile(d / "f.bin", np.complex128).reshape(3, npx)
    e = np.fromfile(d / "e.bin", np.float64).reshape(3, npx)
    u = np.fromfile(d / "u.bin", np.uint8).reshape(3, npx)
    want_e = [orc.encode(np.conj(f[ch]), orc.ENCODE_PHASE) for ch in range(3)]
    for ch in range(3):
        assert np.allclose(e[ch], want_e[ch], rtol=0, atol=4e-16), ch
        # and not the unedited field's encoding
        assert not np.allclose(e[ch], orc.encode(f[ch], orc.ENCODE_PHASE), rtol=0, atol=1e-6)
    want_u = orc.normalize(want_e)
    for ch in range(3):
        assert np.array_equal(u[ch], want_u[ch])
