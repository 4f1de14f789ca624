"""GPU parity of the single-side-band encoder (ohx_encode_ssb,
ophPointCloud::encodeHologram) against the numpy oracle
(oracle/ssb_oracle.py), which runs the reference's FFT round trip with numpy
FFTs; the kernel evaluates it in closed form.

Tolerance: encoded buffer rel-L2 1e-12 (FP64; the oracle's FFT rounding), the
8-bit output exact against the oracle's normaliser on the same encoded
buffers.
"""
import os
import subprocess

import numpy as np
import pytest

from openholo_amd import capi
from oracle import ssb_oracle as so

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "openholo_amd", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def ctx():
    c = capi.Context(0)
    yield c
    c.close()


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("pn,pp,shift", [
    ((1920, 1080), (8e-6, 8e-6), (0.0, 0.5)),       # reference defaults, even sizes
    ((255, 101), (6.4e-6, 8e-6), (0.3, 0.7)),       # odd x odd
    ((64, 33), (3.74e-6, 3.74e-6), (-0.25, 0.0)),   # even x odd
    ((45, 40), (1e-5, 1e-5), (0.5, 0.5)),           # odd x even
])
def test_ssb_matches_oracle(ctx, pn, pp, shift):
    rng = np.random.default_rng(pn[0] + pn[1])
    h = rng.normal(size=(pn[1], pn[0])) + 1j * rng.normal(size=(pn[1], pn[0]))
    npx = pn[0] * pn[1]
    df = ctx.to_device(h.astype(np.complex128).ravel())
    de = ctx.alloc(8 * npx)
    try:
        ctx.encode_ssb(df, pn, pp, de, shift)
        got = de.download(np.float64, npx).reshape(pn[1], pn[0])
        kept = df.download(np.complex128, npx).reshape(pn[1], pn[0])
    finally:
        df.free()
        de.free()
    want = so.encode_ssb(h, pp, shift)
    assert rel(got, want) < 1e-12
    assert np.array_equal(kept, h)                    # the field is left unchanged


def test_ssb_argument_errors(ctx):
    df = ctx.alloc(16 * 16)
    de = ctx.alloc(8 * 16)
    try:
        for pn, pp in [((0, 4), (8e-6, 8e-6)), ((4, 4), (0.0, 8e-6)), ((4, 4), (8e-6, float("nan")))]:
            with pytest.raises(capi.OhxError) as e:
                ctx.encode_ssb(df, pn, pp, de)
            assert e.value.status == -1          # OHX_ERR_INVALID_ARG
    finally:
        df.free()
        de.free()


def test_cli_ssb_encoding(tmp_path, orc):
    """ophPointCloud::encoding(ENCODE_SSB, passband) through the drop-in on the
    reference's RGB configuration and point cloud: encodeHologram on every
    channel's field, then normalize."""
    d = tmp_path
    cfg = os.path.join(GOLD, "Generation_PointCloud_RGB.xml")
    ply = os.path.join(GOLD, "pointcloud_1470.ply")
    r = subprocess.run([os.path.join(BIN, "ophpc_cli"), "-c", cfg, "-i", ply, "-m", "0", "-p", "1", "-f", "1",
                        "-e", "7", "--passband", "0", "-o", str(d / "R"), "--dump-field", str(d / "f.bin"),
                        "--dump-enc", str(d / "e.bin"), "--dump-u8", str(d / "u.bin")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    pn = (1920, 1080)
    f = np.fromfile(d / "f.bin", np.complex128).reshape(3, pn[1], pn[0])
    e = np.fromfile(d / "e.bin", np.float64).reshape(3, pn[1], pn[0])
    u = np.fromfile(d / "u.bin", np.uint8).reshape(3, -1)
    for ch in range(3):
        assert rel(e[ch], so.encode_ssb(f[ch], (10e-6, 10e-6))) < 1e-12, ch
    want_u = orc.normalize([e[ch].ravel() for ch in range(3)])
    for ch in range(3):
        assert np.array_equal(u[ch], want_u[ch])
    assert (d / "R" / "PointCloud_Fresnel_SSB.bmp").exists()
