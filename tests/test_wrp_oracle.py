"""Pins of the WRP oracle (oracle/wrp_oracle.py) on the CPU.

The reference holds no WRP golden outputs (and its own outputs are not
usable: in-place fftShift on the CPU path, atomicAdd misuse on the GPU
path), so the oracle is pinned by independent restatements and physics:
  * the propagation against explicit DFT sums with centred indices (no FFT,
    no shifts), on tiny grids;
  * zero distance: the unnormalised round trip multiplies by 4 pn_x pn_y;
  * propagation by +d then -d returns a compact field (sign of the transfer
    function's exponent, frequency grid);
  * the WRP accumulation against a pixel-major brute-force loop (window
    bounds, truncating pixel index, sign of r for dz < 0, plane clipping).
"""
import math

import numpy as np
import pytest

from oracle import wrp_oracle as wo


def wo_vertex_dtype():
    return np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape", [(4, 6), (5, 3), (6, 6)])
def test_propagation_matches_explicit_dft(shape):
    rng = np.random.default_rng(1)
    src = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    pp, lam, d = (10e-6, 8e-6), 633e-9, 0.01
    a = wo.fresnel_fft(src, pp, lam, d)
    b = wo.fresnel_dft_naive(src, pp, lam, d)
    assert rel(a, b) < 1e-12


def test_zero_distance_scales_by_padded_size():
    rng = np.random.default_rng(2)
    src = rng.normal(size=(12, 16)) + 1j * rng.normal(size=(12, 16))
    out = wo.fresnel_fft(src, (8e-6, 8e-6), 532e-9, 0.0)
    assert rel(out, src * (4 * 12 * 16)) < 1e-13


def test_forward_then_backward_propagation_is_identity():
    ny, nx = 48, 64
    y, x = np.mgrid[0:ny, 0:nx]
    src = np.exp(-(((x - nx / 2) / 4.0) ** 2 + ((y - ny / 2) / 4.0) ** 2)) * np.exp(1j * 0.3 * x)
    pp, lam, d = (10e-6, 10e-6), 633e-9, 1e-3
    n4 = 4 * nx * ny
    fwd = wo.fresnel_fft(src, pp, lam, d) / n4
    assert rel(fwd, src) > 1e-2                    # it did propagate
    back = wo.fresnel_fft(fwd, pp, lam, -d) / n4
    assert rel(back, src) < 1e-9
    # the other sign does not come back
    again = wo.fresnel_fft(fwd, pp, lam, d) / n4
    assert rel(again, src) > 1e-2


def _wrp_bruteforce(vertices, pn, pp, scale, lam, ch, wrp_location):
    pos = vertices["pos"] * np.asarray(scale)
    zmax = wo.MIN_DOUBLE
    for z in pos[:, 2]:
        zmax = max(zmax, z)
    k = 2 * math.pi / lam
    dz = wrp_location - zmax
    w = int(abs(lam * dz / (pp[0] * pp[0] * 2)) * 2)
    out = np.zeros((pn[1], pn[0]), np.complex128)
    for row in range(pn[1]):
        for col in range(pn[0]):
            acc = 0j
            for i in range(len(vertices)):
                tx = int(pos[i, 0] / pp[0]) + (pn[0] >> 1)
                ty = int(pos[i, 1] / pp[1]) + (pn[1] >> 1)
                wx, wy = col - tx, row - ty
                if -w <= wx < w and -w <= wy < w:
                    dx, dy = wx * pp[0], wy * pp[1]
                    r = math.sqrt(dx * dx + dy * dy + dz * dz)
                    r = r if dz > 0 else -r
                    a = vertices["color"][i, ch]
                    acc += complex(a * math.cos(k * r) * math.cos(2 * math.pi) / r,
                                   -a * math.sin(k * r) * math.sin(2 * math.pi) / r)
            out[row, col] = acc
    return out, w


@pytest.mark.parametrize("dz_sign", [1.0, -1.0])
def test_wrp_plane_matches_pixel_loop(dz_sign):
    rng = np.random.default_rng(5)
    pn, pp, lam = (20, 14), (10e-6, 12e-6), 633e-9
    n = 9
    v = np.zeros(n, wo_vertex_dtype())
    v["pos"][:, 0] = rng.uniform(-110e-6, 110e-6, n)      # some windows cross the edges
    v["pos"][:, 1] = rng.uniform(-90e-6, 90e-6, n)
    v["pos"][:, 2] = rng.uniform(0.0, 1e-3, n)
    v["pos"][0, :2] = (-14e-6, -25e-6)                     # negative: C truncation toward zero
    v["color"][:, 1] = rng.uniform(0.1, 1.0, n)
    scale = (1.0, 1.0, 1.0)
    zmax = v["pos"][:, 2].max()
    wrp_location = zmax + dz_sign * 6.5e-4
    want, w = _wrp_bruteforce(v, pn, pp, scale, lam, 1, wrp_location)
    assert w == 4
    got = wo.wrp_plane(v, pn, pp, scale, lam, 1, wrp_location)
    assert np.count_nonzero(want) > 50
    assert np.array_equal(got != 0, want != 0)
    assert rel(got, want) < 1e-14


def test_window_half_width():
    # (int)(|lambda dz / (2 pp^2)| * 2)
    assert wo.window_half_width(633e-9, 6.5e-4, 10e-6) == int(633e-9 * 6.5e-4 / (2e-10) * 2)
    assert wo.window_half_width(532e-9, -1e-3, 8e-6) == 8
    assert wo.window_half_width(532e-9, 1e-9, 8e-6) == 0

