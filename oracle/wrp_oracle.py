"""CPU oracle of the WRP (wavefront recording plane) generator, in numpy FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference check.  The product
package never imports this module.

Restates the reference's CPU path:
  * ophWRP::autoScaling (ophgen/src/ophWRP.cpp:69-118, the `#if 1` branch):
    positions times scale, zmax = max(MIN_DOUBLE, max z);
  * ophWRP::calculateWRPCPU (ophgen/src/ophWRP.cpp:321-435): one WRP per
    channel; every point adds a window of spherical-wave terms at depth
    dz = wrp_location - zmax (the same dz for all points, as in the
    reference's CPU and GPU paths, ophWRPKernel.cu:191);
  * ophGen::Fresnel_FFT (ophgen/src/ophGen.cpp:516-581): zero padding to
    2 pnX x 2 pnY, centred forward FFT, angular-spectrum transfer function
    exp(j 2 pi d sqrt(1/lambda^2 - fx^2 - fy^2)), centred unnormalised
    inverse FFT, centre crop.

Two deliberate departures from the reference's code, both its defects:
  * Openholo::fft2 (Openholo/src/Openholo.cpp:698-751) applies fftShift
    with src == dst in Fresnel_FFT's calls (the plan-bound input buffer is
    the source), which overwrites half of each quadrant swap; the oracle
    shifts out of place, as the function intends.
  * the random-phase option draws from an unseeded RNG (CPU) or curand
    (GPU); results with it are not reproducible, so only bRandomPhase = false
    is restated (the phase factor is then cos(2 pi) = 1 and
    sin(2 pi) = -2.449e-16, kept exactly as the reference computes them).
"""
from __future__ import annotations

import math

import numpy as np

MIN_DOUBLE = 2.2250738585072014e-308
TWO_PI = 2.0 * math.pi


def scaled_positions(vertices: np.ndarray, scale):
    """ophWRP::autoScaling: scaled x, y, z and zmax (ophWRP.cpp:84-93)."""
    pos = vertices["pos"] * np.asarray(scale, dtype=np.float64)
    zmax = MIN_DOUBLE
    for z in pos[:, 2]:
        if zmax < z:
            zmax = z
    return pos, zmax


def window_half_width(lam: float, dz: float, pp_x: float) -> int:
    """w of ophWRP.cpp:366-369: (int)(|lambda dz / (2 pp_x^2)| * 2)."""
    ppxx = pp_x * pp_x * 2
    return int(abs(lam * dz / ppxx) * 2)


def c_trunc_div(a: float, b: float) -> int:
    """C's (int)(a / b): truncation toward zero."""
    return int(a / b)


def wrp_plane(vertices, pn, pp, scale, lam: float, channel: int, wrp_location: float) -> np.ndarray:
    """The WRP of one channel (ophWRP.cpp:344-420), complex128 (pnY, pnX),
    row index = tmpY, column index = tmpX as the reference addresses it."""
    pn_x, pn_y = int(pn[0]), int(pn[1])
    pp_x, pp_y = float(pp[0]), float(pp[1])
    pos, zmax = scaled_positions(vertices, scale)
    k = TWO_PI / lam
    dz = wrp_location - zmax
    dzz = dz * dz
    sign = dz > 0.0
    w = window_half_width(lam, dz, pp_x)
    hx, hy = pn_x >> 1, pn_y >> 1
    cr, ci = math.cos(TWO_PI * 1.0), math.sin(TWO_PI * 1.0)
    out = np.zeros((pn_y, pn_x), np.complex128)
    if w <= 0:
        return out
    wr = np.arange(-w, w)
    dx = wr * pp_x
    dy = wr * pp_y
    # the window's terms depend on (wx, wy) only (dz is common)
    rr = np.sqrt(dx[None, :] * dx[None, :] + (dy * dy)[:, None] + dzz)
    r = rr if sign else -rr
    kr = k * r
    for i in range(len(vertices)):
        x, y = pos[i, 0], pos[i, 1]
        a = float(vertices["color"][i, channel])
        tx = c_trunc_div(x, pp_x) + hx
        ty = c_trunc_div(y, pp_y) + hy
        re = (a * np.cos(kr) * cr) / r
        im = (-a * np.sin(kr) * ci) / r
        x0, y0 = tx - w, ty - w
        xs, ys = max(x0, 0), max(y0, 0)
        xe, ye = min(tx + w, pn_x), min(ty + w, pn_y)
        if xs >= xe or ys >= ye:
            continue
        out[ys:ye, xs:xe] += (re + 1j * im)[ys - y0:ye - y0, xs - x0:xe - x0]
    return out


def _shift(a: np.ndarray) -> np.ndarray:
    """Openholo::fftShift (Openholo.cpp:755-773), out of place."""
    ny, nx = a.shape
    return np.roll(a, (-(ny >> 1), -(nx >> 1)), axis=(0, 1))


def fresnel_fft(src: np.ndarray, pp, lam: float, distance: float) -> np.ndarray:
    """ophGen::Fresnel_FFT (ophGen.cpp:516-581) with out-of-place shifts."""
    pn_y, pn_x = src.shape
    pp_x, pp_y = float(pp[0]), float(pp[1])
    ss_x, ss_y = pn_x * pp_x, pn_y * pp_y
    dx2, dy2 = 2 * pn_x, 2 * pn_y
    hx, hy = pn_x >> 1, pn_y >> 1
    temp = np.zeros((dy2, dx2), np.complex128)
    temp[hy:hy + pn_y, hx:hx + pn_x] = src
    temp = _shift(np.fft.fft2(_shift(temp)))            # FFTW_FORWARD, unnormalised
    fx = (np.arange(dx2) - pn_x) / (ss_x * 2)
    fy = (np.arange(dy2) - pn_y) / (ss_y * 2)
    lut_x, lut_y = fx * fx, fy * fy
    sqrt_part = np.sqrt(1.0 / (lam * lam) - lut_x[None, :] - lut_y[:, None])
    tmp = 2 * math.pi * distance
    temp = temp * np.exp(1j * (tmp * sqrt_part))
    temp = _shift(np.fft.ifft2(_shift(temp)) * (dx2 * dy2))   # FFTW_BACKWARD, unnormalised
    return temp[hy:hy + pn_y, hx:hx + pn_x].copy()


def generate(vertices, pn, pp, scale, lam: float, channel: int, wrp_location: float, distance: float):
    """One channel of ophWRP::generateHologram: (WRP, hologram field)."""
    wrp = wrp_plane(vertices, pn, pp, scale, lam, channel, wrp_location)
    return wrp, fresnel_fft(wrp, pp, lam, distance)


def fresnel_dft_naive(src: np.ndarray, pp, lam: float, distance: float) -> np.ndarray:
    """The same propagation written as explicit DFT sums over the padded grid
    with centred indices (no FFT, no shifts): an independent check of the
    shift and frequency-index conventions of fresnel_fft, for tiny sizes."""
    pn_y, pn_x = src.shape
    pp_x, pp_y = float(pp[0]), float(pp[1])
    dx2, dy2 = 2 * pn_x, 2 * pn_y
    hx, hy = pn_x >> 1, pn_y >> 1
    # centred coordinates: padded sample (Y, X) sits at (Y - pn_y, X - pn_x)
    # after the centring shift, frequency index (i, j) at (i - pn_y, j - pn_x)
    ys = np.arange(dy2) - pn_y
    xs = np.arange(dx2) - pn_x
    pad = np.zeros((dy2, dx2), np.complex128)
    pad[hy:hy + pn_y, hx:hx + pn_x] = src
    ex = np.exp(-2j * np.pi * np.outer(xs, xs) / dx2)        # [freq, x]
    ey = np.exp(-2j * np.pi * np.outer(ys, ys) / dy2)
    spec = ey @ pad @ ex.T
    fx = xs / (pn_x * pp_x * 2)
    fy = ys / (pn_y * pp_y * 2)
    h = np.exp(1j * 2 * np.pi * distance * np.sqrt(1.0 / lam ** 2 - fx[None, :] ** 2 - fy[:, None] ** 2))
    spec = spec * h
    back = np.conj(ey).T @ spec @ np.conj(ex)
    return back[hy:hy + pn_y, hx:hx + pn_x]
