"""CPU oracle of the single-side-band encoder ophPointCloud::encodeHologram,
in numpy FP64.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference check.  The product
package never imports this module.

Restates ophgen/src/ophPointCloud.cpp:193-240 per channel:
  * tmp = fftShift(FFT(fftShift(H))), then tmp = fftShift(IFFT(fftShift(tmp)))
    with Openholo::fft2 (Openholo/src/Openholo.cpp:698-751: shift, FFTW
    execute, shift; unnormalised) and Openholo::fftShift (:753-772:
    out[(i - nx/2) mod nx, (j - ny/2) mod ny] = in[i, j]);
  * enc[i] = Re(tmp[i] * (cos X cos Y - sin X sin Y, 0)) with, for pixel
    (w, h) = (i mod pnX, i / pnX):
      y = (ss_y - pp_y) - pp_y h,     Y = (pi y shift_y) / pp_y,
      x = -ss_x/2 + pp_x w + pp_x/2,  X = (pi x shift_x) / pp_x.
band_limit is a parameter of the reference function that its body never
reads.

Departure from the reference's code, a defect there: both fft2 calls run
fftShift with src == dst (the plan-bound input buffer is the source, set up
by fft2(ivec2, in, sign) at Openholo.cpp:559-585), and the shift's loop runs
in parallel over columns (OpenMP), so what the in-place shift leaves depends
on thread scheduling.  The oracle shifts out of place, as the function
intends; complex_H is left unchanged (the reference's in-place shift also
overwrites it).
"""
from __future__ import annotations

import math

import numpy as np


def fft_shift(a: np.ndarray) -> np.ndarray:
    """Openholo::fftShift, out of place: out[k] = in[(k + n/2) mod n] per axis."""
    ny, nx = a.shape
    return np.roll(a, (-(ny >> 1), -(nx >> 1)), axis=(0, 1))


def carrier(pn, pp, spectrum_shift) -> np.ndarray:
    """cos X cos Y - sin X sin Y per pixel, the reference's operation order
    (ophPointCloud.cpp:206-233)."""
    pn_x, pn_y = int(pn[0]), int(pn[1])
    pp_x, pp_y = float(pp[0]), float(pp[1])
    ss_x, ss_y = pn_x * pp_x, pn_y * pp_y
    half_pp_x, half_ss_x = pp_x / 2, ss_x / 2
    w = np.arange(pn_x, dtype=np.float64)[None, :]
    h = np.arange(pn_y, dtype=np.float64)[:, None]
    y = (ss_y - pp_y) - (pp_y * h)
    Y = (math.pi * y * spectrum_shift[1]) / pp_y
    x = -half_ss_x + pp_x * w + half_pp_x
    X = (math.pi * x * spectrum_shift[0]) / pp_x
    return 1.0 * (np.cos(X) * np.cos(Y) - np.sin(X) * np.sin(Y))


def encode_ssb(field: np.ndarray, pp, spectrum_shift=(0.0, 0.5)) -> np.ndarray:
    """enc (pnY, pnX) FP64 for one channel's complex field (pnY, pnX)."""
    h = np.asarray(field, dtype=np.complex128)
    ny, nx = h.shape
    tmp = fft_shift(np.fft.fft2(fft_shift(h)))
    tmp = fft_shift(np.fft.ifft2(fft_shift(tmp)) * (nx * ny))   # FFTW backward is unnormalised
    return tmp.real * carrier((nx, ny), pp, spectrum_shift)


def round_trip_naive(field: np.ndarray) -> np.ndarray:
    """The same FFT round trip as explicit DFT sums (O(N^2) per axis), for
    tiny grids: forward sign -1, backward sign +1, no normalisation."""
    h = np.asarray(field, dtype=np.complex128)
    ny, nx = h.shape

    def dft(a, sign):
        out = np.zeros_like(a)
        for v in range(ny):
            for u in range(nx):
                acc = 0j
                for j in range(ny):
                    for i in range(nx):
                        acc += a[j, i] * complex(math.cos(2 * math.pi * (u * i / nx + v * j / ny)),
                                                 sign * math.sin(2 * math.pi * (u * i / nx + v * j / ny)))
                out[v, u] = acc
        return out

    def shift(a):
        out = np.empty_like(a)
        for j in range(ny):
            for i in range(nx):
                out[(j - (ny >> 1)) % ny, (i - (nx >> 1)) % nx] = a[j, i]
        return out

    tmp = shift(dft(shift(h), -1))
    return shift(dft(shift(tmp), +1))
