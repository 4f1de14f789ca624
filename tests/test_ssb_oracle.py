"""Pins of the SSB oracle (oracle/ssb_oracle.py) on the CPU.

The reference holds no SSB outputs, and its own output is not reproducible
(in-place shifts under OpenMP), so the oracle is pinned by:
  * the FFT round trip against explicit DFT sums with the shift written as
    the reference's index loop (no numpy FFT, no np.roll), odd and even sizes;
  * for even sizes the round trip is the identity times pnX pnY (the two
    shifts in the middle cancel), so with spectrum_shift = (0, 0) the
    encoding is pnX pnY Re(H);
  * the default carrier (0, 0.5) is cos(pi (pnY - 1 - h) / 2): rows cycle
    1, 0, -1, 0 from the last row upwards.
"""
import numpy as np
import pytest

from oracle import ssb_oracle as so


def rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


@pytest.mark.parametrize("shape", [(4, 6), (5, 3), (3, 4), (5, 5)])
def test_round_trip_matches_dft_sums(shape):
    h = rnd(shape, 1)
    ny, nx = shape
    got = so.fft_shift(np.fft.ifft2(so.fft_shift(so.fft_shift(np.fft.fft2(so.fft_shift(h))))) * (nx * ny))
    want = so.round_trip_naive(h)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # and the encoder uses exactly that round trip
    enc = so.encode_ssb(h, (8e-6, 8e-6), (0.3, 0.7))
    assert np.allclose(enc, want.real * so.carrier((nx, ny), (8e-6, 8e-6), (0.3, 0.7)), rtol=0, atol=1e-11)


def test_even_sizes_are_identity_round_trip():
    h = rnd((12, 16), 2)
    enc = so.encode_ssb(h, (8e-6, 6e-6), (0.0, 0.0))
    assert np.abs(enc - 192 * h.real).max() < 1e-12 * 192 * np.abs(h).max()


def test_odd_size_is_not_identity():
    h = rnd((5, 4), 3)
    enc = so.encode_ssb(h, (8e-6, 8e-6), (0.0, 0.0))
    assert np.abs(enc - 20 * h.real).max() > 1e-3      # the shifts do not cancel for odd n


def test_default_carrier_rows():
    c = so.carrier((6, 8), (8e-6, 8e-6), (0.0, 0.5))
    want = np.array([1.0, 0.0, -1.0, 0.0])
    for h in range(8):
        k = (8 - 1 - h) % 4
        assert np.allclose(c[h], want[k], atol=1e-15)
