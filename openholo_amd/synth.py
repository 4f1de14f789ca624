"""Deterministic synthetic point clouds for the point-cloud hologram workloads.

SplitMix64 with seed ``20261202 + k`` for workload k; a uniform double is
``(u >> 11) * 2**-53``.  Per point the draws are taken in the order x, y, z,
amplitude (and, for the complex-amplitude workload, phase).  Points are stored
as the 56-byte ``Vertex`` record of the reference
(Openholo/src/struct.h:86-106): pos[3], color[3] (all = amplitude), phase.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

VERTEX_DTYPE = np.dtype([("pos", "<f8", (3,)), ("color", "<f8", (3,)), ("phase", "<f8")])

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start .. start+count-1 of the SplitMix64 stream seeded with seed."""
    with np.errstate(over="ignore"):
        idx = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int, start: int = 0) -> np.ndarray:
    return (splitmix64(seed, count, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class Workload:
    """One hologram workload: point distribution and hologram-plane geometry."""
    name: str
    n_points: int
    xy_half: float            # x, y uniform in [-xy_half, +xy_half] m
    z_range: tuple            # z uniform in [z0, z1] m
    pn: tuple                 # (pn_x, pn_y)
    pp: tuple                 # (pp_x, pp_y) m
    lambdas: tuple            # per channel, m
    modes: tuple              # diffraction flags (0 = RS, 1 = Fresnel)
    seed: int
    complex_amplitude: bool = False
    amp_range: tuple = (0.1, 1.0)
    extra: dict = field(default_factory=dict)

    @property
    def n_pixels(self) -> int:
        return self.pn[0] * self.pn[1]

    def pairs(self, n_points: int | None = None) -> int:
        n = self.n_points if n_points is None else n_points
        return n * self.n_pixels * len(self.lambdas)


WORKLOADS = {
    1: Workload("rs-10k-1080p-532nm", 10_000, 4e-3, (0.49, 0.51), (1920, 1080), (8e-6, 8e-6),
                (532e-9,), (0,), 20261202 + 1),
    2: Workload("rs+fresnel-100k-2160p-rgb", 100_000, 8e-3, (0.45, 0.55), (3840, 2160),
                (3.74e-6, 3.74e-6), (638e-9, 532e-9, 473e-9), (0, 1), 20261202 + 2),
    3: Workload("fresnel-1m-2160p-532nm-complex", 1_000_000, 8e-3, (0.4, 0.6), (3840, 2160),
                (3.74e-6, 3.74e-6), (532e-9,), (1,), 20261202 + 3, complex_amplitude=True),
    5: Workload("rs+fresnel-2m-4320p-rgb", 2_000_000, 12e-3, (0.3, 0.7), (7680, 4320),
                (3.74e-6, 3.74e-6), (638e-9, 532e-9, 473e-9), (0, 1), 20261202 + 5),
}


def points(w: Workload, n: int | None = None) -> np.ndarray:
    """First n points (default all) of workload w as 56-byte Vertex records."""
    n = w.n_points if n is None else int(n)
    nd = 5 if w.complex_amplitude else 4
    u = uniform01(w.seed, n * nd).reshape(n, nd)
    v = np.zeros(n, VERTEX_DTYPE)
    v["pos"][:, 0] = -w.xy_half + 2 * w.xy_half * u[:, 0]
    v["pos"][:, 1] = -w.xy_half + 2 * w.xy_half * u[:, 1]
    v["pos"][:, 2] = w.z_range[0] + (w.z_range[1] - w.z_range[0]) * u[:, 2]
    a = w.amp_range[0] + (w.amp_range[1] - w.amp_range[0]) * u[:, 3]
    v["color"][:] = a[:, None]
    if w.complex_amplitude:
        v["phase"] = 2 * math.pi * u[:, 4]
    return v


def random_cloud(seed: int, n: int, xy_half: float, z_range, amp_range=(0.1, 1.0),
                 complex_amplitude: bool = False) -> np.ndarray:
    """Ad-hoc cloud with the same draw order (used by tests and the serving stream)."""
    w = Workload("adhoc", n, xy_half, tuple(z_range), (1, 1), (1.0, 1.0), (1.0,), (0,), seed,
                 complex_amplitude=complex_amplitude, amp_range=tuple(amp_range))
    return points(w)


def serving_requests(count: int, seed: int = 20261202 + 4):
    """The mixed request stream: point count log-uniform in [1e3, 5e5],
    resolution uniform over {1920x1080 at 8 um, 3840x2160 at 3.74 um},
    wavelength uniform over {638, 532, 473} nm, Fresnel with probability 0.7
    else RS; each request's points drawn as in the 10k-point RS workload
    (x, y in +-4 mm, z in [0.49, 0.51] m, amplitude in [0.1, 1]) from its own
    seed.  Draws per request, in order: count, resolution, wavelength, mode."""
    u = uniform01(seed, 4 * count).reshape(count, 4)
    reqs = []
    for i in range(count):
        n = int(round(10 ** (3 + u[i, 0] * (math.log10(5e5) - 3))))
        hi = u[i, 1] >= 0.5
        reqs.append({
            "n": n,
            "pn": (3840, 2160) if hi else (1920, 1080),
            "pp": (3.74e-6, 3.74e-6) if hi else (8e-6, 8e-6),
            "lambda": (638e-9, 532e-9, 473e-9)[min(2, int(u[i, 2] * 3))],
            "diff": 1 if u[i, 3] < 0.7 else 0,
            "seed": seed * 1000 + i,
        })
    return reqs
