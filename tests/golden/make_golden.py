"""Writes tests/golden/single_point.json: 50-digit field values of single
point sources, computed with mpmath from the textbook impulse responses

  Rayleigh-Sommerfeld (first kind, r >> lambda):
      h = a * z / (j * lambda * r^2) * exp(j k r)
  Fresnel:
      h = a * exp(j k z) / (j * lambda * z) * exp(j k (dx^2 + dy^2) / (2 z))

on the reference's pixel grid x_i = -ss_x/2 + (i - 1 + off_x) pp_x,
y_j = -ss_y/2 + (pn_y - j + off_y) pp_y (ophgen/src/ophGen.cpp:413,422,
ss = pn * pp).  These are complex-exponential forms, not the sin / -cos pair
the reference evaluates, so a sign, index or factor mistake in the oracle
shows up as a mismatch.  Inputs are doubles; mpmath takes them exactly.

Run:  python tests/golden/make_golden.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50

CASES = [
    # name, pn, pp, offset, point (x, y, z), amplitude, lambda, mode, pixels [(col,row)]
    dict(name="rs_tiny", pn=(64, 48), pp=(10e-6, 10e-6), off=(0, 0),
         p=(12.3e-6, -7.7e-6, 2.0e-3), a=0.75, lam=633e-9, mode=0,
         pix=[(32, 24), (33, 24), (30, 27), (35, 22), (34, 25)]),
    dict(name="fresnel_tiny", pn=(64, 48), pp=(10e-6, 10e-6), off=(0, 0),
         p=(12.3e-6, -7.7e-6, 2.0e-3), a=0.75, lam=633e-9, mode=1,
         pix=[(32, 24), (33, 24), (30, 27), (35, 22), (36, 21)]),
    dict(name="rs_offset", pn=(40, 30), pp=(8e-6, 6e-6), off=(3, -2),
         p=(-20.0e-6, 15.0e-6, 3.0e-3), a=0.4, lam=532e-9, mode=0,
         pix=[(14, 15), (16, 13), (18, 17), (15, 16)]),
    dict(name="fresnel_offset", pn=(40, 30), pp=(8e-6, 6e-6), off=(3, -2),
         p=(-20.0e-6, 15.0e-6, 3.0e-3), a=0.4, lam=532e-9, mode=1,
         pix=[(14, 15), (16, 13), (18, 17), (12, 19)]),
    dict(name="rs_far_1080p", pn=(1920, 1080), pp=(8e-6, 8e-6), off=(0, 0),
         p=(1.234e-3, -2.345e-3, 0.5), a=0.9, lam=532e-9, mode=0,
         pix=[(0, 0), (1919, 0), (0, 1079), (1919, 1079), (960, 540), (777, 333)]),
    dict(name="fresnel_far_2160p", pn=(3840, 2160), pp=(3.74e-6, 3.74e-6), off=(0, 0),
         p=(-7.5e-3, 6.25e-3, 0.45), a=0.3, lam=473e-9, mode=1,
         pix=[(0, 0), (3839, 0), (0, 2159), (3839, 2159), (1920, 1080), (123, 2001)]),
]


def field(c, col, row):
    pnx, pny = c["pn"]
    ppx, ppy = (mp.mpf(v) for v in c["pp"])
    x = -mp.mpf(pnx) * ppx / 2 + (col - 1 + c["off"][0]) * ppx
    y = -mp.mpf(pny) * ppy / 2 + (pny - row + c["off"][1]) * ppy
    px, py, pz = (mp.mpf(v) for v in c["p"])
    a, lam = mp.mpf(c["a"]), mp.mpf(c["lam"])
    k = 2 * mp.pi / lam
    dx, dy = x - px, y - py
    j = mp.mpc(0, 1)
    if c["mode"] == 0:
        r = mp.sqrt(dx * dx + dy * dy + pz * pz)
        h = a * pz / (j * lam * r * r) * mp.exp(j * k * r)
    else:
        h = a * mp.exp(j * k * pz) / (j * lam * pz) * mp.exp(j * k * (dx * dx + dy * dy) / (2 * pz))
    return h


def main():
    out = []
    for c in CASES:
        vals = []
        for col, row in c["pix"]:
            h = field(c, col, row)
            vals.append([col, row, mp.nstr(h.real, 30), mp.nstr(h.imag, 30)])
        d = dict(c)
        d["values"] = vals
        del d["pix"]
        out.append(d)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "single_point.json")
    with open(path, "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "dps": 50, "cases": out}, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
