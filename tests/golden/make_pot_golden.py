"""Generate tests/golden/pot_golden.json: high-precision values for the scalar
potential's Green function.

Two independent families of values, both at 40 significant digits (mpmath):

1. The potential antiderivative of src/green_functions.jl:13-22,
      P(x, y, z) = -z^2 atan(xy/(zr))/2 - y^2 atan(xz/(yr))/2 - x^2 atan(yz/(xr))/2
                   + yz log(x + r) + xz log(y + r) + xy log(z + r),
   at sample points, together with the check that its mixed third
   derivative is 1/r there (the property that makes it the IGF kernel).
2. Cell integrals  J(a, b, k) = integral over the cell
      [(a-1/2)dx, (a+1/2)dx] x [(b-1/2)dy, (b+1/2)dy] x [(k-1/2)dz, (k+1/2)dz]
   of 1/r.  The z integral is done in closed form (asinh(z/rho), rho the
   distance from the z axis) and the remaining 2-D integral by adaptive
   quadrature, split at x = 0 and y = 0 so that the axis is only ever an
   endpoint; these values do not use P at all.

Run:  python tests/golden/make_pot_golden.py   (takes about a minute)
"""
from __future__ import annotations

import json
import os

import mpmath as mp

mp.mp.dps = 40

HERE = os.path.dirname(os.path.abspath(__file__))


def P(x, y, z):
    r = mp.sqrt(x * x + y * y + z * z)
    return (-z * z * mp.atan(x * y / (z * r)) / 2 - y * y * mp.atan(x * z / (y * r)) / 2
            - x * x * mp.atan(y * z / (x * r)) / 2
            + y * z * mp.log(x + r) + x * z * mp.log(y + r) + x * y * mp.log(z + r))


def _split(lo, hi):
    return [lo, mp.mpf(0), hi] if lo < 0 < hi else [lo, hi]


def cell_integral(center, h):
    """Integral of 1/r over the cell with the given center and sides."""
    lo = [c - s / 2 for c, s in zip(center, h)]
    hi = [c + s / 2 for c, s in zip(center, h)]

    def inner(u, v):
        rho = mp.sqrt(u * u + v * v)
        return mp.asinh(hi[2] / rho) - mp.asinh(lo[2] / rho)

    return mp.quad(inner, _split(lo[0], hi[0]), _split(lo[1], hi[1]))


def main():
    points = [
        (0.3, 0.2, 0.1), (-0.5, 0.7, 1.1), (2.5, -1.5, 0.5), (0.05, 3.0, -2.0),
        (10.5, 7.5, -3.5), (-31.5, 0.5, 15.5), (1e-3, 2e-3, -1.5e-3), (4.0, 0.25, 6.5),
    ]
    pvals = []
    for p in points:
        x, y, z = (mp.mpf(t) for t in p)
        val = P(x, y, z)
        d3 = mp.diff(lambda a, b, c: P(a, b, c), (x, y, z), (1, 1, 1))
        r = mp.sqrt(x * x + y * y + z * z)
        assert abs(d3 - 1 / r) <= mp.mpf("1e-25") / r, p
        pvals.append({"point": list(p), "P": mp.nstr(val, 36)})

    spacings = [(1.0, 1.0, 1.0), (0.7, 1.3, 2.1)]
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (3, -2, 5), (-7, 4, -1),
             (12, 9, 6), (-20, 15, 10)]
    cvals = []
    for h in spacings:
        hm = [mp.mpf(s) for s in h]
        for cell in cells:
            center = [mp.mpf(a) * s for a, s in zip(cell, hm)]
            cvals.append({"h": list(h), "cell": list(cell),
                          "J": mp.nstr(cell_integral(center, hm), 30)})
    out = {
        "about": "Potential Green-function golden values (see make_pot_golden.py); mpmath dps=40",
        "P": pvals,
        "cells": cvals,
    }
    with open(os.path.join(HERE, "pot_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {len(pvals)} P values and {len(cvals)} cell integrals")


if __name__ == "__main__":
    main()
