"""Generate tests/golden/igf_golden.json: high-precision Green-function vectors.

Two independent families of values, both at 40 significant digits (mpmath):

1. The field antiderivative of src/green_functions.jl:35-38,
      F(x, y, z) = x*atan(yz/(r x)) - z*log(r + y) + y*log((r - z)/(r + z))/2,
   at sample points, together with the check that its mixed third
   derivative is x/r^3 there (the property that makes it the IGF kernel).
2. Cell integrals  I_c(a, b, k) = integral over the cell
      [(a-1/2)dx, (a+1/2)dx] x [(b-1/2)dy, (b+1/2)dy] x [(k-1/2)dz, (k+1/2)dz]
   of x/r^3 (c = x), y/r^3 (c = y), z/r^3 (c = z).  The integral along the
   component's own axis is done in closed form (d/dx of -1/r is x/r^3) and
   the remaining 2-D integral of 1/r by adaptive quadrature, so these values
   do not use F at all.

Run:  python tests/golden/make_igf_golden.py   (takes about a minute)
"""
from __future__ import annotations

import json
import os

import mpmath as mp

mp.mp.dps = 40

HERE = os.path.dirname(os.path.abspath(__file__))


def F(x, y, z):
    r = mp.sqrt(x * x + y * y + z * z)
    return x * mp.atan((y * z) / (r * x)) - z * mp.log(r + y) + y * mp.log((r - z) / (r + z)) / 2


def cell_integral(comp, center, h):
    """Integral of (component)/r^3 over the cell with the given center and sides."""
    lo = [c - s / 2 for c, s in zip(center, h)]
    hi = [c + s / 2 for c, s in zip(center, h)]
    # the component's own axis first: int q/r^3 dq = -1/r
    ax = comp
    others = [i for i in range(3) if i != ax]

    def inner(u, v):
        p = [None, None, None]
        p[others[0]] = u
        p[others[1]] = v
        p[ax] = hi[ax]
        r1 = mp.sqrt(sum(t * t for t in p))
        p[ax] = lo[ax]
        r0 = mp.sqrt(sum(t * t for t in p))
        return 1 / r0 - 1 / r1

    return mp.quad(inner, [lo[others[0]], hi[others[0]]], [lo[others[1]], hi[others[1]]])


def main():
    points = [
        (0.3, 0.2, 0.1), (-0.5, 0.7, 1.1), (2.5, -1.5, 0.5), (0.05, 3.0, -2.0),
        (10.5, 7.5, -3.5), (-31.5, 0.5, 15.5), (1e-3, 2e-3, -1.5e-3),
    ]
    fvals = []
    for p in points:
        x, y, z = (mp.mpf(t) for t in p)
        f = F(x, y, z)
        d3 = mp.diff(lambda a, b, c: F(a, b, c), (x, y, z), (1, 1, 1))
        r = mp.sqrt(x * x + y * y + z * z)
        assert abs(d3 - x / r ** 3) <= mp.mpf("1e-25") * abs(x / r ** 3), p
        fvals.append({"point": list(p), "F": mp.nstr(f, 36)})

    spacings = [(1.0, 1.0, 1.0), (0.7, 1.3, 2.1)]
    cells = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1), (3, -2, 5), (-7, 4, -1),
             (12, 9, 6), (-20, 15, 10)]
    cvals = []
    for h in spacings:
        for cell in cells:
            center = [mp.mpf(a) * mp.mpf(s) for a, s in zip(cell, h)]
            ints = [cell_integral(c, center, [mp.mpf(s) for s in h]) for c in range(3)]
            cvals.append({"h": list(h), "cell": list(cell), "I": [mp.nstr(v, 30) for v in ints]})
    out = {
        "about": "Green-function golden values (see make_igf_golden.py); mpmath dps=40",
        "F": fvals,
        "cells": cvals,
    }
    with open(os.path.join(HERE, "igf_golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {len(fvals)} F values and {len(cvals)} cell integrals")


if __name__ == "__main__":
    main()
