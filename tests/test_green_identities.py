"""CPU: the identities the Float32 Green build (k_green_f3 in sc_green.hip)
uses to evaluate the three field components of a node from shared terms.

With f(x, y, z) = x atan(yz/(rx)) - z log(r + y) + y/2 log((r - z)/(r + z))
(the field antiderivative of src/green_functions.jl:35-38) and the component
permutation F_x = f(u, v, w), F_y = f(v, w, u), F_z = f(w, u, v), the kernel
forms
  * log((r - w)/(r + w)) as log(u^2 + v^2) - 2 log(r + w), cyclically;
  * atan(uv/(rw)) as sign(uvw) pi/2 - atan(vw/(ru)) - atan(wu/(rv)).
These tests restate those formulas in numpy and compare them with 40-digit
mpmath values of f, at nodes like the kernel's (half-integer multiples of
the spacings, image-side z offsets, a zero z coordinate).  The restated
formula is checked against f itself, not against a copy of the kernel's.
"""
from __future__ import annotations

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")
mp.mp.dps = 40


def f_exact(x, y, z):
    x, y, z = mp.mpf(x), mp.mpf(y), mp.mpf(z)
    r = mp.sqrt(x * x + y * y + z * z)
    t1 = 0 if x == 0 else x * mp.atan(y * z / (r * x))
    return t1 - z * mp.log(r + y) + y * mp.log((r - z) / (r + z)) / 2


def shared_terms(u, v, w):
    """The three components from shared logs and two atans."""
    uu, vv, ww = u * u, v * v, w * w
    r = np.sqrt((uu + vv) + ww)
    Lu, Lv, Lw = np.log(r + u), np.log(r + v), np.log(r + w)
    Quv, Qvw, Qwu = np.log(uu + vv), np.log(vv + ww), np.log(ww + uu)
    ax = np.arctan((v * w) / (r * u))
    ay = np.arctan((w * u) / (r * v))
    if w == 0.0:
        hp = 0.0
    else:
        hp = -np.pi / 2 if (np.signbit(u) != np.signbit(v)) != np.signbit(w) else np.pi / 2
    az = (hp - ax) - ay
    return ((u * ax - w * Lv) + (v * (Quv - 2 * Lw)) * 0.5,
            (v * ay - u * Lw) + (w * (Qvw - 2 * Lu)) * 0.5,
            (w * az - v * Lu) + (u * (Qwu - 2 * Lv)) * 0.5)


def _nodes(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(n):
        d = rng.uniform(1e-5, 1e-3, 3)
        i, j = rng.integers(0, 130, 2)
        k = int(rng.integers(-128, 130))
        u, v = (i - 0.5) * d[0], (j - 0.5) * d[1]
        w = (k + 0.5) * d[2] + (rng.uniform(-0.05, 0.05) if t % 3 == 0 else 0.0)
        out.append((u, v, w))
    out.append((-0.5e-4, 0.5e-4, 0.0))      # zero z coordinate (image side)
    out.append((-0.5e-4, -0.5e-4, 3e-3))    # both in-plane coordinates negative
    return out


def test_three_components_from_shared_terms():
    worst = 0.0
    for u, v, w in _nodes(600, 1):
        got = shared_terms(u, v, w)
        ref = (f_exact(u, v, w), f_exact(v, w, u), f_exact(w, u, v))
        scale = 10.0 * float(np.sqrt(u * u + v * v + w * w))
        for g, e in zip(got, ref):
            worst = max(worst, abs(g - float(e)) / scale)
    # the one-component Float64 evaluation of f reaches the same level
    assert worst < 5e-15, worst


@pytest.mark.parametrize("sign", [(1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1), (-1, -1, -1)])
def test_atan_sum_is_signed_half_pi(sign):
    rng = np.random.default_rng(2)
    for _ in range(200):
        u, v, w = (mp.mpf(s * a) for s, a in zip(sign, rng.uniform(1e-4, 1e-2, 3)))
        r = mp.sqrt(u * u + v * v + w * w)
        tot = (mp.atan(v * w / (r * u)) + mp.atan(w * u / (r * v)) + mp.atan(u * v / (r * w)))
        assert abs(tot - mp.sign(u * v * w) * mp.pi / 2) < mp.mpf("1e-35")


def test_quotient_log_identity():
    rng = np.random.default_rng(3)
    for _ in range(200):
        u, v, w = (mp.mpf(a) for a in rng.uniform(-1e-2, 1e-2, 3))
        r = mp.sqrt(u * u + v * v + w * w)
        lhs = mp.log((r - w) / (r + w))
        rhs = mp.log(u * u + v * v) - 2 * mp.log(r + w)
        assert abs(lhs - rhs) < mp.mpf("1e-35") * (1 + abs(lhs))
