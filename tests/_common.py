"""Shared helpers for the tests: seeded particle sets, layout conversion, error norms."""
from __future__ import annotations

import numpy as np


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def gaussian(n, sigma, seed, dtype=np.float64):
    """n particles, independent normal coordinates with the given sigmas (metres)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n) * sigma[0]
    y = rng.standard_normal(n) * sigma[1]
    z = rng.standard_normal(n) * sigma[2]
    return x.astype(dtype), y.astype(dtype), z.astype(dtype)


def config_a(n=100_000, seed=0, dtype=np.float64):
    """3-D Gaussian, sigma = (1, 1, 2) mm, 1 nC in equal macro-charges."""
    x, y, z = gaussian(n, (1e-3, 1e-3, 2e-3), seed, dtype)
    q = np.full(n, 1e-9 / n, dtype=dtype)
    return x, y, z, q


def mesh_to_np(t, grid):
    """Flat device/host torch storage of a mesh array -> Fortran-ordered numpy array."""
    a = t.detach().cpu().numpy()
    if a.size == int(np.prod(grid)):
        return a.reshape(tuple(grid), order="F")
    return a.reshape(tuple(grid) + (3,), order="F")


def gaussian_field(pts, sigma, Q):
    """Electric field (V/m) of a Gaussian charge Q (C) with rms sizes sigma (m)
    at the points pts (k, 3), from 1/r = (2/sqrt(pi)) int_0^inf exp(-r^2 u^2) du
    convolved with the density:
      E_c = Q/(4 pi eps0) (2/sqrt(pi)) int_0^inf 2 u^2 x_c / (1 + 2 u^2 s_c^2)
            prod_i (1 + 2 u^2 s_i^2)^(-1/2) exp(-u^2 x_i^2 / (1 + 2 u^2 s_i^2)) du,
    by adaptive quadrature (1/(4 pi eps0) = c^2 * 1e-7)."""
    from scipy import integrate
    fpei = 299792458.0 ** 2 * 1.0e-7
    s2 = np.asarray(sigma, dtype=np.float64) ** 2
    out = np.zeros((len(pts), 3))
    for k, p in enumerate(np.asarray(pts, dtype=np.float64)):
        def common(u):
            d = 1.0 + 2.0 * u * u * s2
            return np.prod(d ** -0.5) * np.exp(-np.sum(u * u * p * p / d)), d
        for c in range(3):
            if p[c] == 0.0:
                continue

            def f(u, c=c):
                g, d = common(u)
                return 2.0 * u * u * p[c] / d[c] * g

            # the integrand peaks near u ~ 1/|p|: split there for the quadrature
            brk = 1.0 / max(np.sqrt(np.sum(p * p)), 1e-12)
            val = integrate.quad(f, 0.0, brk, epsabs=0, epsrel=1e-10, limit=200)[0]
            val += integrate.quad(f, brk, np.inf, epsabs=0, epsrel=1e-10, limit=200)[0]
            out[k, c] = fpei * Q * (2.0 / np.sqrt(np.pi)) * val
    return out
