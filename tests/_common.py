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
