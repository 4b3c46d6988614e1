"""GPU: the tile-order permutation (sc_sort_permutation).

It is a permutation; the particles it orders have non-decreasing deposit
tiles (16^3 cells, x-fastest tile numbering, computed here on the host from
the same Float64 arithmetic); and a Float32 step on the reordered particles
gives bit-identical rho and efield, and the same per-particle fields, as on
the original order (with host and with device-resident bounds).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from _common import gaussian

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _sc():
    import spacecharge
    return spacecharge


def _bunch(n, seed, dtype):
    x, y, z = gaussian(n, (1e-3, 1.5e-3, 3e-3), seed)
    q = np.full(n, 1e-9 / n)
    return [torch.as_tensor(a, dtype=dtype, device=DEV) for a in (x, y, z, q)]


@pytest.mark.parametrize("grid", [(32, 32, 32), (17, 40, 70), (128, 128, 128)])
def test_permutation_groups_by_tile(grid):
    sc = _sc()
    x, y, z, _ = _bunch(300_000, 1, torch.float64)
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float64)
    perm = sc.sort_permutation(mesh, x, y, z)
    torch.cuda.synchronize()
    p = perm.cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(p), np.arange(x.numel()))
    tiles = np.zeros(x.numel(), dtype=np.int64)
    ntile = [(g - 1 + 15) // 16 for g in grid]
    stride = 1
    for a, t in enumerate((x, y, z)):
        u = (t.cpu().numpy() - mesh.min_bounds[a]) / mesh.delta[a]
        cell = np.clip(np.floor(u), 0, grid[a] - 2).astype(np.int64)
        tiles += (cell // 16) * stride
        stride *= ntile[a]
    assert np.all(np.diff(tiles[p]) >= 0)


@pytest.mark.parametrize("device_bounds", [False, True])
def test_step_is_independent_of_particle_order(device_bounds):
    sc = _sc()
    grid = (64, 64, 64)
    x, y, z, q = _bunch(500_000, 2, torch.float32)
    q = q * torch.linspace(0.5, 1.5, q.numel(), device=DEV)
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float32)
    if device_bounds:
        mesh.update_bounds(x, y, z)
    perm = sc.sort_permutation(mesh, x, y, z)
    xs, ys, zs, qs = (torch.index_select(t, 0, perm) for t in (x, y, z, q))
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh)
    E = sc.interpolate_field(mesh, x, y, z)
    rho, ef = mesh.rho.clone(), mesh.efield.clone()
    sc.deposit(mesh, xs, ys, zs, qs)
    sc.solve(mesh)
    Es = sc.interpolate_field(mesh, xs, ys, zs)
    torch.cuda.synchronize()
    assert torch.equal(mesh.rho, rho)
    assert torch.equal(mesh.efield, ef)
    for c in range(3):
        assert torch.equal(Es[c], torch.index_select(E[c], 0, perm))
