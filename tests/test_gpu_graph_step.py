"""GPU: the single-bunch step (deposit + solve + interpolate on a mesh with
host bounds) captured into a graph.

The Green-function build of the solve forks onto the library's second stream
where the deposit starts and rejoins the caller's stream before the z
convolution, in the eager step and inside a capture alike.  Each replay must
equal the eager step bit for bit (Float32 mesh: fixed-point deposit, no
atomics elsewhere), also with eager work between the replays: eager steps,
and an eager solve issued right behind a replay with no deposit of its own,
whose Green build must then wait for the replay's z convolution before it
rewrites the spectra that convolution reads.
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
    x, y, z = gaussian(n, (1e-3, 1e-3, 2e-3), seed)
    q = np.full(n, 1e-9 / n)
    return [torch.as_tensor(a, dtype=dtype, device=DEV) for a in (x, y, z, q)]


def _step(sc, mesh, x, y, z, q, cathode=False):
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh, at_cathode=cathode)
    return sc.interpolate_field(mesh, x, y, z)


@pytest.mark.parametrize("cathode", [False, True])
def test_graph_step_equals_eager_step(cathode):
    sc = _sc()
    grid = (64, 64, 64)
    x0, y0, z0, q = _bunch(400_000, 21, torch.float32)
    if cathode:
        z0 = z0 - z0.min() + 1e-5
    x, y, z = x0.clone(), y0.clone(), z0.clone()
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float32)
    ref = sc.Mesh3D(grid, x, y, z, T=torch.float32)   # same bounds, eager only
    for _ in range(2):  # workspace, bins and the second stream exist before the capture
        _step(sc, mesh, x, y, z, q, cathode)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        Eg = _step(sc, mesh, x, y, z, q, cathode)
    for rep in range(4):
        scale = (1.0, 0.9, 1.1, 1.0)[rep]
        x.copy_(x0 * scale)
        y.copy_(y0 * scale)
        z.copy_(z0 * scale if not cathode else z0)
        torch.cuda.synchronize()
        print(f"replay {rep} cathode={cathode}", flush=True)
        graph.replay()
        if rep % 2:
            # eager solve queued behind the replay, no deposit of its own
            sc.solve(mesh, at_cathode=cathode)
            Es = mesh.efield.clone()
        else:
            _step(sc, mesh, x, y, z, q, cathode)
        torch.cuda.synchronize()
        Eh = _step(sc, ref, x, y, z, q, cathode)
        torch.cuda.synchronize()
        for c in range(3):
            assert torch.equal(Eg[c], Eh[c]), (rep, c)
        if rep % 2:
            assert torch.equal(Es, ref.efield), rep
