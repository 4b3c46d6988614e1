"""GPU: the batched calls (many bunches per call on the library's stream pool)
against a loop of the single-bunch calls on identically built meshes.

Float32 meshes are compared bit for bit: their deposit accumulates in 64-bit
fixed point (exact, so independent of the order in which the chunk atomics
land), and the solve and the interpolation have no atomics.  Float64 meshes
accumulate the deposit with Float64 atomics, whose order varies from run to
run, so they are compared to 1e-12 relative.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from _common import gaussian

pytestmark = pytest.mark.gpu

DEV = "cuda:0"

# uneven bunches: particle counts and grids of the batched configuration,
# including a non-power-of-two grid (generic transform engine)
BUNCHES = [(1_500, (16, 16, 16)), (20_000, (32, 32, 32)), (200_000, (64, 64, 128)),
           (5_000, (24, 20, 40)), (80_000, (8, 8, 8)), (300_000, (32, 32, 64))]


def _sc():
    import spacecharge
    return spacecharge


def _bunch(b, n, T):
    rng = np.random.default_rng(100 + b)
    sig = rng.uniform(0.5e-3, 3e-3, 3)
    x, y, z = gaussian(n, tuple(sig), 200 + b)
    q = np.full(n, 1e-10 / n)
    return [torch.as_tensor(a, dtype=T, device=DEV) for a in (x, y, z, q)]


def _meshes(parts, T, gamma):
    sc = _sc()
    return [sc.Mesh3D(grid, p[0], p[1], p[2], T=T, gamma=gamma)
            for (n, grid), p in zip(BUNCHES, parts)]


def _run_pair(T, gamma=1.0, cathode=False):
    sc = _sc()
    parts = [_bunch(b, n, T) for b, (n, _) in enumerate(BUNCHES)]
    flags = [bool(cathode) and b % 2 == 0 for b in range(len(BUNCHES))]
    single = _meshes(parts, T, gamma)
    single_E = []
    for m, p, f in zip(single, parts, flags):
        sc.deposit(m, *p)
        sc.solve(m, at_cathode=f)
        single_E.append(sc.interpolate_field(m, p[0], p[1], p[2]))
    batch = _meshes(parts, T, gamma)
    sc.deposit_batched(batch, *[[p[k] for p in parts] for k in range(4)])
    sc.solve_batched(batch, at_cathode=flags)
    batch_E = sc.interpolate_batched(batch, *[[p[k] for p in parts] for k in range(3)])
    torch.cuda.synchronize()
    return single, single_E, batch, batch_E


def _rel(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).abs().max() / b.abs().max().clamp_min(1e-300))


def test_batched_float32_bit_identical_to_loop():
    single, sE, batch, bE = _run_pair(torch.float32)
    for b in range(len(BUNCHES)):
        assert torch.equal(batch[b].rho, single[b].rho), f"rho, bunch {b}"
        assert torch.equal(batch[b].efield, single[b].efield), f"efield, bunch {b}"
        for c in range(3):
            assert torch.equal(bE[b][c], sE[b][c]), f"E[{c}], bunch {b}"


def test_batched_float32_cathode_and_gamma():
    single, sE, batch, bE = _run_pair(torch.float32, gamma=5.0, cathode=True)
    for b in range(len(BUNCHES)):
        assert torch.equal(batch[b].efield, single[b].efield), f"efield, bunch {b}"
        for c in range(3):
            assert torch.equal(bE[b][c], sE[b][c]), f"E[{c}], bunch {b}"


def test_batched_float64_matches_loop():
    single, sE, batch, bE = _run_pair(torch.float64, cathode=True)
    worst = 0.0
    for b in range(len(BUNCHES)):
        worst = max(worst, _rel(batch[b].rho, single[b].rho), _rel(batch[b].efield, single[b].efield))
        for c in range(3):
            worst = max(worst, _rel(bE[b][c], sE[b][c]))
    print(f"batched vs loop, Float64: max rel diff {worst:.2e}")
    assert worst <= 1e-12


def test_float32_deposit_is_deterministic():
    """Two deposits of the same particles give bit-identical Float32 rho (the
    fixed-point chunk sums do not depend on the order of the LDS atomics)."""
    sc = _sc()
    x, y, z, q = _bunch(9, 400_000, torch.float32)
    m = sc.Mesh3D((64, 64, 64), x, y, z, T=torch.float32)
    sc.deposit(m, x, y, z, q)
    first = m.rho.clone()
    for _ in range(3):
        sc.deposit(m, x, y, z, q)
        torch.cuda.synchronize()
        assert torch.equal(m.rho, first)


def test_batched_argument_errors():
    sc = _sc()
    x, y, z, q = _bunch(0, 1000, torch.float32)
    m = sc.Mesh3D((8, 8, 8), x, y, z, T=torch.float32)
    with pytest.raises(sc.SpaceChargeError, match="twice"):
        sc.deposit_batched([m, m], [x, x], [y, y], [z, z], [q, q])
    with pytest.raises(sc.SpaceChargeError, match="at least one"):
        sc.solve_batched([])
    with pytest.raises(sc.SpaceChargeError, match="one flag per mesh"):
        sc.solve_batched([m], at_cathode=[True, False])


def test_batched_step_captured_in_a_graph():
    """The batched step (stream pool with fork/join events) captured into one
    graph and replayed with new positions copied into the captured inputs,
    eager work in between: every replay equals the eager batched step
    (Float32, bit for bit)."""
    sc = _sc()
    T = torch.float32
    parts = [_bunch(b, n, T) for b, (n, _) in enumerate(BUNCHES)]
    live = [[t.clone() for t in p] for p in parts]
    meshes = _meshes(live, T, 1.0)
    eager_meshes = _meshes(live, T, 1.0)

    def step(ms, ps):
        sc.deposit_batched(ms, *[[p[k] for p in ps] for k in range(4)])
        sc.solve_batched(ms)
        return sc.interpolate_batched(ms, *[[p[k] for p in ps] for k in range(3)])

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(meshes, live)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        Eg = step(meshes, live)
    for rep in range(3):
        for p, p0 in zip(live, parts):
            for k in range(3):
                p[k].copy_(p0[k] * (1.0 + 0.05 * rep))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        Ee = step(eager_meshes, live)
        torch.cuda.synchronize()
        for b in range(len(BUNCHES)):
            for c in range(3):
                assert torch.equal(Eg[b][c], Ee[b][c]), (rep, b, c)
