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


def _batch_run(sc, meshes, parts, clear=True, at_cathode=False):
    sc.deposit_batched(meshes, *[[p[k] for p in parts] for k in range(4)], clear=clear)
    rho = [m.rho.clone() for m in meshes]
    sc.solve_batched(meshes, at_cathode=at_cathode)
    E = sc.interpolate_batched(meshes, *[[p[k] for p in parts] for k in range(3)])
    torch.cuda.synchronize()
    return rho, E


@pytest.mark.parametrize("cath", ["free", "cathode", "mixed"])
@pytest.mark.parametrize("T", [torch.float32, torch.float64])
def test_grouped_launch_equals_stream_pool(T, cath):
    """20 uneven bunches (three descriptor-store launches) with per-bunch
    deposit tilings, meshes of four shapes (three of them form classes of
    five for the grouped solve; 24x20x40 has lengths the register engine
    does not take): the grouped deposit, solve and interpolation (one launch
    per phase or kernel) against the per-bunch launch sequences on the
    stream pool, with clear=True and clear=False, in free space, with the
    cathode image, and with both (classes split by the flag); bit for bit in
    Float32, 1e-12 in Float64."""
    sc = _sc()
    rng = np.random.default_rng(5)
    specs = [(int(10 ** rng.uniform(3, 5.5)), [(16, 16, 16), (32, 32, 32), (24, 20, 40),
                                              (64, 64, 64)][b % 4]) for b in range(20)]
    parts = [_bunch(300 + b, n, T) for b, (n, _) in enumerate(specs)]
    tilings = [(10 + b % 5, 10 + b % 4) for b in range(20)]

    def meshes():
        ms = [sc.Mesh3D(g, p[0], p[1], p[2], T=T) for (_, g), p in zip(specs, parts)]
        for m, t in zip(ms, tilings):
            m.set_deposit_tiling(*t)
        return ms

    flags = {"free": False, "cathode": True, "mixed": [b % 2 == 1 for b in range(20)]}[cath]
    try:
        out = {}
        for grouped in (True, False):
            sc.set_grouped_launch(grouped)
            ms = meshes()
            rho1, E1 = _batch_run(sc, ms, parts, at_cathode=flags)
            rho2, _ = _batch_run(sc, ms, parts, clear=False, at_cathode=flags)  # adds to rho
            out[grouped] = (rho1, E1, rho2)
    finally:
        sc.set_grouped_launch(True)
    (ga, Ea, ra), (gb, Eb, rb) = out[True], out[False]
    for b in range(len(specs)):
        for a, c in ((ga[b], gb[b]), (ra[b], rb[b])) + tuple((Ea[b][k], Eb[b][k]) for k in range(3)):
            if T == torch.float32:
                assert torch.equal(a, c), b
            else:
                assert _rel(a, c) <= 1e-12, b


def test_batches_that_do_not_group_fall_back():
    """A bunch without particles, or meshes of two types, take the per-bunch
    path and still equal the single-bunch calls."""
    sc = _sc()
    parts = [_bunch(400 + b, n, torch.float32) for b, n in enumerate((5000, 8000, 3000))]
    empty = [t[:0] for t in parts[1]]
    cases = [([torch.float32] * 3, [parts[0], empty, parts[2]]),
             ([torch.float32, torch.float64, torch.float32], parts)]
    for types, ps in cases:
        ms = [sc.Mesh3D((16, 16, 16), (-5e-3,) * 3, (5e-3,) * 3, T=T) for T in types]
        ref = [sc.Mesh3D((16, 16, 16), (-5e-3,) * 3, (5e-3,) * 3, T=T) for T in types]
        rho, E = _batch_run(sc, ms, ps)
        for m, p, r, e in zip(ref, ps, rho, E):
            sc.deposit(m, *p)
            sc.solve(m)
            Es = sc.interpolate_field(m, p[0], p[1], p[2])
            torch.cuda.synchronize()
            if m.T == torch.float32:
                assert torch.equal(r, m.rho)
                for c in range(3):
                    assert torch.equal(e[c], Es[c])
            else:  # Float64 deposits add in LDS-atomic order
                assert _rel(r, m.rho) <= 1e-12
                for c in range(3):
                    assert _rel(e[c], Es[c]) <= 1e-12


def test_grouped_step_captured_in_a_graph():
    """Twelve bunches of two shapes (classes of six for the grouped solve):
    the batched step captured once and replayed with new positions, an eager
    grouped step on other meshes of the same context between replays (it
    rewrites the descriptor tables): every replay equals the eager step bit
    for bit (Float32)."""
    sc = _sc()
    T = torch.float32
    specs = [(20_000 + 5_000 * b, (32, 32, 32) if b % 2 else (16, 16, 32)) for b in range(12)]
    parts = [_bunch(500 + b, n, T) for b, (n, _) in enumerate(specs)]
    live = [[t.clone() for t in p] for p in parts]
    meshes = [sc.Mesh3D(g, p[0], p[1], p[2], T=T) for (_, g), p in zip(specs, live)]
    eager = [sc.Mesh3D(g, p[0], p[1], p[2], T=T) for (_, g), p in zip(specs, live)]

    def step(ms):
        sc.deposit_batched(ms, *[[p[k] for p in live] for k in range(4)])
        sc.solve_batched(ms, at_cathode=[b % 3 == 0 for b in range(12)])
        return sc.interpolate_batched(ms, *[[p[k] for p in live] for k in range(3)])

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step(meshes)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        Eg = step(meshes)
    for rep in range(3):
        for p, p0 in zip(live, parts):
            for k in range(3):
                p[k].copy_(p0[k] * (1.0 - 0.03 * rep))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        Ee = step(eager)
        torch.cuda.synchronize()
        for b in range(12):
            assert torch.equal(meshes[b].efield, eager[b].efield), (rep, b)
            for c in range(3):
                assert torch.equal(Eg[b][c], Ee[b][c]), (rep, b, c)


@pytest.mark.parametrize("nb", [12, 256, 300])
def test_batched_edge_sizes_equal_the_loop(nb):
    """Bunches of one particle and meshes of 2x2x2 among ordinary ones, and
    batch sizes at (256: 32 descriptor-store launches) and past (300: the
    per-bunch path) the grouped-launch limit: Float32 results bit-identical
    to the single-bunch loop."""
    sc = _sc()
    T = torch.float32
    grids = [(2, 2, 2), (3, 2, 5), (8, 8, 8), (16, 16, 16)]
    specs = [(1 if b % 5 == 0 else 50 + 37 * (b % 7), grids[b % 4]) for b in range(nb)]
    parts = [_bunch(700 + b, n, T) for b, (n, _) in enumerate(specs)]
    bounds = ((-6e-3,) * 3, (6e-3,) * 3)
    batch = [sc.Mesh3D(g, *bounds, T=T) for _, g in specs]
    single = [sc.Mesh3D(g, *bounds, T=T) for _, g in specs]
    rho, E = _batch_run(sc, batch, parts)
    for m, p, r, e in zip(single, parts, rho, E):
        sc.deposit(m, *p)
        sc.solve(m)
        Es = sc.interpolate_field(m, p[0], p[1], p[2])
        torch.cuda.synchronize()
        assert torch.equal(r, m.rho)
        for c in range(3):
            assert torch.equal(e[c], Es[c])


def test_grouped_interpolation_node_pair_layout_matches_oracle():
    """Float64 meshes whose interpolation copy uses node pairs along x (the
    layout the library picks for Float64 meshes) and meshes forced to one
    entry per node, in one grouped
    interpolation: every bunch's fields bit-identical to the oracle's
    interpolation of that mesh's own efield, particles on and beyond the
    upper x face included."""
    import oracle as orc
    sc = _sc()
    T = torch.float64
    specs = [(3_000, (16, 16, 16)), (20_000, (32, 32, 32)), (7_000, (24, 20, 40)),
             (50_000, (64, 64, 64)), (1_000, (9, 14, 5))]
    parts = [_bunch(500 + b, n, T) for b, (n, _) in enumerate(specs)]
    meshes = [sc.Mesh3D(g, p[0], p[1], p[2], T=T) for (_, g), p in zip(specs, parts)]
    for b, (m, p) in enumerate(zip(meshes, parts)):
        m.set_field_layout(2 if b % 2 == 0 else 1)
        hi, dl = m.max_bounds, m.delta
        p[0][:2] = torch.tensor([hi[0], hi[0] + 0.5 * dl[0]], dtype=T)
    sc.deposit_batched(meshes, *[[p[k] for p in parts] for k in range(4)])
    sc.solve_batched(meshes)
    E = sc.interpolate_batched(meshes, *[[p[k] for p in parts] for k in range(3)])
    torch.cuda.synchronize()
    for b, (m, p) in enumerate(zip(meshes, parts)):
        ef = np.asfortranarray(m.efield.cpu().numpy())
        xs, ys, zs = (t.cpu().numpy() for t in p[:3])
        ref = orc.interpolate(ef, xs, ys, zs, m.min_bounds, m.delta)
        for c in range(3):
            assert np.array_equal(E[b][c].cpu().numpy(), ref[c]), (b, c)
