"""GPU: the binned deposit's tiling (particles per deposit workgroup and per
binning workgroup) and its autotuner (sc_mesh_autotune_deposit).

  * Every tiling gives the deposit of the default tiling to round-off: the
    per-chunk sums change with the chunking, so Float32 results may differ
    in the last bit (1e-6 of max|rho|), Float64 ones to 1e-13; for one
    tiling Float32 rho is bit-reproducible.
  * Tilings whose chunks split the densest tiles many times over (2^10
    particles per chunk, 300k particles in one 16^3-cell tile) exercise the
    chunk table and the multi-chunk reduction.
  * The autotuner reports a time for every tiling, keeps the fastest and
    leaves rho equal to a plain deposit with it; the batched step after
    tuning each bunch still equals the loop of single-bunch calls bit for bit.
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


def _bunch(seed, n, T, sigma=(1e-3, 1.5e-3, 2e-3)):
    x, y, z = gaussian(n, sigma, seed)
    q = np.full(n, 1e-9 / n) * np.linspace(0.5, 1.5, n)
    return [torch.as_tensor(a, dtype=T, device=DEV) for a in (x, y, z, q)]


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).abs().max() / b.abs().max())


CASES = [((16, 16, 16), 300_000), ((64, 64, 64), 1_000_000), ((24, 20, 40), 50_000),
         ((128, 128, 128), 2_000)]


@pytest.mark.parametrize("grid,n", CASES)
@pytest.mark.parametrize("T,tol", [(torch.float32, 1e-6), (torch.float64, 1e-13)])
def test_every_tiling_gives_the_default_deposit(grid, n, T, tol):
    sc = _sc()
    x, y, z, q = _bunch(1, n, T)
    m = sc.Mesh3D(grid, x, y, z, T=T)
    assert m.deposit_tiling == (13, 12)
    sc.deposit(m, x, y, z, q)
    ref = m.rho.clone()
    worst = 0.0
    for c in sc.CHUNK_LOG2:
        for s in sc.SLICE_LOG2:
            m.set_deposit_tiling(c, s)
            sc.deposit(m, x, y, z, q)
            got = m.rho.clone()
            torch.cuda.synchronize()
            err = _rel(got, ref)
            worst = max(worst, err)
            assert err <= tol, (c, s, err)
            if T == torch.float32:
                sc.deposit(m, x, y, z, q)
                torch.cuda.synchronize()
                assert torch.equal(m.rho, got), (c, s)
    print(f"{grid} n={n} {T}: worst tiling vs default {worst:.2e}")


def test_tiling_applies_to_the_cotangent_deposit():
    """The interpolation adjoint's three-channel deposit at the smallest and
    largest chunks against the default tiling (Float64)."""
    sc = _sc()
    x, y, z, q = _bunch(2, 200_000, torch.float64)
    m = sc.Mesh3D((32, 32, 32), x, y, z, T=torch.float64)
    sc.deposit(m, x, y, z, q)
    sc.solve(m)
    gen = torch.Generator(device=DEV).manual_seed(3)
    gE = [torch.randn(x.numel(), generator=gen, dtype=torch.float64, device=DEV) for _ in range(3)]
    ref = [t.clone() for t in sc.interpolate_adjoint(m, x, y, z, *gE)]
    for c, s in ((10, 10), (14, 13)):
        m.set_deposit_tiling(c, s)
        got = sc.interpolate_adjoint(m, x, y, z, *gE)
        torch.cuda.synchronize()
        assert _rel(got[0], ref[0]) <= 1e-13, (c, s)
        for a in range(1, 4):
            assert torch.equal(got[a], ref[a]), (c, s, a)  # gathers: no tiling involved


@pytest.mark.parametrize("T", [torch.float32, torch.float64])
def test_autotuner_keeps_the_fastest_tiling(T):
    sc = _sc()
    x, y, z, q = _bunch(4, 400_000, T)
    m = sc.Mesh3D((64, 64, 64), x, y, z, T=T)
    times = sc.autotune_deposit(m, x, y, z, q, reps=3)
    assert set(times) == {(c, s) for c in sc.CHUNK_LOG2 for s in sc.SLICE_LOG2}
    assert all(t > 0.0 for t in times.values())
    best = min(times, key=times.get)
    assert m.deposit_tiling == best
    tuned = m.rho.clone()
    sc.deposit(m, x, y, z, q)
    torch.cuda.synchronize()
    if T == torch.float32:
        assert torch.equal(m.rho, tuned)
    else:
        assert _rel(m.rho, tuned) <= 1e-13
    print(f"{T}: tiling {best}, {times[best]:.4f} ms against {times[(13, 12)]:.4f} ms by default")


def test_autotuned_batch_equals_the_loop():
    sc = _sc()
    specs = [(1_500, (16, 16, 16)), (20_000, (32, 32, 32)), (300_000, (64, 64, 128)),
             (5_000, (24, 20, 40))]
    parts = [_bunch(10 + b, n, torch.float32) for b, (n, _) in enumerate(specs)]
    batch = [sc.Mesh3D(g, p[0], p[1], p[2], T=torch.float32) for (_, g), p in zip(specs, parts)]
    single = [sc.Mesh3D(g, p[0], p[1], p[2], T=torch.float32) for (_, g), p in zip(specs, parts)]
    sc.autotune_batched(batch, *[[p[k] for p in parts] for k in range(4)], reps=2)
    for mb, ms in zip(batch, single):
        ms.set_deposit_tiling(*mb.deposit_tiling)
    sc.deposit_batched(batch, *[[p[k] for p in parts] for k in range(4)])
    sc.solve_batched(batch)
    bE = sc.interpolate_batched(batch, *[[p[k] for p in parts] for k in range(3)])
    for m, p in zip(single, parts):
        sc.deposit(m, *p)
        sc.solve(m)
    sE = [sc.interpolate_field(m, p[0], p[1], p[2]) for m, p in zip(single, parts)]
    torch.cuda.synchronize()
    for b in range(len(specs)):
        assert torch.equal(batch[b].rho, single[b].rho), b
        for c in range(3):
            assert torch.equal(bE[b][c], sE[b][c]), (b, c)


def test_tiling_errors():
    sc = _sc()
    x, y, z, q = _bunch(5, 1000, torch.float32)
    m = sc.Mesh3D((8, 8, 8), x, y, z, T=torch.float32)
    for c, s in ((9, 12), (15, 12), (13, 9), (13, 14)):
        with pytest.raises(sc.SpaceChargeError, match="chunk_log2"):
            m.set_deposit_tiling(c, s)
    assert m.deposit_tiling == (13, 12)
    with pytest.raises(sc.SpaceChargeError, match="reps"):
        sc.autotune_deposit(m, x, y, z, q, reps=0)
    with pytest.raises(sc.SpaceChargeError, match="at least one particle"):
        sc.autotune_deposit(m, x[:0], y[:0], z[:0], q[:0])
