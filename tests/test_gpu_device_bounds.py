"""GPU: device-resident bounds (sc_mesh_auto_bounds, NULL bounds in the step).

  * The device computes the Mesh3D particle-extrema bounds (src/mesh.jl:
    95-174) with the same bits as the host path (sc_bounds_from_particles),
    for every particle/mesh type pair, including the zero-spacing patch.
  * A step with device-resident bounds equals the step with the same bounds
    passed from the host: bit for bit on Float32 meshes (deposit in fixed
    point, no atomics elsewhere), to 1e-13 on Float64 ones (deposit atomics
    in varying order); also the cathode, potential and adjoint paths.
  * A tracking loop that re-meshes from the particles every step, and the
    same loop captured in a graph and replayed (positions updated in place
    between replays), equal the eager host-bounds loop.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

from _common import gaussian

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TYPES = [(torch.float64, torch.float64), (torch.float32, torch.float32),
         (torch.float64, torch.float32), (torch.float32, torch.float64)]


def _sc():
    import spacecharge
    return spacecharge


def _bunch(n, seed, dtype, sigma=(1e-3, 1.5e-3, 3e-3), shift=(0.0, 0.0, 0.0)):
    x, y, z = gaussian(n, sigma, seed)
    q = np.full(n, 1e-9 / n)
    arrs = (x + shift[0], y + shift[1], z + shift[2], q)
    return [torch.as_tensor(a, dtype=dtype, device=DEV) for a in arrs]


def _close(a, b, exact):
    if exact:
        return torch.equal(a, b)
    a, b = a.double(), b.double()
    return float((a - b).abs().max()) <= 1e-13 * float(b.abs().max())


@pytest.mark.parametrize("pt,mt", TYPES)
def test_device_bounds_equal_host_bounds_bitwise(pt, mt):
    sc = _sc()
    cases = [_bunch(50_000, 1, pt),
             _bunch(30_000, 2, pt, shift=(0.3, -1.7, 12.5)),      # far from the origin
             _bunch(1, 3, pt)]                                   # one particle: zero patch
    x, y, z, q = _bunch(20_000, 4, pt)
    cases.append([x, torch.full_like(y, 2.5e-3), z, q])          # flat in y: zero patch
    for grid in ((32, 32, 32), (17, 9, 40)):
        for x, y, z, _ in cases:
            host = sc.Mesh3D(grid, x, y, z, T=mt)
            dev = sc.Mesh3D(grid, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), T=mt)
            dev.update_bounds(x, y, z)
            lo, hi, dl = dev.fetch_bounds()
            assert lo == host.min_bounds and hi == host.max_bounds and dl == host.delta, \
                (grid, x.numel(), lo, host.min_bounds, dl, host.delta)


@pytest.mark.parametrize("pt,mt", TYPES)
@pytest.mark.parametrize("cath", [False, True])
def test_step_with_device_bounds_equals_host_bounds(pt, mt, cath):
    sc = _sc()
    grid = (32, 32, 64)
    x, y, z, q = _bunch(200_000, 5, pt, shift=(0.0, 0.0, 6e-3 if cath else 0.0))
    exact = mt == torch.float32
    host = sc.Mesh3D(grid, x, y, z, T=mt, gamma=2.0)
    dev = sc.Mesh3D(grid, (-1.0,) * 3, (1.0,) * 3, T=mt, gamma=2.0)
    dev.update_bounds(x, y, z)
    outs = []
    for m in (host, dev):
        sc.deposit(m, x, y, z, q)
        sc.solve(m, at_cathode=cath)
        E = sc.interpolate_field(m, x, y, z)
        phi = sc.solve_potential(m).clone()
        # random cotangents (uniform ones would make g_x the bunch's net
        # self-force, which cancels to round-off)
        gen = torch.Generator(device=DEV).manual_seed(10)
        gE = [torch.randn(x.numel(), generator=gen, device=DEV, dtype=x.dtype) for _ in range(3)]
        gef, gx1, _, _ = sc.interpolate_adjoint(m, x, y, z, *gE)
        g_rho = sc.solve_adjoint(m, gef, at_cathode=cath)
        gq, gx2, _, _ = sc.deposit_adjoint(m, x, y, z, q, g_rho)
        outs.append((m.rho.clone(), m.efield.clone(), E, phi, gef.clone(), g_rho.clone(), gq,
                     gx1 + gx2))
    torch.cuda.synchronize()
    names = ("rho", "efield", "E", "phi", "g_efield", "g_rho", "g_q", "g_x")
    for name, a, b in zip(names, outs[1], outs[0]):
        if isinstance(a, tuple):
            for c in range(3):
                assert _close(a[c], b[c], exact), (name, c)
        else:
            assert _close(a, b, exact), name


def _step_host(sc, mesh, x, y, z, q):
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh)
    return sc.interpolate_field(mesh, x, y, z)


def test_tracking_loop_remeshes_on_device():
    """Five drift steps, new bounds from the particles each step: device
    bounds (Green reuse on, which device bounds must not use) against a
    fresh host-bounds mesh per step, Float32, bit for bit."""
    sc = _sc()
    grid = (32, 32, 32)
    x, y, z, q = _bunch(100_000, 6, torch.float32)
    v = [t.clone() * 0.05 for t in (x, y, z)]
    dev = sc.Mesh3D(grid, x, y, z, T=torch.float32)
    dev.set_green_reuse(True)
    for step in range(5):
        dev.update_bounds(x, y, z)
        Ed = _step_host(sc, dev, x, y, z, q)
        host = sc.Mesh3D(grid, x, y, z, T=torch.float32)
        Eh = _step_host(sc, host, x, y, z, q)
        for c in range(3):
            assert torch.equal(Ed[c], Eh[c]), (step, c)
        x, y, z = x + v[0] * (step + 1), y + v[1], z - v[2]


def test_graph_captured_step_with_device_bounds():
    """update_bounds + deposit + solve + interpolate captured once, replayed
    with new positions copied into the captured inputs; each replay equals
    the eager host-bounds step, with eager work between replays."""
    sc = _sc()
    grid = (32, 32, 64)
    n = 100_000
    x, y, z, q = _bunch(n, 7, torch.float32)
    xs, ys, zs = x.clone(), y.clone(), z.clone()
    mesh = sc.Mesh3D(grid, xs, ys, zs, T=torch.float32)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):  # allocates the workspace, bins and scratch outside the capture
            mesh.update_bounds(xs, ys, zs)
            _step_host(sc, mesh, xs, ys, zs, q)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        mesh.update_bounds(xs, ys, zs)
        Eg = _step_host(sc, mesh, xs, ys, zs, q)
    rng = np.random.default_rng(8)
    for rep in range(4):
        scale = 1.0 + 0.25 * rep
        shift = rng.normal(0, 1e-3, 3)
        for dst, src, sh in zip((xs, ys, zs), (x, y, z), shift):
            dst.copy_(src * scale + float(sh))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        ref_mesh = sc.Mesh3D(grid, xs, ys, zs, T=torch.float32)
        Eh = _step_host(sc, ref_mesh, xs, ys, zs, q)
        torch.cuda.synchronize()
        for c in range(3):
            assert torch.equal(Eg[c], Eh[c]), (rep, c)


def test_device_bounds_errors():
    sc = _sc()
    from spacecharge import _lib
    x, y, z, q = _bunch(1000, 9, torch.float64)
    m = sc.Mesh3D((8, 8, 8), (-1.0,) * 3, (1.0,) * 3, T=torch.float64)
    lib = _lib.load()
    # no device-resident bounds yet
    assert lib.sc_deposit(m._h, 8, x.data_ptr(), y.data_ptr(), z.data_ptr(), q.data_ptr(), 1000,
                          None, None, 1) != 0
    assert "device-resident" in lib.sc_last_error(sc._ctx(m.device)).decode()
    lo, hi, dl = (_lib.d3((0.0,) * 3) for _ in range(3))
    assert lib.sc_mesh_read_bounds(m._h, lo, hi, dl) != 0
    with pytest.raises(sc.SpaceChargeError, match="empty"):
        m.update_bounds(x[:0], y[:0], z[:0])
    m.update_bounds(x, y, z)
    # mixed NULL and host bounds in a solve
    assert lib.sc_solve(m._h, None, _lib.d3((1.0,) * 3), None, 1.0, 0) != 0
    m2 = sc.Mesh3D((8, 8, 8), x, y, z, T=torch.float64)
    with pytest.raises(sc.SpaceChargeError, match="host bounds"):
        sc.solve_batched([m, m2])
