"""GPU: every workload configuration, end to end on its own sizes, against the oracle.

  A  100,000 Gaussian particles (sigma = 1, 1, 2 mm), 1 nC, on a 32^3 mesh
     with manual bounds at +-5 sigma, gamma = 1, free space: deposit, solve,
     interpolate.  The seed puts particles outside the mesh, so the clamp
     policy runs too.  Float64: rho within 1e-13 of the largest node value,
     mesh and particle fields within 1e-12 rel. L2 of the Float64 oracle per
     component, and at least 3x closer than the Float64 oracle to the same
     algorithm in extended precision.  Float32 (particles and mesh): 1e-5
     against the Float64 oracle.
  B  4,000,000 particles (Gaussian + uniform ellipsoid), 128^3 from the
     particle extrema, Float32, gamma = 1 and 10: mesh and particle fields
     within 1e-5 rel. L2 of the Float64 oracle per component.
  C  2,000,000 cathode-emitted particles (uniform disk, exponential z), 100
     pC, 64x64x128, image charge on: Float32 within 1e-5 of the Float64
     oracle; Float64 with the extended-precision gate of A.
  D  32 bunches (Np log-uniform in [1e4, 2e6], meshes 16^3 .. 128^3),
     Float32, through the grouped batched calls and through a graph replay of
     the same step: every bunch's mesh and particle fields within 1e-5 of the
     Float64 oracle; the replay equal to the eager step bit for bit.
  E  1,000,000 Gaussian particles, 64^3 with fixed bounds, 100 drift-kick
     steps checkpointed every 10, Float64, loss = final rms emittance: the
     directional derivative along a random direction in all positions and
     charges against central differences of the whole run (h = 1e-6 of the
     bunch size and of each charge).
  Device bounds: sc_mesh_auto_bounds gives the oracle's bounds_auto bit for
     bit (all particle/mesh type pairs, zero-spacing patch included).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest
import torch

import oracle as orc
from _common import config_a, gaussian, mesh_to_np, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
THREADS = min(16, os.cpu_count() or 1)


def _sc():
    import spacecharge
    return spacecharge


def _bench():
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import bench
    return bench


def _np64(ts):
    return [t.detach().cpu().numpy().astype(np.float64) for t in ts]


def _oracle_step(x, y, z, q, grid, lo, hi, dl, gamma=1.0, cathode=False, extended=False):
    """Float64 oracle deposit -> solve -> interpolate (and, if asked, the
    solve and interpolation in extended precision on the same rho)."""
    rho = orc.deposit(x, y, z, q, grid, lo, dl, np.float64)
    ef = orc.solve(rho, lo, hi, dl, gamma, cathode, nthreads=THREADS)
    E = orc.interpolate(ef, x, y, z, lo, dl, nthreads=THREADS)
    if not extended:
        return rho, ef, E
    tr = orc.solve_extended(rho, lo, hi, dl, gamma, cathode, nthreads=THREADS)
    Etr = orc.interpolate(np.asfortranarray(tr), x, y, z, lo, dl, nthreads=THREADS)
    return rho, ef, E, tr, Etr


def _gpu_step(mesh, x, y, z, q, cathode=False):
    sc = _sc()
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh, at_cathode=cathode)
    E = sc.interpolate_field(mesh, x, y, z)
    torch.cuda.synchronize()
    grid = mesh.grid_size
    return (mesh_to_np(mesh._rho, grid).astype(np.float64),
            mesh_to_np(mesh._ef, grid).astype(np.float64), _np64(E))


def _errs(got_ef, got_E, ref_ef, ref_E):
    e_mesh = [rel_l2(got_ef[..., c], ref_ef[..., c]) for c in range(3)]
    e_part = [rel_l2(got_E[c], ref_E[c]) for c in range(3)]
    return e_mesh, e_part


def _fmt(v):
    return "[" + ", ".join(f"{e:.2e}" for e in v) + "]"


def _check_fp64_step(label, gpu, ref):
    """Float64 gates: rho 1e-13 of max |rho|; fields 1e-12 of the Float64
    oracle, and 3x closer to the extended-precision result than it."""
    g_rho, g_ef, g_E = gpu
    r_rho, r_ef, r_E, t_ef, t_E = ref
    e_rho = float(np.abs(g_rho - r_rho).max() / np.abs(r_rho).max())
    vs_m, vs_p = _errs(g_ef, g_E, r_ef, r_E)
    gt_m, gt_p = _errs(g_ef, g_E, t_ef, t_E)
    ot_m, ot_p = _errs(r_ef, r_E, t_ef, t_E)
    print(f"{label}: rho {e_rho:.2e}; vs Float64 oracle mesh {_fmt(vs_m)} particles {_fmt(vs_p)}; "
          f"vs extended GPU mesh {_fmt(gt_m)} particles {_fmt(gt_p)}, oracle mesh {_fmt(ot_m)} "
          f"particles {_fmt(ot_p)}")
    assert e_rho <= 1e-13
    return vs_m, vs_p, gt_m, gt_p, ot_m, ot_p


# ---------------------------------------------------------------- config A

A_SEED = 31  # the first seed (from 0) whose 100,000 particles put one outside +-5 sigma


def _config_a_particles(dtype=np.float64):
    x, y, z, q = config_a(100_000, seed=A_SEED, dtype=dtype)
    return x, y, z, q


A_MIN = (-5e-3, -5e-3, -10e-3)
A_MAX = (5e-3, 5e-3, 10e-3)


def test_config_a_seed_exercises_the_clamp():
    x, y, z, _ = _config_a_particles()
    out = (np.abs(x) > 5e-3) | (np.abs(y) > 5e-3) | (np.abs(z) > 10e-3)
    print(f"config A seed {A_SEED}: {int(out.sum())} particles outside +-5 sigma")
    assert out.sum() >= 1


def test_config_a_fp64_full_step():
    sc = _sc()
    grid = (32, 32, 32)
    x, y, z, q = _config_a_particles()
    xt, yt, zt, qt = (torch.as_tensor(a, device=DEV) for a in (x, y, z, q))
    mesh = sc.Mesh3D(grid, A_MIN, A_MAX, T=torch.float64)
    lo, hi, dl = orc.bounds_manual(A_MIN, A_MAX, grid)
    assert (mesh.min_bounds, mesh.max_bounds, mesh.delta) == (lo, hi, dl)
    gpu = _gpu_step(mesh, xt, yt, zt, qt)
    ref = _oracle_step(x, y, z, q, grid, lo, hi, dl, extended=True)
    vs_m, vs_p, gt_m, gt_p, ot_m, ot_p = _check_fp64_step("config A Float64", gpu, ref)
    for c in range(3):
        assert vs_m[c] <= 1e-12 and vs_p[c] <= 1e-12
        assert gt_m[c] <= ot_m[c] / 3 + 1e-15
        assert gt_p[c] <= ot_p[c] / 3 + 1e-15
    # charge conservation including the clamped particles
    assert abs(gpu[0].sum() - q.sum()) <= 1e-12 * np.abs(gpu[0]).sum()


def test_config_a_fp32_full_step():
    sc = _sc()
    grid = (32, 32, 32)
    x, y, z, q = _config_a_particles(np.float32)
    xt, yt, zt, qt = (torch.as_tensor(a, device=DEV) for a in (x, y, z, q))
    mesh = sc.Mesh3D(grid, A_MIN, A_MAX, T=torch.float32)
    lo, hi, dl = orc.bounds_manual(A_MIN, A_MAX, grid, np.float32)
    assert (mesh.min_bounds, mesh.max_bounds, mesh.delta) == (lo, hi, dl)
    _, g_ef, g_E = _gpu_step(mesh, xt, yt, zt, qt)
    x64, y64, z64, q64 = (a.astype(np.float64) for a in (x, y, z, q))
    _, r_ef, r_E = _oracle_step(x64, y64, z64, q64, grid, lo, hi, dl)
    e_m, e_p = _errs(g_ef, g_E, r_ef, r_E)
    print(f"config A Float32 vs Float64 oracle: mesh {_fmt(e_m)} particles {_fmt(e_p)}")
    assert max(e_m) <= 1e-5 and max(e_p) <= 1e-5


# ---------------------------------------------------------------- config B

@pytest.mark.parametrize("gamma", [1.0, 10.0])
def test_config_b_fp32_against_fp64_oracle(gamma):
    sc = _sc()
    bench = _bench()
    grid = (128, 128, 128)
    x, y, z, q = bench.make_particles(4_000_000, torch.float32, torch.device(DEV))
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float32, gamma=gamma, total_charge=1e-9)
    _, g_ef, g_E = _gpu_step(mesh, x, y, z, q)
    x64, y64, z64, q64 = _np64((x, y, z, q))
    want = orc.bounds_auto(*(t.cpu().numpy() for t in (x, y, z)), grid, np.float32)
    assert (mesh.min_bounds, mesh.max_bounds, mesh.delta) == want
    _, r_ef, r_E = _oracle_step(x64, y64, z64, q64, grid, mesh.min_bounds, mesh.max_bounds,
                                mesh.delta, gamma)
    e_m, e_p = _errs(g_ef, g_E, r_ef, r_E)
    print(f"config B Float32 gamma={gamma} vs Float64 oracle: mesh {_fmt(e_m)} "
          f"particles {_fmt(e_p)}")
    assert max(e_m) <= 1e-5 and max(e_p) <= 1e-5


# ---------------------------------------------------------------- config C

def test_config_c_fp32_cathode_against_fp64_oracle():
    sc = _sc()
    bench = _bench()
    grid = (64, 64, 128)
    x, y, z, q = bench.make_cathode_particles(2_000_000, torch.float32, torch.device(DEV))
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float32, total_charge=100e-12)
    # the bunch starts at the cathode: its image sits next to it
    assert mesh.min_bounds[2] < 1e-7
    _, g_ef, g_E = _gpu_step(mesh, x, y, z, q, cathode=True)
    x64, y64, z64, q64 = _np64((x, y, z, q))
    _, r_ef, r_E = _oracle_step(x64, y64, z64, q64, grid, mesh.min_bounds, mesh.max_bounds,
                                mesh.delta, 1.0, True)
    e_m, e_p = _errs(g_ef, g_E, r_ef, r_E)
    print(f"config C Float32 vs Float64 oracle: mesh {_fmt(e_m)} particles {_fmt(e_p)}")
    assert max(e_m) <= 1e-5 and max(e_p) <= 1e-5


def test_config_c_fp64_cathode_against_extended_oracle():
    sc = _sc()
    bench = _bench()
    grid = (64, 64, 128)
    x, y, z, q = bench.make_cathode_particles(2_000_000, torch.float64, torch.device(DEV))
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float64, total_charge=100e-12)
    gpu = _gpu_step(mesh, x, y, z, q, cathode=True)
    x64, y64, z64, q64 = _np64((x, y, z, q))
    lo, hi, dl = orc.bounds_auto(x64, y64, z64, grid)
    assert (mesh.min_bounds, mesh.max_bounds, mesh.delta) == (lo, hi, dl)
    ref = _oracle_step(x64, y64, z64, q64, grid, lo, hi, dl, 1.0, True, extended=True)
    vs_m, vs_p, gt_m, gt_p, ot_m, ot_p = _check_fp64_step("config C Float64", gpu, ref)
    for c in range(3):
        assert gt_m[c] <= ot_m[c] / 3 + 1e-15
        assert gt_p[c] <= ot_p[c] / 3 + 1e-15
        assert vs_m[c] <= max(1e-12, 2 * ot_m[c])
        assert vs_p[c] <= max(1e-12, 2 * ot_p[c])


# ---------------------------------------------------------------- config D

def test_config_d_batched_and_graph_against_fp64_oracle():
    sc = _sc()
    bench = _bench()
    T = torch.float32
    bunches = bench.make_batch(T, torch.device(DEV))
    assert len(bunches) == 32
    assert any(g == (128, 128, 128) for g, _ in bunches)
    parts = [p for _, p in bunches]
    meshes = [sc.Mesh3D(g, p[0], p[1], p[2], T=T) for g, p in bunches]

    def step():
        sc.deposit_batched(meshes, *[[p[k] for p in parts] for k in range(4)])
        sc.solve_batched(meshes)
        return sc.interpolate_batched(meshes, *[[p[k] for p in parts] for k in range(3)])

    E_eager = step()
    torch.cuda.synchronize()
    ef_eager = [m.efield.clone() for m in meshes]
    E_eager = [[t.clone() for t in e] for e in E_eager]
    # the same step captured into a graph and replayed twice, with eager work between
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        E_graph = step()
    for rep in range(2):
        for m in meshes:
            m.efield.zero_()
        graph.replay()
        torch.cuda.synchronize()
        for b, m in enumerate(meshes):
            assert torch.equal(m.efield, ef_eager[b]), f"replay {rep}, efield of bunch {b}"
            for c in range(3):
                assert torch.equal(E_graph[b][c], E_eager[b][c]), f"replay {rep}, E{c} bunch {b}"
    worst_m, worst_p = 0.0, 0.0
    for b, ((grid, p), m) in enumerate(zip(bunches, meshes)):
        x64, y64, z64, q64 = _np64(p)
        _, r_ef, r_E = _oracle_step(x64, y64, z64, q64, grid, m.min_bounds, m.max_bounds,
                                    m.delta)
        g_ef = mesh_to_np(ef_eager[b], grid).astype(np.float64)
        e_m, e_p = _errs(g_ef, _np64(E_eager[b]), r_ef, r_E)
        worst_m, worst_p = max(worst_m, max(e_m)), max(worst_p, max(e_p))
        assert max(e_m) <= 1e-5 and max(e_p) <= 1e-5, (b, grid, p[0].numel(), e_m, e_p)
    print(f"config D, 32 bunches vs Float64 oracle: worst mesh {worst_m:.2e}, "
          f"worst particles {worst_p:.2e}")


# ---------------------------------------------------------------- config E

def test_config_e_tracking_directional_derivative():
    sc = _sc()
    import spacecharge.autograd as scad
    n, sig = 1_000_000, 1e-3
    gen = torch.Generator(device=DEV).manual_seed(17)
    p = torch.randn((n, 3), generator=gen, dtype=torch.float64, device=DEV) * sig
    x, y, z = (p[:, i].contiguous() for i in range(3))
    q = torch.full((n,), 1e-9 / n, dtype=torch.float64, device=DEV)
    q = q * (1.0 + 0.5 * torch.rand(n, generator=gen, dtype=torch.float64, device=DEV))
    mesh = sc.Mesh3D((64, 64, 64), (-8 * sig,) * 3, (8 * sig,) * 3, T=torch.float64)
    nsteps, every, dt, kick = 100, 10, 1.0, 1e-3 * sig / 1e7

    def loss(x0, y0, z0, q0, ckpt):
        zero = torch.zeros_like(x0)
        st = scad.track(mesh, x0, y0, z0, zero, zero, zero, q0, nsteps, dt, kick,
                        checkpoint_every=ckpt)
        return scad.rms_emittance(*st)

    ins = [t.clone().requires_grad_(True) for t in (x, y, z, q)]
    L = loss(*ins, every)
    L.backward()
    v = [torch.randn(n, generator=gen, dtype=torch.float64, device=DEV) for _ in range(4)]
    scale = [sig, sig, sig, q]
    # d/de L(x + e sig vx, y + e sig vy, z + e sig vz, q + e q vq) at e = 0
    dirder = sum(float((ins[a].grad * scale[a] * v[a]).sum()) for a in range(4))
    h = 1e-6
    with torch.no_grad():
        up = [t + h * s * w for t, s, w in zip((x, y, z, q), scale, v)]
        dn = [t - h * s * w for t, s, w in zip((x, y, z, q), scale, v)]
        fd = (float(loss(*up, 0)) - float(loss(*dn, 0))) / (2 * h)
    err = abs(fd - dirder) / abs(fd)
    print(f"config E: loss {float(L):.6e}, directional derivative {dirder:.10e}, "
          f"central difference {fd:.10e}, rel err {err:.2e}")
    assert np.isfinite(dirder) and err <= 1e-5


# ---------------------------------------------------------------- device bounds

@pytest.mark.parametrize("pt,mt", [(torch.float64, torch.float64), (torch.float32, torch.float32),
                                   (torch.float64, torch.float32), (torch.float32, torch.float64)])
def test_device_bounds_equal_oracle_bitwise(pt, mt):
    sc = _sc()
    npt = np.float64 if pt == torch.float64 else np.float32
    mnp = np.float64 if mt == torch.float64 else np.float32
    cases = []
    x, y, z = gaussian(60_000, (1e-3, 1.5e-3, 3e-3), 51)
    cases.append((x, y, z))
    cases.append((x + 0.3, y - 1.7, z + 12.5))                  # far from the origin
    cases.append((x[:1], y[:1], z[:1]))                         # one particle: zero patch
    cases.append((x, np.full_like(y, 2.5e-3), z))               # flat in y: zero patch
    for grid in ((32, 32, 32), (17, 9, 40), (2, 3, 1000)):
        for cx, cy, cz in cases:
            cx, cy, cz = (np.ascontiguousarray(a, dtype=npt) for a in (cx, cy, cz))
            mesh = sc.Mesh3D(grid, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), T=mt)
            mesh.update_bounds(*(torch.as_tensor(a, device=DEV) for a in (cx, cy, cz)))
            got = mesh.fetch_bounds()
            want = orc.bounds_auto(cx, cy, cz, grid, mnp)
            assert got == want, (grid, cx.size, got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_extrema_of_a_device_array(dtype):
    """sc_extrema (the Julia binding's extrema of a device array, which the
    reference's Mesh3D constructor calls on the particles, src/mesh.jl:120-122)
    equals numpy's min and max, also for one element and for negative data."""
    sc = _sc()
    from spacecharge import _lib
    import ctypes
    lib = _lib.load()
    rng = np.random.default_rng(52)
    for a in (rng.standard_normal(1_000_003) * 3e-3 - 7.0, np.array([2.5e-3]),
              -np.abs(rng.standard_normal(4097))):
        a = a.astype(dtype)
        t = torch.as_tensor(a, device=DEV)
        mn, mx = ctypes.c_double(), ctypes.c_double()
        ctx = sc._ctx(t.device)
        rc = lib.sc_extrema(ctx, _lib.SC_F32 if dtype == np.float32 else _lib.SC_F64,
                            t.data_ptr(), t.numel(), ctypes.byref(mn), ctypes.byref(mx))
        assert rc == 0, lib.sc_last_error(ctx).decode()
        assert mn.value == float(a.min()) and mx.value == float(a.max())
    assert lib.sc_extrema(ctx, _lib.SC_F64, t.data_ptr(), 0, ctypes.byref(mn),
                          ctypes.byref(mx)) != 0
