"""The reference's own test suite (test/*.jl), ported to the MI355X API.

Each test names the Julia testset it restates.  Where the reference compares
its GPU path with its CPU path, the CPU side here is the oracle (the C
restatement of the reference CPU path).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.special import erf

import oracle as orc
from _common import gaussian, mesh_to_np

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def sc():
    import spacecharge
    return spacecharge


def t(a, dtype=torch.float64):
    return torch.as_tensor(np.asarray(a, dtype=np.float64), dtype=dtype, device=DEV)


def rho_np(mesh):
    return mesh_to_np(mesh._rho, mesh.grid_size)


def ef_np(mesh):
    return mesh_to_np(mesh._ef, mesh.grid_size)


# ------------------------------------------------ test/test_mesh.jl
def test_mesh_particle_based_constructor():
    S = sc()
    g = (10, 20, 30)
    x, y, z = t([-1.0, 1.0]), t([-2.0, 2.0]), t([-3.0, 3.0])
    m = S.Mesh3D(g, x, y, z)
    assert m.grid_size == g and m.rho.dtype == torch.float64
    assert tuple(m.rho.shape) == g and tuple(m.efield.shape) == g + (3,)
    for a, (mn, mx) in enumerate(((-1, 1), (-2, 2), (-3, 3))):
        assert m.min_bounds[a] < mn and m.max_bounds[a] > mx
        assert m.delta[a] == pytest.approx((m.max_bounds[a] - m.min_bounds[a]) / (g[a] - 1))


def test_mesh_manual_bounds_constructor():
    S = sc()
    g = (10, 20, 30)
    m = S.Mesh3D(g, (-1.0, -2.0, -3.0), (1.0, 2.0, 3.0))
    assert m.min_bounds == (-1.0, -2.0, -3.0) and m.max_bounds == (1.0, 2.0, 3.0)
    for a in range(3):
        assert m.delta[a] == pytest.approx((m.max_bounds[a] - m.min_bounds[a]) / (g[a] - 1))
    assert tuple(m.rho.shape) == g and tuple(m.efield.shape) == g + (3,)


def test_mesh_type_parameters_and_physics():
    S = sc()
    m = S.Mesh3D((5, 5, 5), t([0.0]), t([0.0]), t([0.0]), T=torch.float32)
    assert m.rho.dtype == torch.float32 and m.efield.dtype == torch.float32 and m.rho.is_cuda
    m = S.Mesh3D((5, 5, 5), t([0.0]), t([0.0]), t([0.0]), gamma=2.0, total_charge=1.0)
    assert m.gamma == 2.0 and m.total_charge == 1.0


# ------------------------------------------------ test/test_deposition.jl
def test_deposition_testsets():
    S = sc()
    # Basic Deposition
    m = S.Mesh3D((4, 4, 4), t([0.5]), t([0.5]), t([0.5]))
    S.clear_mesh(m)
    S.deposit(m, t([0.5]), t([0.5]), t([0.5]), t([1.0]))
    r = rho_np(m)
    assert abs(r.sum() - 1.0) <= 1e-10 and (r > 0).any()
    # Multiple Particles
    x, y, z, q = t([0.2, 0.8]), t([0.3, 0.7]), t([0.4, 0.6]), t([1.0, -1.0])
    m = S.Mesh3D((6, 6, 6), x, y, z)
    S.deposit(m, x, y, z, q)
    r = rho_np(m)
    assert abs(r.sum()) <= 1e-10 and (r > 0).any() and (r < 0).any()
    # Charge Accumulation
    m = S.Mesh3D((4, 4, 4), t([0.5]), t([0.5]), t([0.5]))
    S.deposit(m, t([0.5]), t([0.5]), t([0.5]), t([1.0]))
    first = rho_np(m).sum()
    S.deposit(m, t([0.5]), t([0.5]), t([0.5]), t([1.0]), clear=False)
    assert abs(rho_np(m).sum() - 2 * first) <= 1e-10
    # Edge Cases: particle at a grid point, particle at the boundary
    m = S.Mesh3D((4, 4, 4), t([0.0]), t([0.0]), t([0.0]))
    S.deposit(m, t([0.0]), t([0.0]), t([0.0]), t([1.0]))
    assert abs(rho_np(m).sum() - 1.0) <= 1e-10
    mx = m.max_bounds
    m2 = S.Mesh3D((4, 4, 4), t([mx[0]]), t([mx[1]]), t([mx[2]]))
    S.deposit(m2, t([mx[0]]), t([mx[1]]), t([mx[2]]), t([1.0]))
    assert abs(rho_np(m2).sum() - 1.0) <= 1e-10
    # Clear Mesh
    S.clear_mesh(m2)
    assert (rho_np(m2) == 0).all()


def test_deposition_length_mismatch_raises():
    S = sc()
    m = S.Mesh3D((4, 4, 4), t([0.5]), t([0.5]), t([0.5]))
    with pytest.raises(S.SpaceChargeError,
                       match="Particle coordinate and charge arrays must have the same length."):
        S.deposit(m, t([0.5, 0.6]), t([0.5]), t([0.5]), t([1.0]))


# ------------------------------------------------ test/test_interpolation.jl
def test_interpolation_testsets():
    S = sc()
    # Basic Interpolation
    m = S.Mesh3D((4, 4, 4), t([0.5]), t([0.5]), t([0.5]))
    m.efield[0, 0, 0, 0] = 1.0
    m.efield[1, 0, 0, 0] = 2.0
    m.efield[0, 1, 0, 0] = 3.0
    m.efield[0, 0, 1, 0] = 4.0
    Ex, Ey, Ez = S.interpolate_field(m, t([0.5]), t([0.5]), t([0.5]))
    assert np.isfinite(Ex.item()) and abs(Ey.item()) <= 1e-10 and abs(Ez.item()) <= 1e-10
    # Multiple Particles, linear gradient
    x = t([0.25, 0.75])
    m = S.Mesh3D((4, 4, 4), x, x, x)
    ii, jj, kk = torch.meshgrid(*[torch.arange(1, 5, dtype=torch.float64, device=DEV)] * 3,
                                indexing="ij")
    m.efield[..., 0] = ii * 0.1
    m.efield[..., 1] = jj * 0.1
    m.efield[..., 2] = kk * 0.1
    Ex, Ey, Ez = S.interpolate_field(m, x, x, x)
    assert len(Ex) == 2 and torch.isfinite(torch.stack([Ex, Ey, Ez])).all()
    # Interpolation at Grid Points
    x = t([0.0, 1.0])
    m = S.Mesh3D((3, 3, 3), x, x, x)
    m.efield[0, 0, 0, 0] = 1.0
    m.efield[1, 0, 0, 0] = 2.0
    Ex, _, _ = S.interpolate_field(m, x, x, x)
    assert torch.isfinite(Ex).all()
    # Interpolation with Computed Fields.  The reference asserts |E| > 1e7 at
    # the charge itself (test/test_interpolation.jl:109-111); by symmetry that
    # self-field is zero, and the reference meets the bound only through
    # round-off (~1e8 V/m against 8.5e21 V/m one cell away, see
    # test_oracle.py::test_reference_self_field_is_roundoff).  The
    # octant-symmetric Green function here gives ~0 there, so the port checks
    # the field one cell away from the charge instead: large and pointing away.
    m = S.Mesh3D((8, 8, 8), t([0.0]), t([0.0]), t([0.0]))
    S.deposit(m, t([0.0]), t([0.0]), t([0.0]), t([1.0]))
    S.solve(m)
    Ex, Ey, Ez = S.interpolate_field(m, t([0.0]), t([0.0]), t([0.0]))
    for e in (Ex, Ey, Ez):
        assert torch.isfinite(e).all() and abs(e.item()) < 1e-9 * 8.5e21
    d = m.delta
    for c in range(3):
        p = [0.0, 0.0, 0.0]
        p[c] = d[c]
        E = S.interpolate_field(m, t([p[0]]), t([p[1]]), t([p[2]]))
        assert E[c].item() > 1e7
    # Interpolation Edge Cases: constant field
    x = t([0.0, 1.0])
    m = S.Mesh3D((4, 4, 4), x, x, x)
    m.efield.fill_(1.0)
    for e in S.interpolate_field(m, x, x, x):
        assert torch.allclose(e, torch.ones_like(e))


# ------------------------------------------------ test/test_solvers.jl
def test_solver_testsets():
    S = sc()
    o = t([0.0])
    # FreeSpace Solver
    m = S.Mesh3D((16, 16, 16), o, o, o, gamma=2.0)
    S.deposit(m, o, o, o, t([1.0]))
    S.solve(m)
    assert abs(rho_np(m).sum() - 1.0) <= 1e-10
    e = ef_np(m)
    assert (e != 0).any() and np.isfinite(e).all()
    # FreeSpace Solver with Cathode
    zc = t([0.01])
    m = S.Mesh3D((16, 16, 16), o, o, zc, gamma=2.0)
    S.deposit(m, o, o, zc, t([1.0]))
    S.solve(m, at_cathode=True)
    mf = S.Mesh3D((16, 16, 16), o, o, zc, gamma=2.0)
    S.deposit(mf, o, o, zc, t([1.0]))
    S.solve(mf)
    ec, ef = ef_np(m), ef_np(mf)
    assert not np.allclose(ec, ef) and (ec != 0).any() and np.isfinite(ec).all()
    # Solver with Different Gamma
    m1 = S.Mesh3D((8, 8, 8), o, o, o, gamma=1.0)
    S.deposit(m1, o, o, o, t([1.0]))
    S.solve(m1)
    m2 = S.Mesh3D((8, 8, 8), o, o, o, gamma=10.0)
    S.deposit(m2, o, o, o, t([1.0]))
    S.solve(m2)
    assert not np.allclose(ef_np(m1), ef_np(m2))


# ------------------------------------------------ test/test_gpu.jl
def test_gpu_testsets():
    S = sc()
    f32 = torch.float32
    # GPU Mesh Creation
    m = S.Mesh3D((10, 10, 10), t([0.5]), t([0.5]), t([0.5]), T=f32)
    assert m.rho.dtype == f32 and m.rho.is_cuda and m.efield.dim() == 4
    # GPU Deposition
    p = t([0.5], f32)
    m = S.Mesh3D((8, 8, 8), p, p, p, T=f32)
    S.deposit(m, p, p, p, t([1.0], f32))
    r = rho_np(m)
    assert abs(r.sum() - 1.0) <= 1e-6 and (r > 0).any()
    # GPU Interpolation
    m = S.Mesh3D((6, 6, 6), p, p, p, T=f32)
    m.efield[0, 0, 0, 0] = 1.0
    m.efield[1, 0, 0, 0] = 2.0
    Ex, Ey, Ez = S.interpolate_field(m, p, p, p)
    assert Ex.dtype == f32 and Ex.is_cuda and torch.isfinite(torch.stack([Ex, Ey, Ez])).all()
    # GPU Solver
    p0 = t([0.0], f32)
    m = S.Mesh3D((8, 8, 8), p0, p0, p0, T=f32, gamma=2.0)
    S.deposit(m, p0, p0, p0, t([1.0], f32))
    S.solve(m)
    assert (ef_np(m) != 0).any() and abs(rho_np(m).sum() - 1.0) <= 1e-6
    # GPU Multiple Particles: 6^3 -> 12^3 padded transform (radix 3)
    x, y, z, q = t([0.2, 0.8], f32), t([0.3, 0.7], f32), t([0.4, 0.6], f32), t([1.0, -1.0], f32)
    m = S.Mesh3D((6, 6, 6), x, y, z, T=f32)
    S.deposit(m, x, y, z, q)
    S.solve(m)
    Ex, Ey, Ez = S.interpolate_field(m, x, y, z)
    assert len(Ex) == 2 and torch.isfinite(torch.stack([Ex, Ey, Ez])).all()
    assert abs(rho_np(m).sum()) <= 1e-6
    # Default Backend Inference
    pp = t([0.5, 1.0], f32)
    m = S.Mesh3D((6, 6, 6), pp, pp, pp, T=f32)
    assert m.rho.is_cuda and tuple(m.rho.shape) == (6, 6, 6) and tuple(m.efield.shape) == (6, 6, 6, 3)
    # Manual Bounds Constructor GPU
    m = S.Mesh3D((8, 8, 8), (-1.0, -2.0, -3.0), (1.0, 2.0, 3.0), T=f32)
    assert m.min_bounds == (-1.0, -2.0, -3.0) and m.max_bounds == (1.0, 2.0, 3.0)
    # GPU Clear Mesh
    m = S.Mesh3D((4, 4, 4), p, p, p, T=f32)
    S.deposit(m, p, p, p, t([1.0], f32))
    assert abs(rho_np(m).sum() - 1.0) <= 1e-6
    S.clear_mesh(m)
    assert (rho_np(m) == 0).all()


def test_gpu_vs_cpu_consistency():
    """test/test_gpu.jl:158-183: sum(rho) atol 1e-5 and sum(efield) rtol 1e-5
    between the Float32 GPU mesh and the CPU path.  The reference's Float32 CPU
    path builds its Green function in Float32 and its field sum is itself
    1.03e-5 away from the Float64 / extended-precision value here, so the CPU
    side of the comparison is the Float64 path; the Float32 CPU value is
    checked against it separately."""
    S = sc()
    x = np.array([0.5])
    m = S.Mesh3D((4, 4, 4), t(x, torch.float32), t(x, torch.float32), t(x, torch.float32),
                 T=torch.float32)
    S.deposit(m, t(x, torch.float32), t(x, torch.float32), t(x, torch.float32),
              t([1.0], torch.float32))
    S.solve(m)
    lo, hi, dl = orc.bounds_auto(x, x, x, (4, 4, 4), np.float32)
    rho32 = orc.deposit(x, x, x, np.array([1.0]), (4, 4, 4), lo, dl, np.float32)
    rho64 = orc.deposit(x, x, x, np.array([1.0]), (4, 4, 4), lo, dl, np.float64)
    ef32 = orc.solve(rho32, lo, hi, dl).astype(np.float64)
    ef64 = orc.solve(rho64, lo, hi, dl)
    gpu = ef_np(m).astype(np.float64).sum()
    print(f"sum(efield): GPU f32 {gpu:.10e}, CPU f64 {ef64.sum():.10e}, CPU f32 {ef32.sum():.10e}")
    assert rho_np(m).sum() == pytest.approx(rho32.sum(), abs=1e-5)
    assert gpu == pytest.approx(ef64.sum(), rel=1e-5)
    assert ef32.sum() == pytest.approx(ef64.sum(), rel=2e-5)


# ------------------------------------------------ test/analytical_test.jl
@pytest.mark.parametrize("T", [torch.float64, torch.float32])
def test_analytical_isotropic_gaussian(T):
    S = sc()
    eps0 = 8.8541878128e-12
    n, Q, sigma = 1_000_000, 1e-9, 1e-3
    x, y, z = gaussian(n, (sigma, sigma, sigma), seed=123)
    xt, yt, zt = t(x, T), t(y, T), t(z, T)
    qt = torch.full((n,), Q / n, dtype=T, device=DEV)
    m = S.Mesh3D((32, 32, 32), xt, yt, zt, T=T, total_charge=Q)
    S.deposit(m, xt, yt, zt, qt)
    S.solve(m)
    lo, dl = m.min_bounds, m.delta
    zc = np.array([lo[2] + k * dl[2] for k in range(32)])
    ix = int(np.argmin(np.abs([lo[0] + i * dl[0] for i in range(32)])))
    iy = int(np.argmin(np.abs([lo[1] + j * dl[1] for j in range(32)])))
    r = np.abs(zc)
    ana = Q / (4 * np.pi * eps0 * r ** 3) * (erf(r / (np.sqrt(2) * sigma)) - np.sqrt(2 / np.pi) *
                                             r / sigma * np.exp(-(r / sigma) ** 2 / 2)) * zc
    got = ef_np(m)[ix, iy, :, 2].astype(np.float64)
    err = np.abs(got - ana).max() / np.abs(ana).max()
    print(f"analytical Gaussian {T}: max error {err:.4f}")
    assert err < 0.10


@pytest.mark.parametrize("T", [torch.float32, torch.float64])
@pytest.mark.parametrize("gamma,grid", [(1.0, (128, 128, 128)), (10.0, (128, 128, 512))])
def test_analytical_anisotropic_gaussian_128(T, gamma, grid):
    """The Gaussian half of the single-bunch configuration: 2M particles,
    sigma = (1, 1, 3) mm, 1 nC, auto-bounds.  The field on the three axes
    against the 1-D integral form, within 1%.  At gamma the lab field is the
    rest frame's (sigma_z stretched by gamma) with the transverse components
    multiplied by gamma.  Quasi-random (scrambled Sobol) sampling keeps the
    shot noise below the discretisation error.  At gamma = 10 the mesh is
    refined in z: the on-axis Ez of the rest frame's long, thin bunch needs
    rest-frame z cells no longer than sigma_x (with 128 z cells they are 2.4
    sigma_x and the reference algorithm is 6% off, converging as dz^2)."""
    import spacecharge as sc
    from scipy.stats import norm, qmc
    from _common import gaussian_field
    Q, sig = 1e-9, (1e-3, 1e-3, 3e-3)
    n = 2_000_000
    u = np.clip(qmc.Sobol(d=3, scramble=True, seed=3).random(n), 1e-12, 1 - 1e-12)
    g = norm.ppf(u)
    x, y, z = (g[:, a] * sig[a] for a in range(3))
    xt, yt, zt = (torch.as_tensor(a, dtype=T, device=DEV) for a in (x, y, z))
    qt = torch.full_like(xt, Q / n)
    mesh = sc.Mesh3D(grid, xt, yt, zt, T=T, gamma=gamma)
    sc.deposit(mesh, xt, yt, zt, qt)
    sc.solve(mesh)
    ef = mesh_to_np(mesh._ef, grid).astype(np.float64)
    lo, dl = mesh.min_bounds, mesh.delta
    ix = [int(np.argmin(np.abs(lo[a] + np.arange(grid[a]) * dl[a]))) for a in range(3)]
    rest_sig = (sig[0], sig[1], gamma * sig[2])
    worst = 0.0
    for c in range(3):
        idx = list(ix)
        pts = []
        for k in range(grid[c]):
            idx[c] = k
            pts.append([lo[a] + idx[a] * dl[a] for a in range(3)])
        pts = np.array(pts)
        rest = pts * np.array([1.0, 1.0, gamma])
        ana = gaussian_field(rest, rest_sig, Q)[:, c] * (gamma if c < 2 else 1.0)
        sl = list(ix)
        sl[c] = slice(None)
        got = ef[tuple(sl) + (c,)]
        err = np.abs(got - ana).max() / np.abs(ana).max()
        worst = max(worst, err)
        print(f"anisotropic Gaussian {grid} {T} gamma={gamma} axis {c}: max err {err:.4f}")
    assert worst < 0.01
