"""GPU: the scalar potential solve (sc_solve_potential) against the oracle.

The reference defines the potential's Green function
(potential_green_function, src/green_functions.jl:13-22) but no solver that
uses it; the oracle's solve_potential applies it exactly as the reference's
field solve applies the field's (src/solvers/free_space.jl:56-101), and is
pinned on the CPU (tests/test_oracle.py).

Tolerances:
  * Float64 mesh: the same yardstick as the field solve -- the algorithm in
    x87 extended precision (oracle.solve_potential_extended).  The GPU must
    be within 1e-12 relative L2 of it and at least as close as the Float64
    oracle (plus 1e-14).
  * Float32 mesh: relative L2 <= 1e-5 against the extended result.  (The
    oracle's own Float32 restatement differences the antiderivative in
    Float32 and is ~1e-3 off at 64^3; the GPU differences it in Float64.)
  * Properties without the oracle: -grad(phi) against the GPU field solve
    (x, y and z / gamma^2), the closed-form potential of a spherical
    Gaussian, linearity in rho.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.special import erf

import oracle as orc
from _common import config_a, gaussian, mesh_to_np, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _sc():
    import spacecharge
    return spacecharge


def _mesh(grid, T, gamma, n=20_000, seed=4, sigma=(1e-3, 1e-3, 2e-3), generic=False):
    sc = _sc()
    x, y, z = gaussian(n, sigma, seed)
    q = np.full(n, 1e-9 / n)
    xt, yt, zt, qt = (torch.as_tensor(a, dtype=torch.float64, device=DEV) for a in (x, y, z, q))
    mesh = sc.Mesh3D(grid, xt, yt, zt, T=T, gamma=gamma)
    if generic:
        mesh.force_generic_transforms()
    sc.deposit(mesh, xt, yt, zt, qt)
    return mesh


CASES = [((4, 4, 4), 1.0), ((8, 8, 8), 2.0), ((16, 8, 12), 1.0), ((10, 14, 20), 3.0),
         ((32, 32, 32), 1.0), ((24, 20, 40), 10.0), ((64, 32, 128), 2.0), ((64, 64, 128), 1.0),
         # 2n with a prime factor above 7: transforms padded to a 7-smooth length
         ((11, 13, 17), 1.0), ((33, 31, 20), 2.0)]


@pytest.mark.parametrize("generic", [False, True])
@pytest.mark.parametrize("grid,gamma", CASES)
def test_potential_fp64_matches_extended(grid, gamma, generic):
    sc = _sc()
    mesh = _mesh(grid, torch.float64, gamma, generic=generic)
    phi = sc.solve_potential(mesh)
    torch.cuda.synchronize()
    assert phi.shape == tuple(grid)
    got = phi.cpu().numpy()
    rho = np.asfortranarray(mesh_to_np(mesh._rho, grid))
    truth = orc.solve_potential_extended(rho, mesh.delta, mesh.gamma)
    ref = orc.solve_potential(rho, mesh.delta, mesh.gamma)
    e_gpu, e_orc = rel_l2(got, truth), rel_l2(ref, truth)
    print(f"potential f64 {grid} gamma={gamma} generic={generic}: vs extended GPU {e_gpu:.2e} "
          f"oracle {e_orc:.2e}")
    assert e_gpu <= 1e-12
    assert e_gpu <= e_orc + 1e-14


@pytest.mark.parametrize("generic", [False, True])
@pytest.mark.parametrize("grid,gamma", CASES)
def test_potential_fp32_matches_extended(grid, gamma, generic):
    sc = _sc()
    mesh = _mesh(grid, torch.float32, gamma, generic=generic)
    phi = sc.solve_potential(mesh)
    torch.cuda.synchronize()
    assert phi.dtype == torch.float32
    got = phi.cpu().numpy()
    rho = np.asfortranarray(mesh_to_np(mesh._rho, grid).astype(np.float64))
    truth = orc.solve_potential_extended(rho, mesh.delta, mesh.gamma)
    err = rel_l2(got, truth)
    print(f"potential f32 {grid} gamma={gamma} generic={generic}: vs extended {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("grid,T", [((6, 6, 1500), torch.float32), ((6, 2048, 5), torch.float32),
                                    ((6, 700, 6), torch.float64), ((1024, 5, 6), torch.float64),
                                    ((6, 6, 3000), torch.float32), ((6, 2048, 5), torch.float64)])
def test_potential_long_axis(grid, T):
    """Transforms beyond 2048 Float32 / 1024 Float64 points (twice the
    engine's values per lane; 512-lane workgroups beyond 4096 / 2048),
    against the extended-precision potential."""
    sc = _sc()
    mesh = _mesh(grid, T, 1.0)
    phi = sc.solve_potential(mesh)
    torch.cuda.synchronize()
    got = phi.cpu().numpy().astype(np.float64)
    rho = np.asfortranarray(mesh_to_np(mesh._rho, grid).astype(np.float64))
    truth = orc.solve_potential_extended(rho, mesh.delta, mesh.gamma)
    err = rel_l2(got, truth)
    print(f"potential long axis {T} {grid}: vs extended {err:.2e}")
    assert err <= (1e-12 if T == torch.float64 else 1e-5)


@pytest.mark.parametrize("gamma", [1.0, 2.5])
def test_potential_gradient_is_field(gamma):
    """-dphi/dx, -dphi/dy = Ex, Ey and -dphi/dz = gamma^2 Ez (lab-frame
    potential of a bunch moving along z), by central differences in the
    interior, against the GPU field solve of the same rho."""
    sc = _sc()
    grid = (64, 64, 64)
    mesh = _mesh(grid, torch.float64, gamma, n=400_000, sigma=(1e-3, 1e-3, 1e-3))
    phi = sc.solve_potential(mesh).cpu().numpy()
    sc.solve(mesh)
    torch.cuda.synchronize()
    ef = mesh_to_np(mesh._ef, grid)
    s = slice(8, -8)
    for ax, fac in ((0, 1.0), (1, 1.0), (2, 1.0 / gamma ** 2)):
        g = -fac * (np.roll(phi, -1, ax) - np.roll(phi, 1, ax)) / (2 * mesh.delta[ax])
        err = rel_l2(g[s, s, s], ef[s, s, s, ax])
        print(f"gamma={gamma} axis {ax}: -grad phi vs E rel L2 {err:.2e}")
        assert err < 5e-3


def test_potential_analytical_isotropic_gaussian():
    """phi(r) = Q erf(r / (sqrt(2) sigma)) / (4 pi eps0 r): 2M particles,
    sigma = 1 mm, 1 nC, 64^3 auto-bounds, along the z axis."""
    sc = _sc()
    eps0 = 8.8541878128e-12
    n, Q, sigma = 2_000_000, 1e-9, 1e-3
    x, y, z = gaussian(n, (sigma, sigma, sigma), seed=7)
    q = np.full(n, Q / n)
    xt, yt, zt, qt = (torch.as_tensor(a, device=DEV) for a in (x, y, z, q))
    mesh = sc.Mesh3D((64, 64, 64), xt, yt, zt, T=torch.float64)
    sc.deposit(mesh, xt, yt, zt, qt)
    phi = sc.solve_potential(mesh).cpu().numpy()
    lo, dl = mesh.min_bounds, mesh.delta
    ix = int(np.argmin(np.abs([lo[0] + i * dl[0] for i in range(64)])))
    iy = int(np.argmin(np.abs([lo[1] + j * dl[1] for j in range(64)])))
    zc = np.array([lo[2] + k * dl[2] for k in range(64)])
    r = np.sqrt((lo[0] + ix * dl[0]) ** 2 + (lo[1] + iy * dl[1]) ** 2 + zc ** 2)
    ana = Q / (4 * np.pi * eps0) * erf(r / (np.sqrt(2) * sigma)) / r
    err = np.abs(phi[ix, iy, :] - ana).max() / np.abs(ana).max()
    print(f"Gaussian potential: max rel err on axis {err:.2e}")
    assert err < 0.01


def test_potential_linear_in_rho():
    """Two charge sets deposited on one mesh (same spacing): phi(2 rho1 - 3 rho2)
    = 2 phi(rho1) - 3 phi(rho2)."""
    sc = _sc()
    grid = (32, 32, 64)
    m = _mesh(grid, torch.float64, 1.5, seed=11)
    rho1 = m.rho.clone()
    p1 = sc.solve_potential(m).clone()
    x, y, z = gaussian(30_000, (0.8e-3, 1.2e-3, 1.5e-3), 12)
    xt, yt, zt = (torch.as_tensor(a, device=DEV) for a in (x, y, z))
    sc.deposit(m, xt, yt, zt, torch.full_like(xt, 2e-9 / 30_000))
    rho2 = m.rho.clone()
    p2 = sc.solve_potential(m).clone()
    m.rho.copy_(2.0 * rho1 - 3.0 * rho2)
    p3 = sc.solve_potential(m)
    torch.cuda.synchronize()
    lhs = p3.cpu().numpy()
    rhs = 2.0 * p1.cpu().numpy() - 3.0 * p2.cpu().numpy()
    assert rel_l2(lhs, rhs) < 1e-13


@pytest.mark.parametrize("reuse", [False, True])
def test_potential_and_field_solves_share_the_workspace(reuse):
    """Field, potential, field on one mesh: each result equals a solve on a
    fresh mesh bit for bit (the cached Green spectra are rebuilt when the
    solve kind changes, also with the opt-in reuse)."""
    sc = _sc()
    grid = (32, 32, 64)
    m = _mesh(grid, torch.float32, 2.0, seed=13)
    if reuse:
        m.set_green_reuse(True)
    fresh = _mesh(grid, torch.float32, 2.0, seed=13)
    sc.solve(fresh)
    ef_ref = fresh.efield.clone()
    fresh2 = _mesh(grid, torch.float32, 2.0, seed=13)
    phi_ref = sc.solve_potential(fresh2).clone()
    for _ in range(2):
        sc.solve(m)
        torch.cuda.synchronize()
        assert torch.equal(m.efield, ef_ref)
        phi = sc.solve_potential(m)
        torch.cuda.synchronize()
        assert torch.equal(phi, phi_ref)
    sc.solve(m)
    torch.cuda.synchronize()
    assert torch.equal(m.efield, ef_ref)


def test_potential_out_argument_and_errors():
    sc = _sc()
    grid = (8, 8, 8)
    m = _mesh(grid, torch.float64, 1.0, n=1000)
    out = torch.empty(512, dtype=torch.float64, device=DEV)
    phi = sc.solve_potential(m, out=out)
    assert phi.data_ptr() == out.data_ptr()
    with pytest.raises(sc.SpaceChargeError, match="out must be"):
        sc.solve_potential(m, out=torch.empty(511, dtype=torch.float64, device=DEV))
    with pytest.raises(sc.SpaceChargeError, match="out must be"):
        sc.solve_potential(m, out=torch.empty(512, dtype=torch.float32, device=DEV))
    from spacecharge import _lib
    lib = _lib.load()
    assert lib.sc_solve_potential(m._h, _lib.d3(m.delta), 1.0, None) != 0
    assert lib.sc_solve_potential(m._h, _lib.d3(m.delta), -1.0, out.data_ptr()) != 0
    assert lib.sc_solve_potential(m._h, _lib.d3((0.0, 1.0, 1.0)), 1.0, out.data_ptr()) != 0
    assert lib.sc_solve_potential(None, _lib.d3(m.delta), 1.0, out.data_ptr()) != 0
