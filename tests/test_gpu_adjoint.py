"""GPU: the vector-Jacobian products (sc_deposit_adjoint, sc_solve_adjoint,
sc_interpolate_adjoint) and the torch.autograd layer.

Yardsticks (the reference has no derivatives, so the oracle's adjoints are
derived from its forward functions and pinned on the CPU in test_oracle.py):
  * solve: against the oracle's transpose in extended precision, for the
    bunch's own field as cotangent (Float64: within 1e-12 and at least as
    close as the Float64 oracle; Float32: 1e-5) and for white noise (Float64:
    at least as close as the Float64 oracle; Float32: 1e-4), and the
    dot-product identity <S rho, g> = <rho, S^T g> with the GPU's own forward
    solve.
  * deposit / interpolate: against the oracle's VJPs (Float64 to 1e-12 of
    the largest value; Float32 to 1e-5).
  * the whole step: central differences of the GPU forward step (Float64,
    steps of 1e-4 cell, exact up to round-off because the loss is quadratic
    in one particle's position inside its cell) for single particles,
    through deposit, solve and interpolate with the mesh held fixed;
    torch.autograd through the wrappers equals the explicit chain.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle as orc
from _common import gaussian, mesh_to_np, rel_l2

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _sc():
    import spacecharge
    return spacecharge


def _bunch(n, seed, dtype=torch.float64, sigma=(1e-3, 1e-3, 2e-3), zshift=0.0):
    x, y, z = gaussian(n, sigma, seed)
    z = z + zshift
    q = np.full(n, 1e-9 / n)
    return [torch.as_tensor(a, dtype=dtype, device=DEV) for a in (x, y, z, q)]


SOLVE_CASES = [((8, 8, 8), 1.0, False), ((16, 8, 12), 2.0, False), ((24, 20, 40), 10.0, False),
               ((32, 32, 64), 1.0, False), ((8, 8, 8), 1.0, True), ((16, 16, 32), 2.0, True),
               ((12, 10, 6), 1.0, True), ((32, 64, 32), 1.5, True),
               # 2n with a prime factor above 7 (transforms padded to a 7-smooth length)
               ((11, 13, 17), 1.0, False), ((13, 11, 19), 2.0, True)]


@pytest.mark.parametrize("generic", [False, True])
@pytest.mark.parametrize("grid,gamma,cath", SOLVE_CASES)
def test_solve_adjoint_matches_oracle(grid, gamma, cath, generic):
    """Two cotangents: the bunch's own field (the cotangent of the field
    energy, smooth) and white noise.  White noise makes S^T g a near-
    cancelling sum, where the reference algorithm in Float64 is itself only
    good to ~1e-11 relative; there the gate is the Float64 oracle's own
    error."""
    sc = _sc()
    x, y, z, q = _bunch(20_000, 31, zshift=4e-3 if cath else 0.0)
    rng = np.random.default_rng(32)
    white = np.asfortranarray(rng.standard_normal(tuple(grid) + (3,)))
    m0 = sc.Mesh3D(grid, x, y, z, T=torch.float64, gamma=gamma)
    sc.deposit(m0, x, y, z, q)
    rho = np.asfortranarray(mesh_to_np(m0._rho, grid))
    field = orc.solve(rho, m0.min_bounds, m0.max_bounds, m0.delta, m0.gamma, cath)
    field = np.asfortranarray(field / np.abs(field).max())
    for kind, g in (("field", field), ("white", white)):
        truth = orc.solve_transpose_extended(g, m0.min_bounds, m0.max_bounds, m0.delta, m0.gamma,
                                             cath)
        ref = orc.solve_transpose(g, m0.min_bounds, m0.max_bounds, m0.delta, m0.gamma, cath)
        e_orc = rel_l2(ref, truth)
        for T in (torch.float64, torch.float32):
            mesh = sc.Mesh3D(grid, x, y, z, T=T, gamma=gamma)
            if generic:
                mesh.force_generic_transforms()
            got = sc.solve_adjoint(mesh, torch.as_tensor(g, device=DEV), at_cathode=cath)
            torch.cuda.synchronize()
            assert got.shape == tuple(grid) and got.dtype == T
            if T == torch.float32:
                # the Float32 mesh's bounds are Float32 values: its own yardstick
                truth = orc.solve_transpose_extended(g, mesh.min_bounds, mesh.max_bounds,
                                                     mesh.delta, mesh.gamma, cath)
            e_gpu = rel_l2(got.cpu().numpy().astype(np.float64), truth)
            print(f"solve adjoint {kind} {T} {grid} gamma={gamma} cathode={cath} "
                  f"generic={generic}: vs extended GPU {e_gpu:.2e} oracle(f64) {e_orc:.2e}")
            if T == torch.float64:
                assert e_gpu <= e_orc + 1e-14
                if kind == "field":
                    assert e_gpu <= 1e-12
            else:
                assert e_gpu <= (1e-5 if kind == "field" else 1e-4)


@pytest.mark.parametrize("grid,cath,T", [((6, 6, 1500), False, torch.float32),
                                         ((5, 6, 2048), True, torch.float32),
                                         ((1000, 6, 5), False, torch.float64),
                                         ((6, 5, 1024), True, torch.float64),
                                         ((5, 6, 4096), True, torch.float32),
                                         ((2048, 6, 5), False, torch.float64),
                                         ((6, 5, 2048), True, torch.float64)])
def test_solve_adjoint_long_axis(grid, cath, T):
    """Transforms beyond 2048 Float32 / 1024 Float64 points, where the engine
    holds twice the values per lane (on 512-lane workgroups beyond 4096 /
    2048 points); the bunch's own field as cotangent."""
    sc = _sc()
    x, y, z, q = _bunch(20_000, 33, zshift=4e-3 if cath else 0.0)
    mesh = sc.Mesh3D(grid, x, y, z, T=T)
    sc.deposit(mesh, x, y, z, q)
    rho = np.asfortranarray(mesh_to_np(mesh._rho, grid).astype(np.float64))
    field = orc.solve(rho, mesh.min_bounds, mesh.max_bounds, mesh.delta, mesh.gamma, cath)
    g = np.asfortranarray(field / np.abs(field).max())
    got = sc.solve_adjoint(mesh, torch.as_tensor(g, device=DEV), at_cathode=cath)
    torch.cuda.synchronize()
    assert got.shape == tuple(grid) and got.dtype == T
    args = (g, mesh.min_bounds, mesh.max_bounds, mesh.delta, mesh.gamma, cath)
    truth = orc.solve_transpose_extended(*args)
    e_gpu = rel_l2(got.cpu().numpy().astype(np.float64), truth)
    if T == torch.float64:
        e_orc = rel_l2(orc.solve_transpose(*args), truth)
        print(f"solve adjoint long axis {T} {grid} cathode={cath}: vs extended GPU {e_gpu:.2e} "
              f"oracle(f64) {e_orc:.2e}")
        assert e_gpu <= 1e-12 and e_gpu <= e_orc + 1e-14
    else:
        print(f"solve adjoint long axis {T} {grid} cathode={cath}: vs extended {e_gpu:.2e}")
        assert e_gpu <= 1e-5


@pytest.mark.parametrize("cath", [False, True])
def test_solve_adjoint_dot_product_identity(cath):
    """<S rho, g> = <rho, S^T g> with the GPU's forward and transposed solves
    (64^3, Float64)."""
    sc = _sc()
    grid = (64, 64, 64)
    x, y, z, q = _bunch(200_000, 33, zshift=5e-3 if cath else 0.0)
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float64, gamma=2.0)
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh, at_cathode=cath)
    g = torch.randn(grid + (3,), dtype=torch.float64, device=DEV,
                    generator=torch.Generator(device=DEV).manual_seed(34))
    lhs = float((mesh.efield * g).sum())
    rho_g = sc.solve_adjoint(mesh, g, at_cathode=cath)
    rhs = float((mesh.rho * rho_g).sum())
    scale = float((mesh.efield.abs() * g.abs()).sum())
    print(f"dot identity cathode={cath}: {lhs:.16e} vs {rhs:.16e}")
    assert abs(lhs - rhs) <= 1e-13 * scale


@pytest.mark.parametrize("pt,mt,layout", [
    (torch.float64, torch.float64, 1), (torch.float32, torch.float32, 1),
    (torch.float64, torch.float32, 1), (torch.float32, torch.float64, 1),
    (torch.float64, torch.float64, 2), (torch.float32, torch.float64, 2)])
def test_deposit_and_interpolate_adjoint_match_oracle(pt, mt, layout):
    """layout: the interpolation's field copy, one entry per node (1) or node
    pairs along x (2, Float64 meshes); particles on and beyond the upper x
    face included."""
    sc = _sc()
    grid = (17, 12, 20)
    x, y, z, q = _bunch(30_000, 35, dtype=pt)
    mesh = sc.Mesh3D(grid, x, y, z, T=mt)
    mesh.set_field_layout(layout)
    hi, lo, dl = mesh.max_bounds, mesh.min_bounds, mesh.delta
    x[:3] = torch.tensor([hi[0], hi[0] + 0.3 * dl[0], lo[0]], dtype=pt)
    y[:3] = torch.tensor([0.0, hi[1], lo[1]], dtype=pt)
    rng = np.random.default_rng(36)
    g_rho = np.asfortranarray(rng.standard_normal(grid))
    ef = np.asfortranarray(rng.standard_normal(grid + (3,)))
    mesh.efield.copy_(torch.as_tensor(ef, device=DEV))
    gE = [torch.as_tensor(rng.standard_normal(x.numel()), dtype=pt, device=DEV) for _ in range(3)]
    gq, gx, gy, gz = sc.deposit_adjoint(mesh, x, y, z, q, torch.as_tensor(g_rho, device=DEV))
    gef, hx, hy, hz = sc.interpolate_adjoint(mesh, x, y, z, *gE)
    torch.cuda.synchronize()
    assert gq.dtype == pt and hx.dtype == pt and gef.dtype == mt
    xn, yn, zn, qn = (t.cpu().numpy().astype(np.float64) for t in (x, y, z, q))
    mnp = np.float64 if mt == torch.float64 else np.float32
    g_m = g_rho.astype(mnp).astype(np.float64)
    ef_m = np.asfortranarray(ef.astype(mnp).astype(np.float64))
    # the cell frame in the kernel's arithmetic type (the wider of the two)
    fdt = np.float32 if (pt == torch.float32 and mt == torch.float32) else np.float64
    r_gq, r_gx, r_gy, r_gz = orc.deposit_vjp(xn, yn, zn, qn, g_m, mesh.min_bounds, mesh.delta, fdt)
    gEn = [t.cpu().numpy().astype(np.float64) for t in gE]
    r_gef, r_hx, r_hy, r_hz = orc.interpolate_vjp(ef_m, xn, yn, zn, gEn, mesh.min_bounds,
                                                  mesh.delta, fdt)
    tol = 1e-12 if (pt == torch.float64 and mt == torch.float64) else 1e-5
    for name, a, b in (("g_q", gq, r_gq), ("g_x dep", gx, r_gx), ("g_y dep", gy, r_gy),
                       ("g_z dep", gz, r_gz), ("g_x itp", hx, r_hx), ("g_y itp", hy, r_hy),
                       ("g_z itp", hz, r_hz)):
        a = a.cpu().numpy().astype(np.float64)
        err = np.abs(a - b).max() / np.abs(b).max()
        print(f"{pt}->{mt} {name}: max rel err {err:.2e}")
        assert err <= tol, name
    got_ef = mesh_to_np(gef.permute(3, 2, 1, 0).contiguous().view(-1), grid).astype(np.float64)
    err = np.abs(got_ef - r_gef).max() / np.abs(r_gef).max()
    print(f"{pt}->{mt} g_efield: max rel err {err:.2e}")
    assert err <= (1e-13 if mt == torch.float64 else 1e-5)


def _step_loss(sc, mesh, x, y, z, q, gE, cath):
    sc.deposit(mesh, x, y, z, q)
    sc.solve(mesh, at_cathode=cath)
    E = sc.interpolate_field(mesh, x, y, z)
    return float(sum((e * w).sum() for e, w in zip(E, gE)))


@pytest.mark.parametrize("cath", [False, True])
def test_step_gradient_matches_central_differences(cath):
    """L = sum_p gE . E(p) after deposit -> solve -> interpolate; dL/dpos and
    dL/dq from the VJP chain against central differences of the GPU step."""
    sc = _sc()
    grid = (32, 32, 32)
    n = 20_000
    x, y, z, q = _bunch(n, 37, zshift=4e-3 if cath else 0.0)
    q = q * torch.linspace(0.5, 1.5, n, dtype=torch.float64, device=DEV)
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float64, gamma=1.5)
    rng = np.random.default_rng(38)
    gE = [torch.as_tensor(rng.standard_normal(n), device=DEV) for _ in range(3)]
    # forward, then the VJP chain (position terms of interpolation and deposit summed)
    _step_loss(sc, mesh, x, y, z, q, gE, cath)
    gef, gx1, gy1, gz1 = sc.interpolate_adjoint(mesh, x, y, z, *gE)
    g_rho = sc.solve_adjoint(mesh, gef, at_cathode=cath)
    gq, gx2, gy2, gz2 = sc.deposit_adjoint(mesh, x, y, z, q, g_rho)
    grad = [(gx1 + gx2).cpu().numpy(), (gy1 + gy2).cpu().numpy(), (gz1 + gz2).cpu().numpy()]
    gq = gq.cpu().numpy()
    # Within a cell rho and the interpolation weights are linear in a
    # particle's position, so L is quadratic in it and central differences
    # are exact up to round-off: particles at least 5% of a cell from every
    # face, steps of 1e-4 cell.
    pos_np = [t.cpu().numpy() for t in (x, y, z)]
    frac = [((pos_np[a] - mesh.min_bounds[a]) / mesh.delta[a]) % 1.0 for a in range(3)]
    ok = np.flatnonzero(np.all([(f > 0.05) & (f < 0.95) for f in frac], axis=0))
    worst = 0.0
    for p in ok[np.linspace(0, ok.size - 1, 6).astype(int)]:
        p = int(p)
        for a, t in enumerate((x, y, z)):
            h = 1e-4 * mesh.delta[a]
            pos = [x, y, z]
            t_up, t_dn = t.clone(), t.clone()
            t_up[p] += h
            t_dn[p] -= h
            pos[a] = t_up
            up = _step_loss(sc, mesh, *pos, q, gE, cath)
            pos[a] = t_dn
            dn = _step_loss(sc, mesh, *pos, q, gE, cath)
            fd = (up - dn) / (2 * h)
            err = abs(fd - grad[a][p]) / abs(grad[a][p])
            worst = max(worst, err)
            assert err < 1e-5, (p, a, fd, grad[a][p])
        dq = 1e-3 * float(q[p])
        q_up, q_dn = q.clone(), q.clone()
        q_up[p] += dq
        q_dn[p] -= dq
        fd = (_step_loss(sc, mesh, x, y, z, q_up, gE, cath) -
              _step_loss(sc, mesh, x, y, z, q_dn, gE, cath)) / (2 * dq)
        err = abs(fd - gq[p]) / abs(gq[p])
        worst = max(worst, err)
        assert err < 1e-5, (p, "q", fd, gq[p])
    print(f"step gradient vs central differences, cathode={cath}: worst rel err {worst:.2e}")


def test_autograd_step_equals_explicit_chain():
    sc = _sc()
    import spacecharge.autograd as scad
    grid = (32, 32, 64)
    n = 50_000
    x, y, z, q = _bunch(n, 39)
    mesh = sc.Mesh3D(grid, x, y, z, T=torch.float64, gamma=2.0)
    xs, ys, zs, qs = (t.clone().requires_grad_(True) for t in (x, y, z, q))
    Ex, Ey, Ez = scad.step(mesh, xs, ys, zs, qs)
    rng = np.random.default_rng(40)
    gE = [torch.as_tensor(rng.standard_normal(n), device=DEV) for _ in range(3)]
    loss = (Ex * gE[0]).sum() + (Ey * gE[1]).sum() + (Ez * gE[2]).sum()
    loss.backward()
    # explicit chain on a second mesh of the same geometry
    m2 = sc.Mesh3D(grid, x, y, z, T=torch.float64, gamma=2.0)
    sc.deposit(m2, x, y, z, q)
    sc.solve(m2)
    E2 = sc.interpolate_field(m2, x, y, z)
    for a, b in zip((Ex, Ey, Ez), E2):
        assert float((a.detach() - b).abs().max()) <= 1e-12 * float(b.abs().max())
    gef, gx1, gy1, gz1 = sc.interpolate_adjoint(m2, x, y, z, *gE)
    gq, gx2, gy2, gz2 = sc.deposit_adjoint(m2, x, y, z, q, sc.solve_adjoint(m2, gef))
    for name, a, b in (("x", xs.grad, gx1 + gx2), ("y", ys.grad, gy1 + gy2),
                       ("z", zs.grad, gz1 + gz2), ("q", qs.grad, gq)):
        err = float((a - b).abs().max() / b.abs().max())
        print(f"autograd vs chain d/d{name}: {err:.2e}")
        assert err < 1e-11, name


def test_adjoint_argument_errors():
    sc = _sc()
    from spacecharge import _lib
    x, y, z, q = _bunch(100, 41)
    mesh = sc.Mesh3D((8, 8, 8), x, y, z, T=torch.float64)
    lib = _lib.load()
    g = torch.zeros(512 * 3, dtype=torch.float64, device=DEV)
    out = torch.zeros(512, dtype=torch.float64, device=DEV)
    lo, dl, hi = _lib.d3(mesh.min_bounds), _lib.d3(mesh.delta), _lib.d3(mesh.max_bounds)
    assert lib.sc_solve_adjoint(mesh._h, lo, hi, dl, 1.0, 0, None, out.data_ptr()) != 0
    assert lib.sc_solve_adjoint(mesh._h, lo, hi, dl, 0.0, 0, g.data_ptr(), out.data_ptr()) != 0
    gx = torch.zeros(100, dtype=torch.float64, device=DEV)
    # position outputs must come as a set
    assert lib.sc_deposit_adjoint(mesh._h, 8, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                                  q.data_ptr(), 100, lo, dl, out.data_ptr(), None, gx.data_ptr(),
                                  None, None, 0) != 0
    assert "all NULL or all non-NULL" in lib.sc_last_error(sc._ctx(mesh.device)).decode()
    # position gradients need the charges
    assert lib.sc_deposit_adjoint(mesh._h, 8, x.data_ptr(), y.data_ptr(), z.data_ptr(), None, 100,
                                  lo, dl, out.data_ptr(), None, gx.data_ptr(), gx.data_ptr(),
                                  gx.data_ptr(), 0) != 0
    with pytest.raises(sc.SpaceChargeError, match="shape"):
        sc.solve_adjoint(mesh, torch.zeros(8, 8, 8, dtype=torch.float64, device=DEV))


def test_tracking_gradient_matches_central_differences():
    """100 drift-kick steps (Float64, 32^3 mesh with fixed bounds, segments of
    10 steps checkpointed, trajectories recomputed from the stored fields),
    loss = final rms emittance: the gradient with respect to initial
    positions and charges against central differences of the whole run
    (steps of 1e-6 of the bunch size and of the charge), against the same run
    without checkpointing, and against torch.utils.checkpoint re-running
    whole steps."""
    sc = _sc()
    import spacecharge.autograd as scad
    n, sig = 2000, 1e-3
    x, y, z, q = _bunch(n, 42, sigma=(sig, sig, sig))
    q = q * torch.linspace(0.5, 1.5, n, dtype=torch.float64, device=DEV)
    mesh = sc.Mesh3D((32, 32, 32), (-6 * sig,) * 3, (6 * sig,) * 3, T=torch.float64)
    nsteps, dt = 100, 1.0
    kick = 1e-4 * sig / 9e6  # ~1e-4 sigma of momentum per step at the bunch's peak field

    def loss_of(x0, y0, z0, q0, every, recompute="trajectory"):
        zero = torch.zeros_like(x0)
        st = scad.track(mesh, x0, y0, z0, zero, zero, zero, q0, nsteps, dt, kick,
                        checkpoint_every=every, recompute=recompute)
        return scad.rms_emittance(*st)

    xs, ys, zs, qs = (t.clone().requires_grad_(True) for t in (x, y, z, q))
    L = loss_of(xs, ys, zs, qs, 10)
    L.backward()
    grads = [xs.grad.clone(), ys.grad.clone(), zs.grad.clone(), qs.grad.clone()]
    for every, recompute in ((0, "trajectory"), (10, "step")):
        xs2, ys2, zs2, qs2 = (t.clone().requires_grad_(True) for t in (x, y, z, q))
        L2 = loss_of(xs2, ys2, zs2, qs2, every, recompute)
        # Float64 deposits add in LDS-atomic order, so runs differ in the last bits
        assert abs(float(L2) - float(L)) <= 1e-12 * abs(float(L))
        L2.backward()
        for a, b in zip(grads, (xs2.grad, ys2.grad, zs2.grad, qs2.grad)):
            assert float((a - b).abs().max()) <= 1e-10 * float(b.abs().max()), (every, recompute)
    # particles at least 5% of a cell from every face
    pos = [t.cpu().numpy() for t in (x, y, z)]
    frac = [((pos[a] - mesh.min_bounds[a]) / mesh.delta[a]) % 1.0 for a in range(3)]
    ok = np.flatnonzero(np.all([(f > 0.05) & (f < 0.95) for f in frac], axis=0))
    worst = 0.0
    with torch.no_grad():
        for p in ok[[0, ok.size // 3, 2 * ok.size // 3]]:
            p = int(p)
            for a in range(4):
                h = 1e-6 * (sig if a < 3 else float(q[p]))
                args_up = [x.clone(), y.clone(), z.clone(), q.clone()]
                args_dn = [x.clone(), y.clone(), z.clone(), q.clone()]
                args_up[a][p] += h
                args_dn[a][p] -= h
                fd = (float(loss_of(*args_up, 0)) - float(loss_of(*args_dn, 0))) / (2 * h)
                g = float(grads[a][p])
                err = abs(fd - g) / abs(g)
                worst = max(worst, err)
                assert err < 1e-4, (p, a, fd, g)
    print(f"tracking gradient vs central differences: worst rel err {worst:.2e}")


@pytest.mark.parametrize("T,tol", [(torch.float64, 1e-11), (torch.float32, 2e-4)])
@pytest.mark.parametrize("cath", [False, True])
def test_tracking_segments_of_uneven_length(T, tol, cath):
    """25 steps in segments of 10, 10 and 5 with trajectories recomputed from
    the stored fields, nonzero initial momenta, loss touching every output:
    the same loss and the same gradients (all seven inputs) as the run that
    keeps every step for autograd."""
    sc = _sc()
    import spacecharge.autograd as scad
    n, sig = 5000, 1e-3
    x, y, z, q = _bunch(n, 43, sigma=(sig, sig, 2 * sig))
    if cath:
        z = z + 9 * sig
    x, y, z, q = (t.to(T) for t in (x, y, z, q))
    gen = torch.Generator(device=DEV).manual_seed(44)
    p0 = [1e-3 * sig * torch.randn(n, generator=gen, dtype=torch.float64, device=DEV).to(T)
          for _ in range(3)]
    zlo, zhi = (0.5 * sig, 17.5 * sig) if cath else (-8 * sig, 8 * sig)
    mesh = sc.Mesh3D((32, 32, 48), (-6 * sig, -6 * sig, zlo), (6 * sig, 6 * sig, zhi), T=T)
    kick = 1e-4 * sig / 9e6

    def run(every):
        ins = [t.clone().requires_grad_(True) for t in (x, y, z, *p0, q)]
        st = scad.track(mesh, *ins[:6], ins[6], 25, 1.0, kick, checkpoint_every=every,
                        at_cathode=cath)
        w = torch.linspace(0.5, 1.5, n, dtype=T, device=DEV)
        L = scad.rms_emittance(*st) + sum(((s_ * w).sum() for s_ in st)) * 1e-3
        L.backward()
        return float(L), [t.grad for t in ins]

    L1, g1 = run(10)
    L0, g0 = run(0)
    # Float32 meshes deposit in fixed point (bit-reproducible); Float64 ones
    # add in LDS-atomic order
    if T == torch.float32:
        assert L1 == L0
    else:
        assert abs(L1 - L0) <= 1e-12 * abs(L0)
    for i, (a, b) in enumerate(zip(g1, g0)):
        err = float((a - b).abs().max()) / float(b.abs().max())
        assert err <= tol, (i, err)
