"""Differentiable forms of deposit, solve and interpolate_field for torch.autograd.

The reference has no derivatives; these wrap the library's vector-Jacobian
products (sc_deposit_adjoint, sc_solve_adjoint, sc_interpolate_adjoint) as
torch.autograd.Function objects, so that a tracking loop written in torch
(drift-kick steps through the space-charge field) can be differentiated with
respect to the particles' positions and charges:

    import spacecharge.autograd as scad
    rho = scad.deposit(mesh, x, y, z, q)
    efield = scad.solve(mesh, rho)
    Ex, Ey, Ez = scad.interpolate(mesh, efield, x, y, z)
    # or: Ex, Ey, Ez = scad.step(mesh, x, y, z, q)

The mesh geometry (bounds, spacing, gamma) is a constant of the
differentiation: derivatives are taken with it held fixed, and each backward
pass checks that it has not changed since the forward pass.  The functions
use the mesh's rho and efield as scratch: they return copies, and the
backward of ``interpolate`` writes the saved field back into mesh.efield
before it runs.  ``solve`` switches the mesh's Green-function reuse on
(Mesh3D.set_green_reuse) and leaves it on: the cached spectra are keyed on
the spacing, gamma, cathode offset and solve kind, so results are the same
bits as with a rebuild, and a tracking loop on a fixed mesh builds them once
for its forward solves and once for its adjoint solves.
"""
from __future__ import annotations

import torch
import torch.utils.checkpoint

from . import (SpaceChargeError, deposit as _deposit, deposit_adjoint, interpolate_adjoint,
               interpolate_field, solve as _solve, solve_adjoint)

__all__ = ["deposit", "solve", "interpolate", "step", "track", "rms_emittance"]


def _geometry(mesh):
    # device-resident bounds (Mesh3D.update_bounds) are identified by their version
    return (tuple(mesh.min_bounds), tuple(mesh.max_bounds), tuple(mesh.delta), float(mesh.gamma),
            mesh._device_bounds, mesh._geom_version)


def _check_geometry(ctx):
    if _geometry(ctx.mesh) != ctx.geometry:
        raise SpaceChargeError("the mesh geometry changed between the forward and backward pass")


class _Deposit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, x, y, z, q):
        _deposit(mesh, x, y, z, q)
        ctx.mesh, ctx.geometry = mesh, _geometry(mesh)
        ctx.save_for_backward(x, y, z, q)
        return mesh.rho.clone()

    @staticmethod
    def backward(ctx, g_rho):
        _check_geometry(ctx)
        x, y, z, q = ctx.saved_tensors
        need_pos = any(ctx.needs_input_grad[1:4])
        gq, gx, gy, gz = deposit_adjoint(ctx.mesh, x, y, z, q, g_rho, positions=need_pos)
        return None, gx, gy, gz, gq


def _reuse_green(mesh):
    if not mesh._green_reuse:
        mesh.set_green_reuse(True)


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, rho, at_cathode):
        _reuse_green(mesh)
        mesh.rho.copy_(rho)
        _solve(mesh, at_cathode=at_cathode)
        ctx.mesh, ctx.geometry, ctx.at_cathode = mesh, _geometry(mesh), at_cathode
        return mesh.efield.clone()

    @staticmethod
    def backward(ctx, g_efield):
        _check_geometry(ctx)
        _reuse_green(ctx.mesh)
        return None, solve_adjoint(ctx.mesh, g_efield, at_cathode=ctx.at_cathode), None


class _Interpolate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mesh, efield, x, y, z):
        mesh.efield.copy_(efield)
        Ex, Ey, Ez = interpolate_field(mesh, x, y, z)
        ctx.mesh, ctx.geometry = mesh, _geometry(mesh)
        ctx.save_for_backward(efield, x, y, z)
        return Ex, Ey, Ez

    @staticmethod
    def backward(ctx, gEx, gEy, gEz):
        _check_geometry(ctx)
        efield, x, y, z = ctx.saved_tensors
        gE = [torch.zeros_like(x) if g is None else g for g in (gEx, gEy, gEz)]
        mesh = ctx.mesh
        mesh.efield.copy_(efield)  # the position terms are taken at the forward field
        need_pos = any(ctx.needs_input_grad[2:5])
        g_ef, gx, gy, gz = interpolate_adjoint(mesh, x, y, z, *gE, positions=need_pos)
        return None, g_ef, gx, gy, gz


def deposit(mesh, x, y, z, q):
    """rho (nx, ny, nz) of the CIC deposit of (x, y, z, q), differentiable in x, y, z, q."""
    return _Deposit.apply(mesh, x, y, z, q)


def solve(mesh, rho, at_cathode: bool = False):
    """efield (nx, ny, nz, 3) of rho on the mesh, differentiable in rho."""
    return _Solve.apply(mesh, rho, bool(at_cathode))


def interpolate(mesh, efield, x, y, z):
    """(Ex, Ey, Ez) of efield at (x, y, z), differentiable in efield, x, y, z."""
    return _Interpolate.apply(mesh, efield, x, y, z)


def step(mesh, x, y, z, q, at_cathode: bool = False):
    """The space-charge step deposit -> solve -> interpolate, differentiable."""
    rho = deposit(mesh, x, y, z, q)
    efield = solve(mesh, rho, at_cathode)
    return interpolate(mesh, efield, x, y, z)


class _TrackSegment(torch.autograd.Function):
    """k drift-kick steps as one node of the autograd graph.  The forward
    pass keeps the segment's starting state and every step's mesh field
    (nx*ny*nz*3 values, independent of the particle count); the backward
    pass recomputes the segment's particle trajectory from them by
    interpolation alone (the same kernels and operations, so the same bits
    as the forward pass) and runs the steps' vector-Jacobian products in
    reverse:

        p' = p + kick * E(x),  x' = x + dt * p'
        Gp = g_p' + dt * g_x'          (cotangent of p')
        g_E = kick * Gp
        g_x = g_x' + (interpolation and deposit position terms of g_E),
        g_p = Gp,  g_q += deposit charge term."""

    @staticmethod
    def forward(ctx, mesh, k, dt, kick, at_cathode, x, y, z, px, py, pz, q):
        _reuse_green(mesh)
        ctx.save_for_backward(x, y, z, px, py, pz, q)
        fields = []
        for _ in range(k):
            _deposit(mesh, x, y, z, q)
            _solve(mesh, at_cathode=at_cathode)
            fields.append(mesh._ef.clone())
            Ex, Ey, Ez = interpolate_field(mesh, x, y, z)
            px, py, pz = px + kick * Ex, py + kick * Ey, pz + kick * Ez
            x, y, z = x + dt * px, y + dt * py, z + dt * pz
        ctx.mesh, ctx.geometry, ctx.fields = mesh, _geometry(mesh), fields
        ctx.dt, ctx.kick, ctx.at_cathode = dt, kick, at_cathode
        return x, y, z, px, py, pz

    @staticmethod
    def backward(ctx, *gout):
        _check_geometry(ctx)
        mesh, fields, dt, kick = ctx.mesh, ctx.fields, ctx.dt, ctx.kick
        x, y, z, px, py, pz, q = ctx.saved_tensors
        # the segment's positions, recomputed from the stored fields
        pos = []
        for f in fields:
            pos.append((x, y, z))
            mesh._ef.copy_(f)
            Ex, Ey, Ez = interpolate_field(mesh, x, y, z)
            px, py, pz = px + kick * Ex, py + kick * Ey, pz + kick * Ez
            x, y, z = x + dt * px, y + dt * py, z + dt * pz
        ref = ctx.saved_tensors[0]
        gx, gy, gz, gpx, gpy, gpz = (torch.zeros_like(ref) if g is None else g for g in gout)
        gq = torch.zeros_like(ctx.saved_tensors[6])
        for f, (x, y, z) in zip(reversed(fields), reversed(pos)):
            gpx, gpy, gpz = gpx + dt * gx, gpy + dt * gy, gpz + dt * gz
            mesh._ef.copy_(f)
            g_ef, ix, iy, iz = interpolate_adjoint(mesh, x, y, z, kick * gpx, kick * gpy, kick * gpz)
            g_rho = solve_adjoint(mesh, g_ef, at_cathode=ctx.at_cathode)
            dq, dx, dy, dz = deposit_adjoint(mesh, x, y, z, q, g_rho)
            gx, gy, gz = gx + ix + dx, gy + iy + dy, gz + iz + dz
            gq = gq + dq
        ctx.fields = None
        return None, None, None, None, None, gx, gy, gz, gpx, gpy, gpz, gq


def track(mesh, x, y, z, px, py, pz, q, nsteps: int, dt: float, kick: float,
          checkpoint_every: int = 10, at_cathode: bool = False, recompute: str = "trajectory"):
    """Drift-kick tracking through the space-charge field, differentiable in
    the initial state and the charges: each step kicks p += kick * E(x)
    (the field of the current positions, mesh geometry fixed) and drifts
    x += dt * p.  Returns (x, y, z, px, py, pz).

    With checkpoint_every = k > 0 and gradients enabled, the steps run in
    segments of k whose particle states are not kept; the backward pass
    recomputes each segment from its stored starting state:
      recompute="trajectory" (default): every step's mesh field is kept
        (3 nx ny nz values per step, e.g. 3 MB at 64^3 in Float32) and the
        recomputation is interpolation only, with the segment's reverse pass
        hand-written (_TrackSegment);
      recompute="step": torch.utils.checkpoint re-runs whole steps (deposit,
        solve, interpolation) and autograd differentiates them.
    checkpoint_every = 0 keeps every step's particles and field for
    autograd."""
    if recompute not in ("trajectory", "step"):
        raise SpaceChargeError('recompute must be "trajectory" or "step"')

    def segment(k):
        def run(x, y, z, px, py, pz):
            for _ in range(k):
                Ex, Ey, Ez = step(mesh, x, y, z, q, at_cathode)
                px, py, pz = px + kick * Ex, py + kick * Ey, pz + kick * Ez
                x, y, z = x + dt * px, y + dt * py, z + dt * pz
            return x, y, z, px, py, pz
        return run

    state = (x, y, z, px, py, pz)
    done = 0
    while done < nsteps:
        k = min(checkpoint_every, nsteps - done) if checkpoint_every > 0 else nsteps - done
        if checkpoint_every > 0 and torch.is_grad_enabled():
            if recompute == "trajectory":
                state = _TrackSegment.apply(mesh, k, float(dt), float(kick), bool(at_cathode),
                                            *state, q)
            else:
                state = torch.utils.checkpoint.checkpoint(segment(k), *state, use_reentrant=False)
        else:
            state = segment(k)(*state)
        done += k
    return state


def rms_emittance(x, y, z, px, py, pz):
    """Sum over the three planes of the rms emittance
    sqrt(<u^2><p^2> - <u p>^2) (centred moments), in Float64."""
    total = 0.0
    for u, p in ((x, px), (y, py), (z, pz)):
        u = u.double() - u.double().mean()
        p = p.double() - p.double().mean()
        total = total + torch.sqrt((u * u).mean() * (p * p).mean() - (u * p).mean() ** 2)
    return total
