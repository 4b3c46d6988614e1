"""spacecharge -- Python host layer of the MI355X space-charge step.

Mirrors the SpaceCharge.jl API (src/SpaceCharge.jl:17):

    Julia                                     Python
    Mesh3D(grid, x, y, z; T, gamma, ...)      Mesh3D(grid, x, y, z, T=..., gamma=...)
    Mesh3D(grid, min_b, max_b; T, ...)        Mesh3D(grid, min_b, max_b, T=...)
    deposit!(mesh, x, y, z, q; clear=true)    deposit(mesh, x, y, z, q, clear=True)
    clear_mesh!(mesh)                         clear_mesh(mesh)
    solve!(mesh; at_cathode=false)            solve(mesh, at_cathode=False)
    interpolate_field(mesh, x, y, z)          interpolate_field(mesh, x, y, z)

and adds batched forms for many independent bunches (deposit_batched,
solve_batched, interpolate_batched) and the scalar potential
(solve_potential, on the reference's potential_green_function), which the
reference does not have.

Every call forwards to one function of the C ABI (include/spacecharge_mi355.h)
through ctypes; the arithmetic runs in the hand-written gfx950 kernels of
libspacecharge_mi355.so.  Device memory and streams come from torch: mesh
arrays are torch tensors on the GPU, and each call runs on torch's current
stream, asynchronously, like the reference's GPU backend.

``mesh.rho`` is a (nx, ny, nz) tensor and ``mesh.efield`` a (nx, ny, nz, 3)
tensor with the reference's column-major strides, so ``mesh.rho[i, j, k]``
is Julia's ``mesh.rho[i+1, j+1, k+1]``.  Errors raise SpaceChargeError with
the reference's messages where it has one.
"""
from __future__ import annotations

import contextlib
import ctypes
import math
from typing import Sequence, Tuple

import torch

from . import _lib
from ._lib import SC_F32, SC_F64

__all__ = [
    "Mesh3D",
    "deposit",
    "clear_mesh",
    "solve",
    "solve_potential",
    "interpolate_field",
    "sort_permutation",
    "deposit_adjoint",
    "solve_adjoint",
    "interpolate_adjoint",
    "SpaceChargeError",
    "CLIGHT",
    "FPEI",
    "synchronize",
    "KernelTimer",
    "deposit_batched",
    "solve_batched",
    "interpolate_batched",
    "set_stream_pool_size",
    "autotune_deposit",
    "autotune_batched",
    "set_grouped_launch",
]

# src/utils.jl:7-8
CLIGHT = 299792458.0
FPEI = CLIGHT ** 2 * 1.0e-7


class SpaceChargeError(Exception):
    """Raised where the reference raises ErrorException (error(msg))."""


_DT = {torch.float32: SC_F32, torch.float64: SC_F64}
_ctxs = {}
_ctx_streams = {}  # device index -> stream handle last given to that context


# torch's current stream as a raw handle without building a Stream object
# (~2 us of host time per call); checked against the public API once
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_raw_stream_checked = False


def _current_stream(idx: int) -> int:
    global _raw_stream, _raw_stream_checked
    if _raw_stream is not None:
        s = _raw_stream(idx)
        if _raw_stream_checked:
            return s
        if s == torch.cuda.current_stream(idx).cuda_stream:
            _raw_stream_checked = True
            return s
        _raw_stream = None
    return torch.cuda.current_stream(idx).cuda_stream


def _ctx(device: torch.device) -> ctypes.c_void_p:
    idx = device.index if device.index is not None else torch.cuda.current_device()
    c = _ctxs.get(idx)
    if c is None:
        lib = _lib.load()
        h = ctypes.c_void_p()
        rc = lib.sc_ctx_create(idx, ctypes.byref(h))
        if rc != 0:
            raise SpaceChargeError(lib.sc_last_error(None).decode())
        c = h
        _ctxs[idx] = c
    stream = _current_stream(idx)
    if _ctx_streams.get(idx) != stream:
        _lib.load().sc_ctx_set_stream(c, ctypes.c_void_p(stream))
        _ctx_streams[idx] = stream
    return c


def _check(rc: int, ctx) -> None:
    if rc != 0:
        raise SpaceChargeError(_lib.load().sc_last_error(ctx).decode())


_NO_SWITCH = contextlib.nullcontext()


def _on(device: torch.device):
    """torch.cuda.device(device) around a library call, which restores the
    caller's current device afterwards; nothing when it is already current."""
    if device.index is not None and torch.cuda.current_device() == device.index:
        return _NO_SWITCH
    return torch.cuda.device(device)


def _len(a) -> int:
    return a.shape[0] if isinstance(a, torch.Tensor) and a.dim() == 1 else len(a)


def _default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise SpaceChargeError("no GPU visible: the space-charge step runs on MI355X only")
    return torch.device("cuda", torch.cuda.current_device())


def _as_device(a, device: torch.device, dtype=None) -> torch.Tensor:
    # the common case, a flat contiguous device tensor of a supported type,
    # passes as it is (the conversions below cost ~2 us of host time per
    # array, which small meshes and tracking loops feel)
    if (isinstance(a, torch.Tensor) and a.device == device and a.dim() == 1
            and a.dtype in _DT and (dtype is None or a.dtype == dtype) and a.is_contiguous()):
        return a
    t = torch.as_tensor(a, device=device)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if t.dtype not in _DT:
        t = t.to(torch.float64)
    return t.contiguous().reshape(-1)


def _to_T(v: float, T: torch.dtype) -> float:
    if T == torch.float32:
        return float(torch.tensor(v, dtype=torch.float32).item())
    return float(v)


def _check_grid(grid_size) -> Tuple[int, int, int]:
    g = tuple(int(a) for a in grid_size)
    if len(g) != 3:
        raise SpaceChargeError("grid_size must have three elements.")
    if any(a <= 1 for a in g):
        raise SpaceChargeError("All elements of grid_size must be at least 2.")  # src/mesh.jl:106-108
    return g


class Mesh3D:
    """A 3-D Cartesian mesh on one GPU (src/mesh.jl:19-34).

    Mesh3D(grid_size, x, y, z, *, T=torch.float64, gamma=1.0, total_charge=0.0)
        bounds from the particle extrema with the reference's 1e-6*delta
        padding (src/mesh.jl:95-174);
    Mesh3D(grid_size, min_bounds, max_bounds, *, T=..., gamma=..., total_charge=...)
        manual bounds (src/mesh.jl:196-238).
    """

    def __init__(self, grid_size: Sequence[int], *args, T: torch.dtype = torch.float64,
                 gamma: float = 1.0, total_charge: float = 0.0, device=None):
        if T not in _DT:
            raise SpaceChargeError("T must be torch.float32 or torch.float64")
        # argument checks first, in the reference's order (src/mesh.jl:106-116, 206-211)
        g = _check_grid(grid_size)
        if len(args) == 3:
            px, py, pz = args
            lens = [_len(a) for a in (px, py, pz)]
            if min(lens) == 0:
                raise SpaceChargeError("Particle arrays cannot be empty.")
            if not (lens[0] == lens[1] == lens[2]):
                raise SpaceChargeError("Particle coordinate arrays must have the same length.")
            dev = torch.device(device) if device is not None else (
                px.device if isinstance(px, torch.Tensor) and px.is_cuda else _default_device())
            x = _as_device(px, dev)
            y = _as_device(py, dev, x.dtype)
            z = _as_device(pz, dev, x.dtype)
            lo, hi, dl = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
            ctx = _ctx(dev)
            with _on(dev):
                rc = _lib.load().sc_bounds_from_particles(
                    ctx, _DT[x.dtype], x.data_ptr(), y.data_ptr(), z.data_ptr(), x.numel(),
                    _lib.i3(g), _DT[T], lo, hi, dl)
            _check(rc, ctx)
            self.min_bounds = tuple(lo)
            self.max_bounds = tuple(hi)
            self.delta = tuple(dl)
        elif len(args) == 2:
            mn, mx = (tuple(float(v) for v in a) for a in args)
            if len(mn) != 3 or len(mx) != 3:
                raise SpaceChargeError("min_bounds and max_bounds must have three elements.")
            if any(b <= a for a, b in zip(mn, mx)):
                raise SpaceChargeError(
                    "max_bounds must be strictly greater than min_bounds for all dimensions.")
            dev = torch.device(device) if device is not None else _default_device()
            lo = tuple(_to_T(v, T) for v in mn)
            hi = tuple(_to_T(v, T) for v in mx)
            # delta = (T(max) - T(min)) / (n - 1) in T (:219)
            lo_t = torch.tensor(lo, dtype=T)
            hi_t = torch.tensor(hi, dtype=T)
            n1 = torch.tensor([a - 1 for a in g], dtype=T)
            self.min_bounds, self.max_bounds = lo, hi
            self.delta = tuple(float(v) for v in ((hi_t - lo_t) / n1).tolist())
        else:
            raise SpaceChargeError(
                "Mesh3D takes (grid_size, x, y, z) or (grid_size, min_bounds, max_bounds)")
        self.grid_size = g
        self.T = T
        self.device = dev
        self.gamma = _to_T(gamma, T)
        self.total_charge = _to_T(total_charge, T)
        nx, ny, nz = g
        M = nx * ny * nz
        self._rho = torch.zeros(M, dtype=T, device=dev)
        self._ef = torch.zeros(3 * M, dtype=T, device=dev)
        self.rho = self._rho.as_strided((nx, ny, nz), (1, nx, nx * ny))
        self.efield = self._ef.as_strided((nx, ny, nz, 3), (1, nx, nx * ny, M))
        ctx = _ctx(dev)
        h = ctypes.c_void_p()
        with _on(dev):
            rc = _lib.load().sc_mesh_create_with_buffers(
                ctx, _DT[T], nx, ny, nz, self._rho.data_ptr(), self._ef.data_ptr(), ctypes.byref(h))
        _check(rc, ctx)
        self._h = h
        self._device_bounds = False
        self._geom_version = 0
        self._green_reuse = False

    def __del__(self):
        h = getattr(self, "_h", None)
        # at interpreter exit the module globals may already be gone
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h is not None and lib is not None:
            lib.sc_mesh_destroy(h)
            self._h = None

    def __repr__(self) -> str:  # mirrors Base.show (src/mesh.jl:240-246)
        nx, ny, nz = self.grid_size
        lo, hi = self.min_bounds, self.max_bounds
        return (f"Mesh3D{{{str(self.T).replace('torch.', '')}, MI355X}} ({nx}x{ny}x{nz}) "
                f"bounds=[({lo[0]},{lo[1]},{lo[2]}), ({hi[0]},{hi[1]},{hi[2]})] "
                f"gamma={self.gamma}")

    @property
    def workspace_bytes(self) -> int:
        """Device bytes held by the solver workspace (0 before the first solve)."""
        return int(_lib.load().sc_mesh_workspace_bytes(self._h))

    def set_green_reuse(self, enable: bool = True) -> None:
        """Reuse the transformed Green function while delta/gamma/cathode are unchanged."""
        _check(_lib.load().sc_mesh_set_green_reuse(self._h, int(bool(enable))), None)
        self._green_reuse = bool(enable)

    @property
    def deposit_tiling(self):
        """(chunk_log2, slice_log2) of the binned deposit: 2^chunk_log2 particles
        per deposit workgroup, about 2^slice_log2 per binning workgroup."""
        c, s = ctypes.c_int(), ctypes.c_int()
        _check(_lib.load().sc_mesh_get_deposit_tiling(self._h, ctypes.byref(c), ctypes.byref(s)),
               None)
        return c.value, s.value

    def set_deposit_tiling(self, chunk_log2: int, slice_log2: int) -> None:
        """Sets the binned deposit's tiling (chunk_log2 in [10, 14], slice_log2 in
        [10, 13]); autotune_deposit picks it by measurement."""
        rc = _lib.load().sc_mesh_set_deposit_tiling(self._h, int(chunk_log2), int(slice_log2))
        _check(rc, _ctx(self.device))

    def force_generic_transforms(self, enable: bool = True) -> None:
        """Diagnostic: run every transform on the generic mixed-radix engine."""
        _check(_lib.load().sc_mesh_force_generic_transforms(self._h, int(bool(enable))), None)

    def set_field_layout(self, layout: int = 0) -> None:
        """Diagnostic: layout of the interpolation's node-interleaved field
        copy: 1 = one entry per node, 2 = node pairs along x (Float64 meshes
        only; picked by the library for Float64 fields of 64 MiB and more),
        0 = by size."""
        ctx = _ctx(self.device)
        with _on(self.device):
            rc = _lib.load().sc_mesh_set_field_layout(self._h, int(layout))
        _check(rc, ctx)

    def update_bounds(self, particles_x, particles_y, particles_z) -> None:
        """Recompute the bounds from the particle extrema with the rules of
        Mesh3D(grid, x, y, z) (src/mesh.jl:95-174) on the device, without a
        host round trip (no reference counterpart).  From then on the step
        functions read the device-resident bounds; ``min_bounds``,
        ``max_bounds`` and ``delta`` are refreshed only by fetch_bounds()."""
        lens = [_len(a) for a in (particles_x, particles_y, particles_z)]
        if min(lens) == 0:
            raise SpaceChargeError("Particle arrays cannot be empty.")
        if not (lens[0] == lens[1] == lens[2]):
            raise SpaceChargeError("Particle coordinate arrays must have the same length.")
        x, y, z = _particles(self, (particles_x, particles_y, particles_z), "xyz")
        ctx = _ctx(self.device)
        with _on(self.device):
            rc = _lib.load().sc_mesh_auto_bounds(self._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                                 z.data_ptr(), x.numel())
        _check(rc, ctx)
        self._device_bounds = True
        self._geom_version += 1

    def fetch_bounds(self):
        """Copy the device-resident bounds into min_bounds, max_bounds and
        delta (synchronises) and return them."""
        lo, hi, dl = (ctypes.c_double * 3)(), (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        ctx = _ctx(self.device)
        with _on(self.device):
            rc = _lib.load().sc_mesh_read_bounds(self._h, lo, hi, dl)
        _check(rc, ctx)
        self.min_bounds, self.max_bounds, self.delta = tuple(lo), tuple(hi), tuple(dl)
        return self.min_bounds, self.max_bounds, self.delta


def _bounds_lhd(mesh: Mesh3D):
    """(min_bounds, max_bounds, delta) arguments, NULL for device-resident
    bounds.  The ctypes arrays are kept on the mesh while its bounds stay
    the same (the library reads them during the call only)."""
    if mesh._device_bounds:
        return None, None, None
    key = (mesh.min_bounds, mesh.max_bounds, mesh.delta)
    cached = mesh.__dict__.get("_bounds_args")
    if cached is None or cached[0] != key:
        cached = (key, (_lib.d3(key[0]), _lib.d3(key[1]), _lib.d3(key[2])))
        mesh._bounds_args = cached
    return cached[1]


def _bounds_ld(mesh: Mesh3D):
    """(min_bounds, delta) arguments: NULL when the mesh's bounds live on the device."""
    lo, _, dl = _bounds_lhd(mesh)
    return lo, dl


def _particles(mesh: Mesh3D, arrays, names) -> list:
    out = []
    first = None
    for a in arrays:
        t = _as_device(a, mesh.device, first)
        if first is None:
            first = t.dtype
        out.append(t)
    return out


def clear_mesh(mesh: Mesh3D) -> None:
    """clear_mesh! (src/deposition.jl:10-12): rho = 0."""
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        _check(_lib.load().sc_clear_mesh(mesh._h), ctx)


def deposit(mesh: Mesh3D, particles_x, particles_y, particles_z, particles_q, *,
            clear: bool = True) -> None:
    """deposit! (src/deposition.jl:218-247): CIC charge deposit into mesh.rho."""
    lens = [_len(a) for a in (particles_x, particles_y, particles_z, particles_q)]
    if not all(n == lens[0] for n in lens):
        raise SpaceChargeError("Particle coordinate and charge arrays must have the same length.")
    x, y, z, q = _particles(mesh, (particles_x, particles_y, particles_z, particles_q), "xyzq")
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_deposit(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                    z.data_ptr(), q.data_ptr(), x.numel(),
                                    *_bounds_ld(mesh), int(bool(clear)))
    _check(rc, ctx)


def solve(mesh: Mesh3D, *, at_cathode: bool = False) -> None:
    """solve! (src/solvers/free_space.jl:14-47): mesh.efield from mesh.rho."""
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_solve(mesh._h, *_bounds_lhd(mesh), float(mesh.gamma),
                                  int(bool(at_cathode)))
    _check(rc, ctx)


def solve_potential(mesh: Mesh3D, out=None):
    """Scalar potential of mesh.rho in free space (no reference solver; the
    Green function is potential_green_function, src/green_functions.jl:13-22):
    FPEI * gamma/(dx dy dz gamma) * (rho (*) cell integral of 1/r), the
    lab-frame potential, so -grad_x, -grad_y give Ex, Ey and -grad_z/gamma^2
    gives Ez.  Returns a (nx, ny, nz) tensor with rho's column-major strides
    (a view of ``out``, a contiguous tensor of nx*ny*nz elements, if given)."""
    nx, ny, nz = mesh.grid_size
    n = nx * ny * nz
    if out is None:
        out = torch.empty(n, dtype=mesh.rho.dtype, device=mesh.device)
    elif (out.dtype != mesh.rho.dtype or out.device != mesh.rho.device or out.numel() != n
          or not out.is_contiguous()):
        raise SpaceChargeError("out must be a contiguous tensor of the mesh's dtype, device and size")
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_solve_potential(mesh._h, _bounds_ld(mesh)[1], float(mesh.gamma),
                                            out.data_ptr())
    _check(rc, ctx)
    return out.as_strided((nx, ny, nz), (1, nx, nx * ny))


def interpolate_field(mesh: Mesh3D, particles_x, particles_y, particles_z):
    """interpolate_field (src/interpolation.jl:100-128): (Ex, Ey, Ez) in the particle dtype."""
    lens = [_len(a) for a in (particles_x, particles_y, particles_z)]
    if not all(n == lens[0] for n in lens):
        raise SpaceChargeError("Particle coordinate arrays must have the same length.")
    x, y, z = _particles(mesh, (particles_x, particles_y, particles_z), "xyz")
    Ex, Ey, Ez = (torch.empty_like(x) for _ in range(3))
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_interpolate(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                        z.data_ptr(), x.numel(), *_bounds_ld(mesh),
                                        Ex.data_ptr(), Ey.data_ptr(), Ez.data_ptr())
    _check(rc, ctx)
    return Ex, Ey, Ez


CHUNK_LOG2 = range(10, 15)
SLICE_LOG2 = range(10, 14)


def autotune_deposit(mesh: Mesh3D, particles_x, particles_y, particles_z, particles_q, *,
                     reps: int = 5) -> dict:
    """Times the deposit of these particles at every tiling (chunk_log2 in
    CHUNK_LOG2, slice_log2 in SLICE_LOG2; HIP events, `reps` timed calls
    each), keeps the fastest on the mesh and leaves mesh.rho holding the
    deposit.  Returns {(chunk_log2, slice_log2): ms per call}.  Synchronises;
    not allowed under graph capture.  No reference counterpart."""
    x, y, z, q = _particles(mesh, (particles_x, particles_y, particles_z, particles_q), "xyzq")
    if not (x.numel() == y.numel() == z.numel() == q.numel()):
        raise SpaceChargeError("Particle coordinate and charge arrays must have the same length.")
    times = (ctypes.c_float * (len(CHUNK_LOG2) * len(SLICE_LOG2)))()
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_mesh_autotune_deposit(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                                  z.data_ptr(), q.data_ptr(), x.numel(),
                                                  *_bounds_ld(mesh), int(reps), times)
    _check(rc, ctx)
    ns = len(SLICE_LOG2)
    return {(c, s): float(times[i * ns + j]) for i, c in enumerate(CHUNK_LOG2)
            for j, s in enumerate(SLICE_LOG2)}


def autotune_batched(meshes: Sequence[Mesh3D], xs, ys, zs, qs, *, reps: int = 5) -> list:
    """autotune_deposit for every bunch of a batch (each timed alone); the
    tilings then apply to deposit_batched and the single-bunch calls alike."""
    return [autotune_deposit(m, x, y, z, q, reps=reps) for m, x, y, z, q in zip(meshes, xs, ys, zs, qs)]


def sort_permutation(mesh: Mesh3D, particles_x, particles_y, particles_z) -> torch.Tensor:
    """Tile order of the particles on this mesh (no reference counterpart):
    an int32 tensor ``perm`` such that ``x[perm]`` groups the particles by the
    deposit's 16^3-cell tiles.  Applied to every per-particle array every
    few tracking steps, it makes the deposit and interpolation touch the mesh
    in local order; the step's results do not depend on the order."""
    x, y, z = _particles(mesh, (particles_x, particles_y, particles_z), "xyz")
    perm = torch.empty(x.numel(), dtype=torch.int32, device=mesh.device)
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_sort_permutation(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                             z.data_ptr(), x.numel(), *_bounds_ld(mesh),
                                             perm.data_ptr())
    _check(rc, ctx)
    return perm


# ---- vector-Jacobian products (no reference counterpart) ----

def _fortran_flat(mesh: Mesh3D, t, shape, name):
    """Flat tensor of the mesh's dtype and device holding a (nx, ny, nz[, 3])
    array in the reference's column-major order (a view when ``t`` already
    has those strides)."""
    t = torch.as_tensor(t, device=mesh.device)
    if tuple(t.shape) != tuple(shape):
        raise SpaceChargeError(f"{name} must have shape {tuple(shape)}")
    t = t.to(mesh.rho.dtype)
    return t.permute(*reversed(range(t.dim()))).contiguous().view(-1)


def _mesh_views(flat, mesh: Mesh3D, ncomp):
    nx, ny, nz = mesh.grid_size
    if ncomp == 1:
        return flat.as_strided((nx, ny, nz), (1, nx, nx * ny))
    return flat.as_strided((nx, ny, nz, 3), (1, nx, nx * ny, nx * ny * nz))


def deposit_adjoint(mesh: Mesh3D, particles_x, particles_y, particles_z, particles_q, g_rho, *,
                    positions: bool = True):
    """Vector-Jacobian product of deposit at the cotangent ``g_rho`` (shape
    (nx, ny, nz)): returns (g_q, g_x, g_y, g_z) in the particle dtype, with
    g_q = sum_8 w g_rho and g_pos = q * d/dpos sum_8 w g_rho (the mesh
    geometry held fixed); the position terms are None unless ``positions``."""
    x, y, z, q = _particles(mesh, (particles_x, particles_y, particles_z, particles_q), "xyzq")
    g = _fortran_flat(mesh, g_rho, mesh.grid_size, "g_rho")
    gq = torch.empty_like(x)
    gx, gy, gz = (torch.empty_like(x) for _ in range(3)) if positions else (None, None, None)
    ptr = (lambda t: 0 if t is None else t.data_ptr())
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_deposit_adjoint(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                            z.data_ptr(), q.data_ptr(), x.numel(),
                                            *_bounds_ld(mesh),
                                            g.data_ptr(), gq.data_ptr(), ptr(gx), ptr(gy), ptr(gz), 0)
    _check(rc, ctx)
    return gq, gx, gy, gz


def solve_adjoint(mesh: Mesh3D, g_efield, *, at_cathode: bool = False):
    """Vector-Jacobian product of solve: g_rho = S^T g_efield for the cotangent
    ``g_efield`` (shape (nx, ny, nz, 3)); returns a (nx, ny, nz) tensor with
    rho's strides.  Uses the mesh's solver workspace, not rho or efield."""
    g = _fortran_flat(mesh, g_efield, tuple(mesh.grid_size) + (3,), "g_efield")
    out = torch.empty(math.prod(mesh.grid_size), dtype=mesh.rho.dtype, device=mesh.device)
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_solve_adjoint(mesh._h, *_bounds_lhd(mesh), float(mesh.gamma),
                                          int(bool(at_cathode)), g.data_ptr(), out.data_ptr())
    _check(rc, ctx)
    return _mesh_views(out, mesh, 1)


def interpolate_adjoint(mesh: Mesh3D, particles_x, particles_y, particles_z, g_Ex, g_Ey, g_Ez, *,
                        positions: bool = True):
    """Vector-Jacobian product of interpolate_field at the mesh's current
    efield: returns (g_efield, g_x, g_y, g_z), g_efield a (nx, ny, nz, 3)
    tensor with efield's strides (the deposit of the cotangents) and the
    position terms (None unless ``positions``) in the particle dtype."""
    x, y, z = _particles(mesh, (particles_x, particles_y, particles_z), "xyz")
    gE = [_as_device(a, mesh.device, x.dtype) for a in (g_Ex, g_Ey, g_Ez)]
    if not all(t.numel() == x.numel() for t in gE):
        raise SpaceChargeError("Field cotangents must have one value per particle.")
    gef = torch.empty(3 * math.prod(mesh.grid_size), dtype=mesh.rho.dtype, device=mesh.device)
    gx, gy, gz = (torch.empty_like(x) for _ in range(3)) if positions else (None, None, None)
    ptr = (lambda t: 0 if t is None else t.data_ptr())
    ctx = _ctx(mesh.device)
    with _on(mesh.device):
        rc = _lib.load().sc_interpolate_adjoint(mesh._h, _DT[x.dtype], x.data_ptr(), y.data_ptr(),
                                                z.data_ptr(), x.numel(), *_bounds_ld(mesh),
                                                gE[0].data_ptr(),
                                                gE[1].data_ptr(), gE[2].data_ptr(), gef.data_ptr(),
                                                ptr(gx), ptr(gy), ptr(gz), 0)
    _check(rc, ctx)
    return _mesh_views(gef, mesh, 3), gx, gy, gz


# ---- many bunches per call (no reference counterpart) ----

def _batch_device(meshes) -> torch.device:
    if len(meshes) == 0:
        raise SpaceChargeError("a batch needs at least one mesh")
    dev = meshes[0].device
    if any(m.device != dev for m in meshes):
        raise SpaceChargeError("all meshes of a batch must be on one device")
    if len({id(m) for m in meshes}) != len(meshes):
        raise SpaceChargeError("a mesh appears twice in one batch")
    if any(m._device_bounds for m in meshes):
        raise SpaceChargeError("the batched calls take host bounds (call fetch_bounds() first)")
    return dev


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


def _flat3(vals):
    return (ctypes.c_double * (3 * len(vals)))(*[float(a) for v in vals for a in v])


def _batch_particles(meshes, groups, names):
    """Per-bunch particle tensors on the device, one particle dtype for the batch."""
    out = [[] for _ in names]
    dtype = None
    for b, m in enumerate(meshes):
        arrs = [g[b] for g in groups]
        lens = [_len(a) for a in arrs]
        if not all(n == lens[0] for n in lens):
            raise SpaceChargeError("Particle coordinate and charge arrays must have the same length.")
        for k, a in enumerate(arrs):
            t = _as_device(a, m.device, dtype)
            dtype = dtype or t.dtype
            out[k].append(t)
    return out


def deposit_batched(meshes: Sequence[Mesh3D], xs, ys, zs, qs, *, clear: bool = True) -> None:
    """deposit! for many bunches at once (bunch b into meshes[b]); the bunches run
    concurrently on the library's stream pool.  Same results as a loop of deposit()."""
    dev = _batch_device(meshes)
    x, y, z, q = _batch_particles(meshes, (xs, ys, zs, qs), "xyzq")
    n = (ctypes.c_int64 * len(meshes))(*[t.numel() for t in x])
    ctx = _ctx(dev)
    with _on(dev):
        rc = _lib.load().sc_deposit_batched(
            (ctypes.c_void_p * len(meshes))(*[m._h for m in meshes]), len(meshes), _DT[x[0].dtype],
            _ptrs(x), _ptrs(y), _ptrs(z), _ptrs(q), n, _flat3([m.min_bounds for m in meshes]),
            _flat3([m.delta for m in meshes]), int(bool(clear)))
    _check(rc, ctx)


def solve_batched(meshes: Sequence[Mesh3D], *, at_cathode=False) -> None:
    """solve! for many bunches at once; at_cathode is one flag or one per bunch."""
    dev = _batch_device(meshes)
    flags = list(at_cathode) if isinstance(at_cathode, (list, tuple)) else [at_cathode] * len(meshes)
    if len(flags) != len(meshes):
        raise SpaceChargeError("at_cathode needs one flag per mesh")
    ctx = _ctx(dev)
    with _on(dev):
        rc = _lib.load().sc_solve_batched(
            (ctypes.c_void_p * len(meshes))(*[m._h for m in meshes]), len(meshes),
            _flat3([m.min_bounds for m in meshes]), _flat3([m.max_bounds for m in meshes]),
            _flat3([m.delta for m in meshes]),
            (ctypes.c_double * len(meshes))(*[float(m.gamma) for m in meshes]),
            (ctypes.c_int * len(meshes))(*[int(bool(f)) for f in flags]))
    _check(rc, ctx)


def interpolate_batched(meshes: Sequence[Mesh3D], xs, ys, zs):
    """interpolate_field for many bunches at once: a list of (Ex, Ey, Ez) per bunch."""
    dev = _batch_device(meshes)
    x, y, z = _batch_particles(meshes, (xs, ys, zs), "xyz")
    E = [[torch.empty_like(t) for t in x] for _ in range(3)]
    n = (ctypes.c_int64 * len(meshes))(*[t.numel() for t in x])
    ctx = _ctx(dev)
    with _on(dev):
        rc = _lib.load().sc_interpolate_batched(
            (ctypes.c_void_p * len(meshes))(*[m._h for m in meshes]), len(meshes), _DT[x[0].dtype],
            _ptrs(x), _ptrs(y), _ptrs(z), n, _flat3([m.min_bounds for m in meshes]),
            _flat3([m.delta for m in meshes]), _ptrs(E[0]), _ptrs(E[1]), _ptrs(E[2]))
    _check(rc, ctx)
    return [(E[0][b], E[1][b], E[2][b]) for b in range(len(meshes))]


def set_stream_pool_size(nstreams: int, device=None) -> None:
    """Number of streams the batched calls spread bunches over (default 4)."""
    dev = torch.device(device) if device is not None else _default_device()
    ctx = _ctx(dev)
    _check(_lib.load().sc_ctx_set_pool_size(ctx, int(nstreams)), ctx)


def set_grouped_launch(particles: bool = True, solve=None, device=None) -> None:
    """Batched deposit and interpolation as one launch per phase for all
    bunches (particles=True, the default), and batched solves of same-shape
    small meshes as one launch per kernel (solve, default: as particles);
    False runs one launch sequence per bunch on the stream pool.  No
    reference counterpart."""
    dev = torch.device(device) if device is not None else _default_device()
    ctx = _ctx(dev)
    solve = particles if solve is None else solve
    phases = (1 if particles else 0) | (2 if solve else 0)
    _check(_lib.load().sc_ctx_set_grouped_launch(ctx, phases), ctx)


def synchronize(device=None) -> None:
    """Wait for all queued space-charge work (CUDA.synchronize() in the reference benches)."""
    dev = torch.device(device) if device is not None else _default_device()
    ctx = _ctx(dev)
    _check(_lib.load().sc_synchronize(ctx), ctx)


class KernelTimer:
    """Per-kernel HIP-event timers of the library, on the launching stream.

    with KernelTimer() as t:
        ... space-charge calls ...
    t.results() -> {kernel_name: (total_ms, launches)}
    """

    def __init__(self, device=None):
        self.device = torch.device(device) if device is not None else _default_device()
        self._res = {}

    def __enter__(self):
        ctx = _ctx(self.device)
        lib = _lib.load()
        _check(lib.sc_timing_reset(ctx), ctx)
        _check(lib.sc_timing_enable(ctx, 1), ctx)
        return self

    def __exit__(self, *exc):
        ctx = _ctx(self.device)
        lib = _lib.load()
        _check(lib.sc_timing_enable(ctx, 0), ctx)
        res = {}
        for k in range(lib.sc_kernel_count()):
            ms, n = ctypes.c_double(), ctypes.c_int64()
            _check(lib.sc_timing_read(ctx, k, ctypes.byref(ms), ctypes.byref(n)), ctx)
            if n.value:
                res[lib.sc_kernel_name(k).decode()] = (ms.value, int(n.value))
        self._res = res
        return False

    def results(self) -> dict:
        return dict(self._res)
