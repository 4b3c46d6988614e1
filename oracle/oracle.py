"""ctypes front end of the CPU oracle (libsc_oracle.so, built from sc_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain-C restatement of the
SpaceCharge.jl v1.2.0 CPU path; tests compare the HIP library against it and
bench.py times it as the CPU baseline.  The product (the HIP library and its
host layer) never imports this module.

Arrays follow the reference's Julia layout: ``rho`` is a Fortran-ordered
(nx, ny, nz) array, ``efield`` a Fortran-ordered (nx, ny, nz, 3) array, so
``rho[i, j, k]`` is Julia's ``rho[i+1, j+1, k+1]``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# $SC_ORACLE_LIB selects another build of the same source (the sanitizer
# build of the Makefile's sanitize-test target)
_LIB_PATH = os.environ.get("SC_ORACLE_LIB", os.path.join(_HERE, "libsc_oracle.so"))
_lib = None

F32, F64 = 4, 8
FPEI = 299792458.0 ** 2 * 1.0e-7  # src/utils.jl:7-8


def build() -> str:
    """Compile libsc_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    vp, ll, i, d = ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, ctypes.c_double
    pd = ctypes.POINTER(ctypes.c_double)
    pll = ctypes.POINTER(ctypes.c_longlong)
    lib.sco_bounds_auto.argtypes = [i, vp, vp, vp, ll, pll, i, pd, pd, pd]
    lib.sco_bounds_manual.argtypes = [pd, pd, pll, i, pd, pd, pd]
    lib.sco_deposit.argtypes = [i, i, vp, vp, vp, vp, ll, pll, pd, pd, vp]
    lib.sco_interpolate.argtypes = [i, i, vp, vp, vp, vp, ll, pll, pd, pd, vp, vp, vp, i]
    lib.sco_solve.argtypes = [i, pll, pd, pd, pd, d, i, vp, vp, i]
    lib.sco_field_green_function_f64.argtypes = [d, d, d]
    lib.sco_field_green_function_f64.restype = d
    lib.sco_field_green_function_f32.argtypes = [ctypes.c_float] * 3
    lib.sco_field_green_function_f32.restype = ctypes.c_float
    lib.sco_green_lattice.argtypes = [i, pll, pd, d, i, pd, vp, i]
    lib.sco_solve_extended.argtypes = [pll, pd, pd, pd, d, i, vp, vp, i]
    lib.sco_fft3d.argtypes = [i, pll, vp, i, i]
    lib.sco_fpei.restype = d
    lib.sco_solve_potential.argtypes = [i, pll, pd, d, vp, vp, i]
    lib.sco_solve_potential_extended.argtypes = [pll, pd, d, vp, vp, i]
    lib.sco_potential_green_function_f64.argtypes = [d, d, d]
    lib.sco_potential_green_function_f64.restype = d
    lib.sco_workspace_create.argtypes = [i, pll]
    lib.sco_workspace_create.restype = vp
    lib.sco_workspace_free.argtypes = [i, vp]
    lib.sco_solve_ws.argtypes = [vp, i, pd, pd, pd, d, i, vp, vp, i]
    lib.sco_solve_transpose.argtypes = [i, pll, pd, pd, pd, d, i, vp, vp, i]
    lib.sco_solve_transpose_extended.argtypes = [pll, pd, pd, pd, d, i, vp, vp, i]
    _lib = lib
    return lib


def _code(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return F32
    if dt == np.float64:
        return F64
    raise TypeError(f"unsupported dtype {dt}")


def _d3(v):
    return (ctypes.c_double * 3)(*[float(a) for a in v])


def _ll3(v):
    return (ctypes.c_longlong * 3)(*[int(a) for a in v])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class OracleError(Exception):
    pass


def bounds_auto(x, y, z, grid, T=np.float64):
    """src/mesh.jl:95-174: (min_bounds, max_bounds, delta) as mesh-type values."""
    lib = _load()
    x = np.ascontiguousarray(x)
    y = np.ascontiguousarray(y, dtype=x.dtype)
    z = np.ascontiguousarray(z, dtype=x.dtype)
    lo, hi, dl = _d3([0] * 3), _d3([0] * 3), _d3([0] * 3)
    rc = lib.sco_bounds_auto(_code(x.dtype), _ptr(x), _ptr(y), _ptr(z), x.size, _ll3(grid),
                             _code(T), lo, hi, dl)
    if rc != 0:
        raise OracleError(f"bounds_auto failed ({rc})")
    return tuple(lo), tuple(hi), tuple(dl)


def bounds_manual(mn, mx, grid, T=np.float64):
    """src/mesh.jl:196-238."""
    lib = _load()
    lo, hi, dl = _d3([0] * 3), _d3([0] * 3), _d3([0] * 3)
    rc = lib.sco_bounds_manual(_d3(mn), _d3(mx), _ll3(grid), _code(T), lo, hi, dl)
    if rc != 0:
        raise OracleError(f"bounds_manual failed ({rc})")
    return tuple(lo), tuple(hi), tuple(dl)


def deposit(x, y, z, q, grid, lo, delta, T=np.float64, rho=None):
    """src/deposition.jl:28-86,167-197: serial CIC deposit, accumulating into rho."""
    lib = _load()
    x = np.ascontiguousarray(x)
    y, z, q = (np.ascontiguousarray(a, dtype=x.dtype) for a in (y, z, q))
    if rho is None:
        rho = np.zeros(tuple(grid), dtype=T, order="F")
    assert rho.flags.f_contiguous and rho.dtype == np.dtype(T)
    rc = lib.sco_deposit(_code(x.dtype), _code(T), _ptr(x), _ptr(y), _ptr(z), _ptr(q), x.size,
                         _ll3(grid), _d3(lo), _d3(delta), _ptr(rho))
    if rc != 0:
        raise OracleError(f"deposit failed ({rc})")
    return rho


def interpolate(efield, x, y, z, lo, delta, nthreads=0):
    """src/interpolation.jl:17-128: (Ex, Ey, Ez) in the particle dtype."""
    lib = _load()
    assert efield.flags.f_contiguous and efield.ndim == 4 and efield.shape[3] == 3
    x = np.ascontiguousarray(x)
    y, z = (np.ascontiguousarray(a, dtype=x.dtype) for a in (y, z))
    out = [np.empty_like(x) for _ in range(3)]
    rc = lib.sco_interpolate(_code(x.dtype), _code(efield.dtype), _ptr(efield), _ptr(x), _ptr(y),
                             _ptr(z), x.size, _ll3(efield.shape[:3]), _d3(lo), _d3(delta),
                             _ptr(out[0]), _ptr(out[1]), _ptr(out[2]), nthreads)
    if rc != 0:
        raise OracleError(f"interpolate failed ({rc})")
    return tuple(out)


def solve(rho, lo, hi, delta, gamma=1.0, at_cathode=False, nthreads=0):
    """src/solvers/free_space.jl:14-101: efield, Fortran-ordered (nx, ny, nz, 3)."""
    lib = _load()
    assert rho.flags.f_contiguous and rho.ndim == 3
    ef = np.zeros(rho.shape + (3,), dtype=rho.dtype, order="F")
    rc = lib.sco_solve(_code(rho.dtype), _ll3(rho.shape), _d3(lo), _d3(hi), _d3(delta),
                       float(gamma), int(bool(at_cathode)), _ptr(rho), _ptr(ef), nthreads)
    if rc != 0:
        raise OracleError(f"solve failed ({rc})")
    return ef


class Workspace:
    """The solve workspace of one mesh (src/mesh.jl:49-72): the three padded
    complex arrays and the transform plans, made once and reused by every
    solve (the reference caches them on the mesh after the first solve!)."""

    def __init__(self, grid, T=np.float64):
        self._lib = _load()
        self.grid = tuple(int(g) for g in grid)
        self.T = np.dtype(T)
        self._h = self._lib.sco_workspace_create(_code(self.T), _ll3(self.grid))
        if not self._h:
            raise OracleError("workspace allocation failed")

    def solve(self, rho, lo, hi, delta, gamma=1.0, at_cathode=False, nthreads=0, out=None):
        """solve() on this workspace: the same arithmetic, no allocation or
        planning; writes into ``out`` (a Fortran (nx, ny, nz, 3) array of the
        workspace's type, like mesh.efield) when given."""
        assert rho.flags.f_contiguous and rho.shape == self.grid and rho.dtype == self.T
        if out is None:
            out = np.zeros(rho.shape + (3,), dtype=rho.dtype, order="F")
        assert out.flags.f_contiguous and out.shape == rho.shape + (3,) and out.dtype == self.T
        ef = out
        rc = self._lib.sco_solve_ws(self._h, _code(self.T), _d3(lo), _d3(hi), _d3(delta),
                                    float(gamma), int(bool(at_cathode)), _ptr(rho), _ptr(ef),
                                    nthreads)
        if rc != 0:
            raise OracleError(f"solve_ws failed ({rc})")
        return ef

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.sco_workspace_free(_code(self.T), h)
            self._h = None


def solve_extended(rho, lo, hi, delta, gamma=1.0, at_cathode=False, nthreads=0):
    """The reference algorithm in x87 extended precision (64-bit mantissa),
    rounded to Float64 at the end: a yardstick for Float64 round-off."""
    lib = _load()
    rho = np.asfortranarray(rho, dtype=np.float64)
    ef = np.zeros(rho.shape + (3,), dtype=np.float64, order="F")
    rc = lib.sco_solve_extended(_ll3(rho.shape), _d3(lo), _d3(hi), _d3(delta), float(gamma),
                                int(bool(at_cathode)), _ptr(rho), _ptr(ef), nthreads)
    if rc != 0:
        raise OracleError(f"solve_extended failed ({rc})")
    return ef


def field_green_function(x, y, z, T=np.float64):
    """src/green_functions.jl:35-38, evaluated in T."""
    lib = _load()
    if np.dtype(T) == np.float32:
        return np.float32(lib.sco_field_green_function_f32(x, y, z))
    return lib.sco_field_green_function_f64(x, y, z)


def green_lattice(grid, delta, gamma, icomp, offset=(0.0, 0.0, 0.0), T=np.float64,
                  nthreads=0):
    """src/green_functions.jl:41-112: the (2nx,2ny,2nz) differenced lattice (real part)."""
    lib = _load()
    shape = tuple(2 * int(g) for g in grid)
    cdt = np.complex64 if np.dtype(T) == np.float32 else np.complex128
    out = np.zeros(shape, dtype=cdt, order="F")
    rc = lib.sco_green_lattice(_code(T), _ll3(grid), _d3(delta), float(gamma), int(icomp),
                               _d3(offset), _ptr(out), nthreads)
    if rc != 0:
        raise OracleError(f"green_lattice failed ({rc})")
    return out


def fft3d(a, sign=-1, nthreads=0):
    """Unnormalised in-place 3-D transform of a Fortran-ordered complex array (returns it)."""
    lib = _load()
    assert a.flags.f_contiguous and a.ndim == 3
    T = np.float32 if a.dtype == np.complex64 else np.float64
    rc = lib.sco_fft3d(_code(T), _ll3(a.shape), _ptr(a), int(sign), nthreads)
    if rc != 0:
        raise OracleError(f"fft3d failed ({rc})")
    return a


def solve_potential(rho, delta, gamma=1.0, nthreads=0):
    """Scalar potential (free space) of rho, shape (nx, ny, nz), Fortran order,
    in rho's dtype: FPEI * gamma/(dx dy dz gamma) * (rho (*) cell integral of 1/r)."""
    lib = _load()
    rho = np.asfortranarray(rho)
    phi = np.zeros(rho.shape, dtype=rho.dtype, order="F")
    rc = lib.sco_solve_potential(_code(rho.dtype), _ll3(rho.shape), _d3(delta), float(gamma),
                                 _ptr(rho), _ptr(phi), nthreads)
    if rc:
        raise RuntimeError(f"sco_solve_potential failed ({rc})")
    return phi


def solve_potential_extended(rho, delta, gamma=1.0, nthreads=0):
    """solve_potential in x87 extended precision (Float64 in and out)."""
    lib = _load()
    rho = np.asfortranarray(rho, dtype=np.float64)
    phi = np.zeros(rho.shape, dtype=np.float64, order="F")
    rc = lib.sco_solve_potential_extended(_ll3(rho.shape), _d3(delta), float(gamma), _ptr(rho),
                                          _ptr(phi), nthreads)
    if rc:
        raise RuntimeError(f"sco_solve_potential_extended failed ({rc})")
    return phi


def potential_green_function(x, y, z):
    """src/green_functions.jl:13-22 in Float64."""
    return _load().sco_potential_green_function_f64(x, y, z)


# ---------------------------------------------------------------- adjoints
# The reference has no derivatives; these are the vector-Jacobian products of
# its deposit, solve and interpolation, used to check the library's.


def solve_transpose(g_efield, lo, hi, delta, gamma=1.0, at_cathode=False, nthreads=0):
    """S^T g for the field solve S (rho -> efield): g_efield is a Fortran
    (nx, ny, nz, 3) array; returns the (nx, ny, nz) cotangent of rho, in
    g_efield's dtype (see solve_transpose in sc_oracle_body.h)."""
    lib = _load()
    g = np.asfortranarray(g_efield)
    out = np.zeros(g.shape[:3], dtype=g.dtype, order="F")
    rc = lib.sco_solve_transpose(_code(g.dtype), _ll3(g.shape[:3]), _d3(lo), _d3(hi), _d3(delta),
                                 float(gamma), int(bool(at_cathode)), _ptr(g), _ptr(out), nthreads)
    if rc != 0:
        raise OracleError(f"solve_transpose failed ({rc})")
    return out


def solve_transpose_extended(g_efield, lo, hi, delta, gamma=1.0, at_cathode=False, nthreads=0):
    """solve_transpose in x87 extended precision (Float64 in and out)."""
    lib = _load()
    g = np.asfortranarray(g_efield, dtype=np.float64)
    out = np.zeros(g.shape[:3], dtype=np.float64, order="F")
    rc = lib.sco_solve_transpose_extended(_ll3(g.shape[:3]), _d3(lo), _d3(hi), _d3(delta),
                                          float(gamma), int(bool(at_cathode)), _ptr(g), _ptr(out),
                                          nthreads)
    if rc != 0:
        raise OracleError(f"solve_transpose_extended failed ({rc})")
    return out


def _cic_frame(x, y, z, grid, lo, delta, dtype=np.float64):
    """Per particle: lower cell index (clamped to [0, n-2], src/deposition.jl:
    43-48 with out-of-mesh particles clamped) and fractions.  The normalised
    coordinate (p - lo)/delta is formed in ``dtype`` (the wider of the
    particle and mesh types, as the deposit does), the rest in Float64."""
    idx, frac = [], []
    for p, l, d, n in zip((x, y, z), lo, delta, grid):
        t = ((np.asarray(p, dtype=dtype) - dtype(l)) / dtype(d)).astype(np.float64)
        i = np.clip(np.floor(t), 0, n - 2)
        i = np.where(np.isnan(i), 0, i).astype(np.int64)
        idx.append(i)
        frac.append(t - i)
    return idx, frac


def _trilinear_grad(field, idx, frac, delta):
    """Interpolant of a Fortran (nx, ny, nz) node field at the particles and
    its x, y, z derivatives, cell index held fixed."""
    (i, j, k), (dx, dy, dz) = idx, frac
    w = [(1 - dx, dx), (1 - dy, dy), (1 - dz, dz)]
    dw = [(-1.0 / delta[0], 1.0 / delta[0]), (-1.0 / delta[1], 1.0 / delta[1]),
          (-1.0 / delta[2], 1.0 / delta[2])]
    val = np.zeros_like(dx)
    gx, gy, gz = np.zeros_like(dx), np.zeros_like(dx), np.zeros_like(dx)
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                f = field[i + a, j + b, k + c]
                val += f * w[0][a] * w[1][b] * w[2][c]
                gx += f * dw[0][a] * w[1][b] * w[2][c]
                gy += f * w[0][a] * dw[1][b] * w[2][c]
                gz += f * w[0][a] * w[1][b] * dw[2][c]
    return val, gx, gy, gz


def deposit_vjp(x, y, z, q, g_rho, lo, delta, frame_dtype=np.float64):
    """VJP of deposit at cotangent g_rho (Fortran (nx, ny, nz)): (g_q, g_x,
    g_y, g_z), Float64."""
    idx, frac = _cic_frame(x, y, z, g_rho.shape, lo, delta, frame_dtype)
    val, gx, gy, gz = _trilinear_grad(np.asarray(g_rho, dtype=np.float64), idx, frac, delta)
    q = np.asarray(q, dtype=np.float64)
    return val, q * gx, q * gy, q * gz


def interpolate_vjp(efield, x, y, z, gE, lo, delta, frame_dtype=np.float64):
    """VJP of interpolate at the field efield (Fortran (nx, ny, nz, 3)) and
    cotangents gE = (g_Ex, g_Ey, g_Ez): (g_efield, g_x, g_y, g_z), Float64.
    g_efield is the deposit of each g_Ec (the gather's transpose)."""
    grid = efield.shape[:3]
    gef = np.zeros(tuple(grid) + (3,), order="F")
    for c in range(3):
        gef[..., c] = deposit(np.asarray(x, np.float64), np.asarray(y, np.float64),
                              np.asarray(z, np.float64), np.asarray(gE[c], np.float64), grid, lo,
                              delta)
    idx, frac = _cic_frame(x, y, z, grid, lo, delta, frame_dtype)
    gx, gy, gz = 0.0, 0.0, 0.0
    for c in range(3):
        _, ax, ay, az = _trilinear_grad(np.asarray(efield[..., c], np.float64), idx, frac, delta)
        w = np.asarray(gE[c], dtype=np.float64)
        gx, gy, gz = gx + w * ax, gy + w * ay, gz + w * az
    return gef, gx, gy, gz
