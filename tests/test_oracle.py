"""CPU tests that pin the oracle (the C restatement of the reference CPU path).

Each oracle function is checked against something that does not share its
code or formula:
  * FFT                    numpy.fft (pocketfft) on random data, all radices.
  * field_green_function   40-digit mpmath values (tests/golden/igf_golden.json),
                           whose generator also checks d3F/dxdydz = x/r^3.
  * Green lattice          cell integrals of x/r^3, y/r^3, z/r^3 computed without
                           the antiderivative (golden file); the raw last plane.
  * solve                  brute-force direct sum with cells integrated by scipy;
                           the analytical isotropic Gaussian
                           (test/analytical_test.jl) and an anisotropic one (its
                           1-D integral form, which is itself checked against
                           the isotropic closed form); the cathode solve equals
                           a free-space solve of the explicitly mirrored charge.
  * deposit / interpolate  hand-computed CIC weights at exact binary positions;
                           the reference's own assertions (test/test_deposition.jl,
                           test/test_interpolation.jl); exactness on linear fields;
                           the adjoint identity <interp(E), q> = <E, deposit(q)>.
  * bounds                 the closed form of src/mesh.jl:95-174 and
                           test/test_mesh.jl's assertions.
  * adjoints (VJPs)         solve_transpose = the transpose of the solve's
                           matrix built column by column from unit charges
                           (free space and cathode); deposit / interpolation
                           VJPs = central differences of the forward oracle,
                           and the dot-product identities.
  * potential              40-digit mpmath values of the antiderivative of 1/r
                           (tests/golden/pot_golden.json, whose generator checks
                           d3P/dxdydz = 1/r) and cell integrals of 1/r computed
                           without it; brute-force direct sum; the analytical
                           potential of an isotropic Gaussian.
"""
from __future__ import annotations

import json
import os

import numpy as np
import pytest
from scipy import integrate
from scipy.special import erf

import oracle as orc
from _common import config_a, gaussian, gaussian_field, rel_l2

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "igf_golden.json")))
PGOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pot_golden.json")))
FPEI = 299792458.0 ** 2 * 1.0e-7


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("dims", [(8, 8, 8), (12, 6, 10), (14, 7, 22), (256, 2, 4), (26, 3, 5),
                                  (11, 13, 2), (64, 64, 8)])
@pytest.mark.parametrize("dt", [np.complex128, np.complex64])
def test_fft3d_matches_numpy(dims, dt):
    rng = np.random.default_rng(sum(dims))
    a = (rng.standard_normal(dims) + 1j * rng.standard_normal(dims)).astype(dt)
    a = np.asfortranarray(a)
    ref = np.fft.fftn(a.astype(np.complex128))
    fwd = orc.fft3d(a.copy(order="F"), -1)
    bwd = orc.fft3d(a.copy(order="F"), +1)
    tol = 1e-14 if dt == np.complex128 else 2e-6
    assert np.abs(fwd - ref).max() <= tol * np.abs(ref).max()
    assert np.abs(bwd - np.fft.ifftn(a.astype(np.complex128)) * a.size).max() <= tol * np.abs(ref).max()


# ---------------------------------------------------------------- Green function
def test_field_green_function_golden():
    for e in GOLD["F"]:
        x, y, z = e["point"]
        ref = float(e["F"])
        got = orc.field_green_function(x, y, z)
        assert abs(got - ref) <= 4e-15 * max(abs(ref), 1.0), (e["point"], got, ref)
        got32 = float(orc.field_green_function(x, y, z, np.float32))
        assert abs(got32 - ref) <= 2e-5 * max(abs(ref), 1.0)


@pytest.mark.parametrize("gamma", [1.0, 3.0])
def test_green_lattice_cell_integrals(gamma):
    """Differenced lattice entry = factor * (integral of comp/r^3 over the cell)."""
    grid = (32, 32, 32)
    for h in ({tuple(e["h"]) for e in GOLD["cells"]}):
        # the lattice stretches z by gamma: delta_z = h_z / gamma
        delta = (h[0], h[1], h[2] / gamma)
        vol = h[0] * h[1] * h[2]
        for icomp in (1, 2, 3):
            lat = orc.green_lattice(grid, delta, gamma, icomp).real
            factor = (gamma if icomp < 3 else 1.0) / vol
            for e in GOLD["cells"]:
                if tuple(e["h"]) != h:
                    continue
                a, b, c = e["cell"]
                # 1-based lattice index i <-> displacement (i - n)*dx
                got = lat[a + grid[0] - 1, b + grid[1] - 1, c + grid[2] - 1]
                ref = factor * float(e["I"][icomp - 1])
                dist = max(1.0, float(np.sqrt(a * a + b * b + c * c)))
                # cancellation in the 8-point difference grows like distance^3
                tol = 1e-13 * dist ** 3 * factor + 1e-12 * abs(ref)
                assert abs(got - ref) <= tol, (h, e["cell"], icomp, got, ref)


def test_green_lattice_last_plane_is_raw():
    """src/green_functions.jl:61-66 leaves the last i, j, k planes undifferenced."""
    grid, delta, gamma = (4, 5, 3), (0.5, 0.25, 1.0), 2.0
    lat = orc.green_lattice(grid, delta, gamma, 1).real
    dz = delta[2] * gamma
    i, j, k = 2 * grid[0] - 1, 1, 2  # 0-based, on the last i plane
    u = i * delta[0] + (1 - 2 * grid[0]) * delta[0] / 2
    v = j * delta[1] + (1 - 2 * grid[1]) * delta[1] / 2
    w = k * dz + (1 - 2 * grid[2]) * dz / 2
    factor = gamma / (delta[0] * delta[1] * dz)
    assert lat[i, j, k] == pytest.approx(orc.field_green_function(u, v, w) * factor, rel=1e-14)


def _cell_integrals(comp, span, h):
    """scipy integral of comp/r^3 over every cell with displacement in span."""
    out = {}
    others = [i for i in range(3) if i != comp]
    for a in range(-span[0], span[0] + 1):
        for b in range(-span[1], span[1] + 1):
            for c in range(-span[2], span[2] + 1):
                cen = (a * h[0], b * h[1], c * h[2])
                lo = [cen[i] - h[i] / 2 for i in range(3)]
                hi = [cen[i] + h[i] / 2 for i in range(3)]

                def f(v, u):
                    p = [0.0, 0.0, 0.0]
                    p[others[0]], p[others[1]] = u, v
                    p[comp] = hi[comp]
                    r1 = np.sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2)
                    p[comp] = lo[comp]
                    r0 = np.sqrt(p[0] ** 2 + p[1] ** 2 + p[2] ** 2)
                    return 1.0 / r0 - 1.0 / r1

                val, _ = integrate.dblquad(f, lo[others[0]], hi[others[0]], lo[others[1]],
                                           hi[others[1]], epsabs=1e-13, epsrel=1e-11)
                out[(a, b, c)] = val
    return out


def test_solve_equals_brute_force_direct_sum():
    """E_c(m) = FPEI * factor_c * sum_n rho(n) I_c(m - n): the zero padding, the
    n-1 shift, the 1/P normalisation, the permutation and the factor."""
    grid = (4, 3, 5)
    h = (0.5, 0.8, 0.6)
    gamma = 1.5
    rng = np.random.default_rng(5)
    rho = np.asfortranarray(rng.uniform(-1, 1, grid) * 1e-12)
    lo = (0.0, 0.0, 0.0)
    hi = tuple((g - 1) * d for g, d in zip(grid, h))
    ef = orc.solve(rho, lo, hi, h, gamma)
    hz = (h[0], h[1], h[2] * gamma)
    vol = hz[0] * hz[1] * hz[2]
    span = tuple(g - 1 for g in grid)
    for comp in range(3):
        I = _cell_integrals(comp, span, hz)
        factor = (gamma if comp < 2 else 1.0) / vol
        ref = np.zeros(grid)
        for m in np.ndindex(grid):
            s = 0.0
            for n in np.ndindex(grid):
                s += rho[n] * I[(m[0] - n[0], m[1] - n[1], m[2] - n[2])]
            ref[m] = FPEI * factor * s
        assert rel_l2(ef[..., comp], ref) < 1e-9, comp


def test_cathode_equals_mirrored_free_space():
    """solve!(at_cathode=true) == free-space solve of rho plus its image charge
    (-rho reflected through z = 0) on a mesh that spans both."""
    nx, ny, nz, K = 6, 5, 4, 2
    d = (0.5, 0.25, 0.5)
    gamma = 1.7
    rng = np.random.default_rng(9)
    rho = np.asfortranarray(rng.uniform(0, 1, (nx, ny, nz)) * 1e-12)
    lo = (0.0, 0.0, K * d[2])
    hi = ((nx - 1) * d[0], (ny - 1) * d[1], (K + nz - 1) * d[2])
    ef = orc.solve(rho, lo, hi, d, gamma, at_cathode=True)
    # doubled mesh: z from -zmax to zmax, same spacing
    nzd = 2 * (K + nz - 1) + 1
    rd = np.zeros((nx, ny, nzd), order="F")
    for k in range(nz):
        rd[:, :, 2 * K + nz - 1 + k] += rho[:, :, k]
        rd[:, :, nz - 1 - k] -= rho[:, :, k]
    lod = (0.0, 0.0, -hi[2])
    hid = (hi[0], hi[1], hi[2])
    efd = orc.solve(rd, lod, hid, d, gamma)
    ref = efd[:, :, 2 * K + nz - 1: 2 * K + 2 * nz - 1, :]
    assert rel_l2(ef, ref) < 1e-12
    free = orc.solve(rho, lo, hi, d, gamma)
    assert rel_l2(ef, free) > 1e-3  # the image field matters (test/test_solvers.jl:33-56)


def test_analytical_isotropic_gaussian():
    """Port of test/analytical_test.jl: 1M particles, sigma = 1 mm, Q = 1 nC,
    32^3 auto-bounds; on-axis Ez within 10% of the closed form."""
    eps0 = 8.8541878128e-12
    n, Q, sigma = 1_000_000, 1e-9, 1e-3
    x, y, z = gaussian(n, (sigma, sigma, sigma), seed=123)
    q = np.full(n, Q / n)
    grid = (32, 32, 32)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    ef = orc.solve(rho, lo, hi, dl)
    zc = np.array([lo[2] + k * dl[2] for k in range(grid[2])])
    ix = int(np.argmin(np.abs([lo[0] + i * dl[0] for i in range(grid[0])])))
    iy = int(np.argmin(np.abs([lo[1] + j * dl[1] for j in range(grid[1])])))
    r = np.abs(zc)
    ana = Q / (4 * np.pi * eps0 * r ** 3) * (erf(r / (np.sqrt(2) * sigma)) -
                                             np.sqrt(2 / np.pi) * r / sigma *
                                             np.exp(-(r / sigma) ** 2 / 2)) * zc
    err = np.abs(ef[ix, iy, :, 2] - ana).max() / np.abs(ana).max()
    assert err < 0.10


def test_gaussian_field_formula_reduces_to_isotropic_closed_form():
    """The anisotropic Gaussian's 1-D integral (tests/_common.py) at equal
    sizes equals Q/(4 pi eps0 r^2) (erf(r/(sqrt 2 s)) - sqrt(2/pi) r/s e^(-r^2/2s^2))."""
    Q, s = 1e-9, 1e-3
    r = np.array([0.2, 0.7, 1.5, 3.0, 6.0]) * s
    pts = np.stack([r * 0.6, r * 0.0, r * 0.8], axis=1)
    E = gaussian_field(pts, (s, s, s), Q)
    rr = np.linalg.norm(pts, axis=1)
    Er = FPEI * Q / rr ** 2 * (erf(rr / (np.sqrt(2) * s)) -
                               np.sqrt(2 / np.pi) * rr / s * np.exp(-rr ** 2 / (2 * s * s)))
    ana = Er[:, None] * pts / rr[:, None]
    assert np.abs(E - ana).max() <= 1e-12 * np.abs(ana).max()


def _axis_fields(ef, lo, dl, grid):
    """Nodes and field along the mesh lines through the node nearest the origin."""
    ix = [int(np.argmin(np.abs(lo[a] + np.arange(grid[a]) * dl[a]))) for a in range(3)]
    out = []
    for c in range(3):
        idx = list(ix)
        pts = []
        for k in range(grid[c]):
            idx[c] = k
            pts.append([lo[a] + idx[a] * dl[a] for a in range(3)])
        sl = list(ix)
        sl[c] = slice(None)
        out.append((np.array(pts), ef[tuple(sl) + (c,)]))
    return out


def test_analytical_anisotropic_gaussian():
    """1M particles, sigma = (1, 1, 3) mm, 1 nC, 32^3 auto-bounds: the field
    on the three axes within 10% (as test/analytical_test.jl asks of the
    isotropic case) of the 1-D integral form."""
    Q, sig = 1e-9, (1e-3, 1e-3, 3e-3)
    x, y, z = gaussian(1_000_000, sig, seed=5)
    q = np.full(x.size, Q / x.size)
    grid = (32, 32, 32)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    ef = orc.solve(orc.deposit(x, y, z, q, grid, lo, dl), lo, hi, dl)
    for c, (pts, got) in enumerate(_axis_fields(ef, lo, dl, grid)):
        ana = gaussian_field(pts, sig, Q)[:, c]
        assert np.abs(got - ana).max() / np.abs(ana).max() < 0.10, c


# ---------------------------------------------------------------- deposit / interpolate
def test_deposit_hand_computed_weights():
    grid = (4, 4, 4)
    lo, hi, dl = orc.bounds_manual((0, 0, 0), (3, 3, 3), grid)
    assert dl == (1.0, 1.0, 1.0)
    # fractions (1/4, 1/2, 3/4) inside cell (1, 0, 2): exact binary weights
    x, y, z, q = [1.25], [0.5], [2.75], [8.0]
    rho = orc.deposit(np.array(x), np.array(y), np.array(z), np.array(q), grid, lo, dl)
    exp = np.zeros(grid)
    for (di, wx) in ((0, 0.75), (1, 0.25)):
        for (dj, wy) in ((0, 0.5), (1, 0.5)):
            for (dk, wz) in ((0, 0.25), (1, 0.75)):
                exp[1 + di, 0 + dj, 2 + dk] = 8.0 * wx * wy * wz
    assert np.array_equal(rho, exp)


def test_deposit_reference_assertions():
    """test/test_deposition.jl:4-115 against the oracle."""
    def mesh(grid, x, y, z):
        return orc.bounds_auto(np.array(x, float), np.array(y, float), np.array(z, float), grid)

    def dep(grid, x, y, z, q, b, rho=None):
        lo, hi, dl = b
        return orc.deposit(np.array(x, float), np.array(y, float), np.array(z, float),
                           np.array(q, float), grid, lo, dl, rho=rho)

    b = mesh((4, 4, 4), [0.5], [0.5], [0.5])
    rho = dep((4, 4, 4), [0.5], [0.5], [0.5], [1.0], b)
    assert abs(rho.sum() - 1.0) <= 1e-10 and (rho > 0).any()
    b = mesh((6, 6, 6), [0.2, 0.8], [0.3, 0.7], [0.4, 0.6])
    rho = dep((6, 6, 6), [0.2, 0.8], [0.3, 0.7], [0.4, 0.6], [1.0, -1.0], b)
    assert abs(rho.sum()) <= 1e-10 and (rho > 0).any() and (rho < 0).any()
    b = mesh((4, 4, 4), [0.5], [0.5], [0.5])
    rho = dep((4, 4, 4), [0.5], [0.5], [0.5], [1.0], b)
    first = rho.sum()
    rho = dep((4, 4, 4), [0.5], [0.5], [0.5], [1.0], b, rho=rho)
    assert abs(rho.sum() - 2 * first) <= 1e-10
    b = mesh((4, 4, 4), [0.0], [0.0], [0.0])
    assert abs(dep((4, 4, 4), [0.0], [0.0], [0.0], [1.0], b).sum() - 1.0) <= 1e-10
    mx = b[1]
    b2 = mesh((4, 4, 4), [mx[0]], [mx[1]], [mx[2]])
    assert abs(dep((4, 4, 4), [mx[0]], [mx[1]], [mx[2]], [1.0], b2).sum() - 1.0) <= 1e-10


def test_deposit_particle_at_max_bound():
    """norm = n-1 exactly: all charge on the last node (clamped lower cell)."""
    grid = (5, 5, 5)
    lo, hi, dl = orc.bounds_manual((0, 0, 0), (4, 4, 4), grid)
    rho = orc.deposit(np.array([4.0]), np.array([4.0]), np.array([4.0]), np.array([1.0]), grid,
                      lo, dl)
    assert rho[4, 4, 4] == 1.0 and rho.sum() == 1.0


def test_interpolate_reference_assertions_and_exactness():
    grid = (4, 4, 4)
    lo, hi, dl = orc.bounds_auto(np.array([0.0, 1.0]), np.array([0.0, 1.0]),
                                 np.array([0.0, 1.0]), grid)
    ef = np.ones(grid + (3,), order="F")
    Ex, Ey, Ez = orc.interpolate(ef, np.array([0.0, 1.0]), np.array([0.0, 1.0]),
                                 np.array([0.0, 1.0]), lo, dl)
    assert np.allclose(Ex, 1.0) and np.allclose(Ey, 1.0) and np.allclose(Ez, 1.0)
    # linear fields are reproduced exactly (up to rounding) anywhere in the mesh
    grid = (9, 7, 11)
    lo, hi, dl = orc.bounds_manual((-1, -2, -3), (1, 2, 3), grid)
    ii, jj, kk = np.meshgrid(*[np.arange(g) for g in grid], indexing="ij")
    ef = np.zeros(grid + (3,), order="F")
    ef[..., 0] = 0.3 + 1.5 * ii - 0.25 * jj + 2.0 * kk
    ef[..., 1] = -1.0 + 0.5 * jj
    ef[..., 2] = 4.0 * kk
    rng = np.random.default_rng(3)
    px, py, pz = rng.uniform(-1, 1, 1000), rng.uniform(-2, 2, 1000), rng.uniform(-3, 3, 1000)
    Ex, Ey, Ez = orc.interpolate(ef, px, py, pz, lo, dl)
    fi, fj, fk = (px - lo[0]) / dl[0], (py - lo[1]) / dl[1], (pz - lo[2]) / dl[2]
    assert np.allclose(Ex, 0.3 + 1.5 * fi - 0.25 * fj + 2.0 * fk, rtol=1e-12, atol=1e-12)
    assert np.allclose(Ey, -1.0 + 0.5 * fj, rtol=1e-12, atol=1e-12)
    assert np.allclose(Ez, 4.0 * fk, rtol=1e-12, atol=1e-12)


def test_deposit_interpolate_adjoint():
    """sum_p q_p interp(E)_p == sum_nodes E * deposit(q): same CIC weights."""
    x, y, z, q = config_a(5000, seed=8)
    grid = (12, 10, 14)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rng = np.random.default_rng(4)
    ef = np.asfortranarray(rng.standard_normal(grid + (3,)))
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    out = orc.interpolate(ef, x, y, z, lo, dl)
    for c in range(3):
        lhs = np.dot(q, out[c])
        rhs = np.sum(ef[..., c] * rho)
        assert lhs == pytest.approx(rhs, rel=1e-12)


# ---------------------------------------------------------------- bounds
def test_bounds_closed_form_and_reference_assertions():
    """src/mesh.jl:95-174 and test/test_mesh.jl:23-48."""
    grid = (10, 20, 30)
    x, y, z = np.array([-1.0, 1.0]), np.array([-2.0, 2.0]), np.array([-3.0, 3.0])
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    for a, (mn, mx) in enumerate(((-1, 1), (-2, 2), (-3, 3))):
        d0 = (mx - mn) / (grid[a] - 1)
        assert lo[a] == mn - 1e-6 * d0 and hi[a] == mx + 1e-6 * d0
        assert lo[a] < mn and hi[a] > mx
        assert dl[a] == (hi[a] - lo[a]) / (grid[a] - 1)
    lo, hi, dl = orc.bounds_manual((-1.0, -2.0, -3.0), (1.0, 2.0, 3.0), grid)
    assert lo == (-1.0, -2.0, -3.0) and hi == (1.0, 2.0, 3.0)
    assert dl == (2 / 9, 4 / 19, 6 / 29)
    # a single particle: zero extent -> delta 1e-6, bounds at the particle
    lo, hi, dl = orc.bounds_auto(np.array([0.5]), np.array([0.5]), np.array([0.5]), (4, 4, 4))
    assert dl == (1e-6, 1e-6, 1e-6) and lo == (0.5, 0.5, 0.5)


def test_bounds_float32_rules():
    """Float32 particles: first delta in Float32, padding in Float64, result in T."""
    x = np.array([0.1, 0.7], np.float32)
    lo, hi, dl = orc.bounds_auto(x, x, x, (8, 8, 8), np.float32)
    d0 = np.float32((np.float32(0.7) - np.float32(0.1)) / np.float32(7))
    l64 = float(np.float32(0.1)) - 1e-6 * float(d0)
    h64 = float(np.float32(0.7)) + 1e-6 * float(d0)
    assert lo[0] == float(np.float32(l64)) and hi[0] == float(np.float32(h64))
    assert dl[0] == float(np.float32((h64 - l64) / 7))


def test_bounds_validation():
    with pytest.raises(orc.OracleError):
        orc.bounds_auto(np.zeros(1), np.zeros(1), np.zeros(1), (1, 2, 2))
    with pytest.raises(orc.OracleError):
        orc.bounds_auto(np.zeros(0), np.zeros(0), np.zeros(0), (2, 2, 2))
    with pytest.raises(orc.OracleError):
        orc.bounds_manual((0, 0, 0), (0, 0, 0), (2, 2, 2))
    with pytest.raises(orc.OracleError):
        orc.bounds_manual((1, 1, 1), (0, 0, 0), (2, 2, 2))


def test_float32_reference_path_vs_float64():
    """The reference's own T=Float32 path (Green function in Float32) against
    Float64: recorded to show why the MI355X Float32 solve builds the Green
    function in Float64."""
    x, y, z, q = config_a(20000, seed=6)
    grid = (32, 32, 32)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    ef64 = orc.solve(rho, lo, hi, dl)
    lo32, hi32, dl32 = orc.bounds_auto(x, y, z, grid, np.float32)
    rho32 = orc.deposit(x.astype(np.float32), y.astype(np.float32), z.astype(np.float32),
                        q.astype(np.float32), grid, lo32, dl32, np.float32)
    ef32 = orc.solve(rho32, lo32, hi32, dl32)
    errs = [rel_l2(ef32[..., c], ef64[..., c]) for c in range(3)]
    print("reference Float32 path vs Float64 oracle, rel L2:", errs)
    assert max(errs) < 1e-2


@pytest.mark.parametrize("grid,gamma", [((16, 16, 16), 1.0), ((24, 20, 40), 10.0)])
def test_reference_float64_roundoff_floor(grid, gamma):
    """The reference algorithm's own Float64 error, against the same algorithm
    in x87 extended precision: nonzero and growing with the grid, ~1e-13 at
    16^3 and ~1.5e-12 at 24x20x40 with gamma = 10.  The GPU Float64 parity
    tests are measured against this floor."""
    x, y, z, q = config_a(20000, seed=4)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    e64 = orc.solve(rho, lo, hi, dl, gamma)
    ext = orc.solve_extended(rho, lo, hi, dl, gamma)
    errs = [rel_l2(e64[..., c], ext[..., c]) for c in range(3)]
    print(f"oracle Float64 vs extended {grid}: {errs}")
    assert 1e-16 < max(errs) < 1e-11


def test_reference_self_field_is_roundoff():
    """test/test_interpolation.jl:109-111 asserts |E| > 1e7 at a single charge's
    own node.  The reference path (the oracle) meets it, but only through
    round-off of its asymmetrically rounded lattice positions: the value is
    ~13 orders below the field one cell away, and the exact self-field is 0."""
    x = np.array([0.0])
    g = (8, 8, 8)
    lo, hi, dl = orc.bounds_auto(x, x, x, g)
    rho = orc.deposit(x, x, x, np.array([1.0]), g, lo, dl)
    ef = orc.solve(rho, lo, hi, dl)
    at_charge = np.abs(ef[0, 0, 0, :])
    one_cell = np.array([ef[1, 0, 0, 0], ef[0, 1, 0, 1], ef[0, 0, 1, 2]])
    assert (at_charge > 1e7).all()            # the reference's assertion holds ...
    assert (at_charge < 1e-12 * one_cell).all()  # ... at round-off level


# ---------------------------------------------------------------- scalar potential
def test_potential_green_function_golden():
    for e in PGOLD["P"]:
        x, y, z = e["point"]
        ref = float(e["P"])
        got = orc.potential_green_function(x, y, z)
        assert abs(got - ref) <= 4e-15 * max(abs(ref), 1.0), (e["point"], got, ref)
    assert orc.potential_green_function(0.0, 0.0, 0.0) == 0.0


@pytest.mark.parametrize("gamma", [1.0, 3.0])
def test_potential_lattice_cell_integrals(gamma):
    """Differenced lattice entry = gamma/vol * (integral of 1/r over the cell)."""
    grid = (32, 32, 32)
    for h in ({tuple(e["h"]) for e in PGOLD["cells"]}):
        delta = (h[0], h[1], h[2] / gamma)
        vol = h[0] * h[1] * h[2]
        lat = orc.green_lattice(grid, delta, gamma, 4).real
        factor = gamma / vol
        for e in PGOLD["cells"]:
            if tuple(e["h"]) != h:
                continue
            a, b, c = e["cell"]
            got = lat[a + grid[0] - 1, b + grid[1] - 1, c + grid[2] - 1]
            ref = factor * float(e["J"])
            dist = max(1.0, float(np.sqrt(a * a + b * b + c * c)))
            # the 8-point difference of P ~ r^2 cancels like distance^4
            tol = 1e-14 * dist ** 4 * factor + 1e-12 * abs(ref)
            assert abs(got - ref) <= tol, (h, e["cell"], got, ref)


def _cell_integrals_inv_r(span, h):
    """scipy integral of 1/r over every cell with displacement in span (z in
    closed form, the x-y integral split at the axis)."""
    out = {}
    for a in range(-span[0], span[0] + 1):
        for b in range(-span[1], span[1] + 1):
            for c in range(-span[2], span[2] + 1):
                lo = [(a - 0.5) * h[0], (b - 0.5) * h[1], (c - 0.5) * h[2]]
                hi = [(a + 0.5) * h[0], (b + 0.5) * h[1], (c + 0.5) * h[2]]

                def f(v, u):
                    rho = np.hypot(u, v)
                    return np.arcsinh(hi[2] / rho) - np.arcsinh(lo[2] / rho)

                xs = [lo[0], 0.0, hi[0]] if lo[0] < 0 < hi[0] else [lo[0], hi[0]]
                ys = [lo[1], 0.0, hi[1]] if lo[1] < 0 < hi[1] else [lo[1], hi[1]]
                val = 0.0
                for x0, x1 in zip(xs[:-1], xs[1:]):
                    for y0, y1 in zip(ys[:-1], ys[1:]):
                        val += integrate.dblquad(f, x0, x1, y0, y1, epsabs=1e-13,
                                                 epsrel=1e-11)[0]
                out[(a, b, c)] = val
    return out


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_solve_potential_equals_brute_force_direct_sum(dt):
    """phi(m) = FPEI * gamma/vol * sum_n rho(n) J(m - n) on a ragged grid, with
    the cells integrated by scipy (not through the antiderivative)."""
    grid = (4, 3, 5)
    h = (0.5, 0.8, 0.6)
    gamma = 1.5
    rng = np.random.default_rng(6)
    rho = np.asfortranarray(rng.uniform(-1, 1, grid) * 1e-12)
    phi = orc.solve_potential(rho.astype(dt), h, gamma)
    assert phi.dtype == dt
    hz = (h[0], h[1], h[2] * gamma)
    vol = hz[0] * hz[1] * hz[2]
    J = _cell_integrals_inv_r(tuple(g - 1 for g in grid), hz)
    ref = np.zeros(grid)
    for m in np.ndindex(grid):
        ref[m] = sum(rho[n] * J[(m[0] - n[0], m[1] - n[1], m[2] - n[2])]
                     for n in np.ndindex(grid))
    ref *= FPEI * gamma / vol
    # Float32: the 8-point difference of P (which grows like r^2) cancels
    # about log2(r^2/h^2) bits at the far lattice nodes
    assert rel_l2(phi, ref) < (1e-9 if dt == np.float64 else 2e-5)
    ext = orc.solve_potential_extended(rho, h, gamma)
    assert rel_l2(ext, ref) < 1e-9


def test_potential_analytical_isotropic_gaussian():
    """phi(r) = Q erf(r / (sqrt(2) sigma)) / (4 pi eps0 r) for a spherical
    Gaussian of charge Q: 1M particles, sigma = 1 mm, 1 nC, 32^3 auto-bounds,
    within 2% of the closed form along the z axis (the mesh is cut at the
    bunch edge, which truncates the far charge)."""
    eps0 = 8.8541878128e-12
    n, Q, sigma = 1_000_000, 1e-9, 1e-3
    x, y, z = gaussian(n, (sigma, sigma, sigma), seed=123)
    q = np.full(n, Q / n)
    grid = (32, 32, 32)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    phi = orc.solve_potential(rho, dl)
    zc = np.array([lo[2] + k * dl[2] for k in range(grid[2])])
    ix = int(np.argmin(np.abs([lo[0] + i * dl[0] for i in range(grid[0])])))
    iy = int(np.argmin(np.abs([lo[1] + j * dl[1] for j in range(grid[1])])))
    xc, yc = lo[0] + ix * dl[0], lo[1] + iy * dl[1]
    r = np.sqrt(xc ** 2 + yc ** 2 + zc ** 2)
    ana = Q / (4 * np.pi * eps0) * erf(r / (np.sqrt(2) * sigma)) / r
    err = np.abs(phi[ix, iy, :] - ana).max() / np.abs(ana).max()
    assert err < 0.02, err


def test_potential_gradient_is_field():
    """-grad(phi) by central differences equals the field solve (gamma = 1) in
    the interior, to the O(h^2) of the difference."""
    x, y, z, q = config_a(200_000, seed=4)
    grid = (32, 32, 32)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid)
    rho = orc.deposit(x, y, z, q, grid, lo, dl)
    phi = orc.solve_potential(rho, dl)
    ef = orc.solve(rho, lo, hi, dl)
    s = slice(4, -4)
    for ax in range(3):
        g = -(np.roll(phi, -1, ax) - np.roll(phi, 1, ax)) / (2 * dl[ax])
        e = ef[..., ax]
        err = rel_l2(g[s, s, s], e[s, s, s])
        assert err < 0.015, (ax, err)


# ---------------------------------------------------------------- adjoints
@pytest.mark.parametrize("at_cathode", [False, True])
def test_solve_transpose_is_matrix_transpose(at_cathode):
    """S built column by column (solve of each unit charge), then S^T g."""
    grid = (3, 4, 5)
    d = (0.5, 0.25, 0.4)
    gamma = 1.7
    lo = (0.0, 0.0, 2 * d[2] if at_cathode else -0.3)
    hi = tuple(l + (g - 1) * dd for l, g, dd in zip(lo, grid, d))
    M = int(np.prod(grid))
    A = np.zeros((3 * M, M))
    for n in range(M):
        e = np.zeros(M)
        e[n] = 1.0
        A[:, n] = orc.solve(np.asfortranarray(e.reshape(grid, order="F")), lo, hi, d, gamma,
                            at_cathode).ravel(order="F")
    rng = np.random.default_rng(21)
    g = np.asfortranarray(rng.standard_normal(grid + (3,)))
    ref = (A.T @ g.ravel(order="F")).reshape(grid, order="F")
    got = orc.solve_transpose(g, lo, hi, d, gamma, at_cathode)
    assert rel_l2(got, ref) < 1e-12
    ext = orc.solve_transpose_extended(g, lo, hi, d, gamma, at_cathode)
    assert rel_l2(ext, ref) < 1e-12
    # Float32: the reference's Float32 Green function is ~1e-4 off (differenced in Float32)
    g32 = orc.solve_transpose(g.astype(np.float32), lo, hi, d, gamma, at_cathode)
    assert g32.dtype == np.float32 and rel_l2(g32, ref) < 1e-3
    # the cathode image's transpose is not the image solve's negative
    if at_cathode:
        free = orc.solve_transpose(g, lo, hi, d, gamma, False)
        assert rel_l2(free, ref) > 1e-3


def _particles_in(grid, lo, d, n, seed):
    rng = np.random.default_rng(seed)
    pts = [lo[a] + rng.uniform(0.02, grid[a] - 1.02, n) * d[a] for a in range(3)]
    return pts[0], pts[1], pts[2], rng.uniform(0.5, 1.5, n) * 1e-12


def test_deposit_vjp_matches_central_differences():
    grid, lo, d = (6, 5, 7), (-0.1, 0.2, 0.0), (0.3, 0.25, 0.2)
    x, y, z, q = _particles_in(grid, lo, d, 40, 22)
    rng = np.random.default_rng(23)
    g = np.asfortranarray(rng.standard_normal(grid))
    gq, gx, gy, gz = orc.deposit_vjp(x, y, z, q, g, lo, d)

    def loss(px, py, pz, qq):
        return float(np.sum(orc.deposit(px, py, pz, qq, grid, lo, d) * g))

    # g_q: L is linear in q, so <g_q, dq> = L(dq) exactly
    dq = rng.standard_normal(x.size)
    assert abs(np.dot(gq, dq) - loss(x, y, z, dq)) <= 1e-12 * np.abs(gq).sum() * np.abs(dq).max()
    for p in range(0, 40, 5):
        for a, gref in enumerate((gx, gy, gz)):
            h = 1e-6 * d[a]
            pos = [x.copy(), y.copy(), z.copy()]
            pos[a][p] += h
            up = loss(pos[0][p:p + 1], pos[1][p:p + 1], pos[2][p:p + 1], q[p:p + 1])
            pos[a][p] -= 2 * h
            dn = loss(pos[0][p:p + 1], pos[1][p:p + 1], pos[2][p:p + 1], q[p:p + 1])
            fd = (up - dn) / (2 * h)
            assert abs(fd - gref[p]) <= 1e-7 * max(abs(gref[p]), 1e-30) + 1e-22, (p, a, fd, gref[p])


def test_interpolate_vjp_matches_central_differences():
    grid, lo, d = (6, 5, 7), (-0.1, 0.2, 0.0), (0.3, 0.25, 0.2)
    x, y, z, _ = _particles_in(grid, lo, d, 60, 24)
    rng = np.random.default_rng(25)
    ef = np.asfortranarray(rng.standard_normal(grid + (3,)))
    gE = [rng.standard_normal(x.size) for _ in range(3)]
    gef, gx, gy, gz = orc.interpolate_vjp(ef, x, y, z, gE, lo, d)
    # g_efield: <interp(E'), gE> = <E', g_efield> for any field E'
    e2 = np.asfortranarray(rng.standard_normal(grid + (3,)))
    lhs = sum(np.dot(a, b) for a, b in zip(orc.interpolate(e2, x, y, z, lo, d), gE))
    assert abs(lhs - np.sum(e2 * gef)) <= 1e-12 * np.abs(e2).sum() * np.abs(gef).max()
    # positions: each particle's terms depend on its own position only, so all
    # particles are displaced at once and differenced per particle
    for a, gref in enumerate((gx, gy, gz)):
        h = 1e-6 * d[a]
        pos = [x.copy(), y.copy(), z.copy()]
        pos[a] += h
        up = sum(e * w for e, w in zip(orc.interpolate(ef, *pos, lo, d), gE))
        pos[a] -= 2 * h
        dn = sum(e * w for e, w in zip(orc.interpolate(ef, *pos, lo, d), gE))
        fd = (up - dn) / (2 * h)
        assert np.abs(fd - gref).max() <= 1e-7 * np.abs(gref).max(), a


@pytest.mark.parametrize("T", [np.float64, np.float32])
@pytest.mark.parametrize("cathode", [False, True])
def test_workspace_solve_equals_solve(T, cathode):
    """The persistent workspace (arrays and plans made once, as the reference
    caches them on the mesh) gives the allocating solve's bits, twice in a
    row on one workspace."""
    grid = (12, 10, 14)
    x, y, z = gaussian(4000, (1e-3, 1e-3, 2e-3), 77)
    z = z + (5e-3 if cathode else 0.0)
    q = np.full(4000, 1e-12)
    lo, hi, dl = orc.bounds_auto(x, y, z, grid, T)
    rho = orc.deposit(x.astype(T), y.astype(T), z.astype(T), q.astype(T), grid, lo, dl, T)
    ref = orc.solve(rho, lo, hi, dl, 2.0, cathode)
    ws = orc.Workspace(grid, T)
    out = np.zeros(grid + (3,), dtype=T, order="F")
    for _ in range(2):
        got = ws.solve(rho, lo, hi, dl, 2.0, cathode, out=out)
        assert got is out
        assert np.array_equal(got, ref)
