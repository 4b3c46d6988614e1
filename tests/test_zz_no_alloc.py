"""GPU: sc_interpolate_noalloc, the interpolation that allocates no device
memory (its field copy lives in a fixed per-GPU array), through the C ABI.

Each call is compared with sc_interpolate on the same mesh and particles:
the same kernels, so the same bits.  Two calls in a row, the second on a
smaller mesh and fewer particles (the fixed array keeps the first call's
contents, which the second must not see), then a mesh above the array's
limit, which must be refused with nothing written.
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


def _parts(n, seed, dtype):
    x, y, z = gaussian(n, (1e-3, 1.2e-3, 2.5e-3), seed)
    q = np.full(n, 1e-9 / n)
    return [torch.as_tensor(a, dtype=dtype, device=DEV) for a in (x, y, z, q)]


def _interp(fn, mesh, x, y, z):
    """One call of fn (sc_interpolate or sc_interpolate_noalloc) through the C ABI."""
    sc = _sc()
    from spacecharge import _lib
    out = [torch.full_like(x, float("nan")) for _ in range(3)]
    ctx = sc._ctx(mesh.device)
    rc = fn(mesh._h, _lib.SC_F32 if x.dtype == torch.float32 else _lib.SC_F64, x.data_ptr(),
            y.data_ptr(), z.data_ptr(), x.numel(), _lib.d3(mesh.min_bounds), _lib.d3(mesh.delta),
            out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
    return rc, out, ctx


def test_noalloc_interpolation_equals_interpolate():
    sc = _sc()
    from spacecharge import _lib
    lib = _lib.load()
    cases = [  # (particles, particle type, grid, mesh type)
        (2_000_000, torch.float32, (128, 128, 128), torch.float32),
        (50_000, torch.float64, (40, 24, 32), torch.float64),
    ]
    for n, pt, grid, mt in cases:
        x, y, z, q = _parts(n, 61 + n % 7, pt)
        mesh = sc.Mesh3D(grid, x, y, z, T=mt)
        sc.deposit(mesh, x, y, z, q)
        sc.solve(mesh)
        rc, got, ctx = _interp(lib.sc_interpolate_noalloc, mesh, x, y, z)
        assert rc == 0, lib.sc_last_error(ctx).decode()
        rc, ref, ctx = _interp(lib.sc_interpolate, mesh, x, y, z)
        assert rc == 0, lib.sc_last_error(ctx).decode()
        torch.cuda.synchronize()
        print(f"noalloc {grid} {mt}, {n} particles: max |E| {float(ref[0].abs().max()):.3e}")
        for c in range(3):
            assert bool(torch.isfinite(got[c]).all())
            assert torch.equal(got[c], ref[c]), (grid, c)


def test_noalloc_interpolation_refuses_meshes_above_its_limit():
    sc = _sc()
    from spacecharge import _lib
    lib = _lib.load()
    x, y, z, _ = _parts(1000, 67, torch.float64)
    # 130^3 Float64 nodes need 70.3 MB of field copy; the array holds 64 MiB
    mesh = sc.Mesh3D((130, 130, 130), x, y, z, T=torch.float64)
    rc, out, ctx = _interp(lib.sc_interpolate_noalloc, mesh, x, y, z)
    torch.cuda.synchronize()
    assert rc != 0
    assert "too large for the allocation-free interpolation" in lib.sc_last_error(ctx).decode()
    for c in range(3):
        assert bool(torch.isnan(out[c]).all())
