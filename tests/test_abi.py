"""CPU tests of the boundary: the C-ABI library loads, exports every function
include/spacecharge_mi355.h declares, and the host layer raises the
reference's errors before any device work.  No kernel is launched here."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import spacecharge as sc
from spacecharge import _lib

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in os.listdir(os.path.join(REPO, "include")):
        if not hdr.endswith(".h"):
            continue
        text = open(os.path.join(REPO, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^[A-Za-z_][\w \t\*]*?\b(sc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for required in ("sc_ctx_create", "sc_mesh_create", "sc_bounds_from_particles", "sc_deposit",
                     "sc_solve", "sc_interpolate", "sc_clear_mesh", "sc_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.PROTOTYPES, f"{name} has no ctypes prototype"


def _julia_ccalls():
    """(symbol, number of argument types) of every ccall into the library in
    the Julia binding."""
    text = open(os.path.join(REPO, "spacecharge.jl_amd", "julia", "SpaceChargeMI355.jl")).read()
    out = []
    for m in re.finditer(r"ccall\(\(:(sc_\w+), libsc\),\s*[\w{}]+,\s*\(", text):
        depth, i = 1, m.end()
        while depth:  # the argument-type tuple, up to its closing parenthesis
            depth += {"(": 1, ")": -1}.get(text[i], 0)
            i += 1
        types = text[m.end():i - 1]
        types = re.sub(r"\{[^{}]*\}", "", re.sub(r"\{[^{}]*\}", "", types))
        n = len([t for t in types.split(",") if t.strip()])
        out.append((m.group(1), n))
    return out


def test_julia_binding_calls_declared_symbols():
    """Every ccall of the Julia binding names a function the header declares
    and the library exports, with as many argument types as the ctypes
    prototype (written from the header) has."""
    calls = _julia_ccalls()
    assert len(calls) >= 20
    declared = set(declared_functions())
    lib = _lib.load()
    for name, nargs in calls:
        assert name in declared, name
        assert hasattr(lib, name), name
        assert nargs == len(_lib.PROTOTYPES[name][1]), (name, nargs, len(_lib.PROTOTYPES[name][1]))


def test_host_only_calls():
    lib = _lib.load()
    n = lib.sc_kernel_count()
    names = [lib.sc_kernel_name(k).decode() for k in range(n)]
    for k in ("deposit_tiles", "zconv", "interpolate", "green_f"):
        assert k in names
    assert lib.sc_kernel_name(n) == b""
    # errors without a context are reported through the thread-local slot
    assert lib.sc_ctx_destroy(None) == 0
    assert lib.sc_mesh_create(None, _lib.SC_F32, 4, 4, 4, ctypes.byref(ctypes.c_void_p())) != 0
    assert b"NULL" in lib.sc_last_error(None)
    assert lib.sc_solve(None, _lib.d3((0, 0, 0)), _lib.d3((1, 1, 1)), _lib.d3((1, 1, 1)), 1.0,
                        0) != 0


@pytest.mark.parametrize("grid", [(0, 2, 2), (2, 0, 2), (2, 2, 0), (1, 2, 2), (2, 1, 2),
                                  (2, 2, 1)])
def test_mesh_rejects_small_grid(grid):
    """test/test_mesh.jl:78-84."""
    with pytest.raises(sc.SpaceChargeError, match="All elements of grid_size must be at least 2."):
        sc.Mesh3D(grid, [0.0], [0.0], [0.0])


def test_mesh_rejects_empty_and_mismatched():
    """test/test_mesh.jl:86-90."""
    with pytest.raises(sc.SpaceChargeError, match="Particle arrays cannot be empty."):
        sc.Mesh3D((2, 2, 2), np.zeros(0), np.zeros(0), np.zeros(0))
    with pytest.raises(sc.SpaceChargeError, match="must have the same length"):
        sc.Mesh3D((4, 4, 4), [0.0, 1.0], [0.0], [0.0])


def test_mesh_rejects_bad_manual_bounds():
    """test/test_mesh.jl:92-93."""
    msg = "max_bounds must be strictly greater than min_bounds for all dimensions."
    with pytest.raises(sc.SpaceChargeError, match=msg):
        sc.Mesh3D((2, 2, 2), (0, 0, 0), (0, 0, 0))
    with pytest.raises(sc.SpaceChargeError, match=msg):
        sc.Mesh3D((2, 2, 2), (1, 1, 1), (0, 0, 0))


def test_constants():
    """src/utils.jl:7-8."""
    assert sc.CLIGHT == 299792458.0
    assert sc.FPEI == 299792458.0 ** 2 * 1e-7
