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


JL_PKG = os.path.join(REPO, "spacecharge.jl_amd", "julia", "SpaceChargeMI355")
JL_SRC = os.path.join(JL_PKG, "src", "SpaceChargeMI355.jl")
JL_EXT = os.path.join(JL_PKG, "ext", "SpaceChargeMI355ChainRulesCoreExt.jl")


def _julia_ccalls():
    """(symbol, number of argument types) of every ccall into the library in
    the Julia binding."""
    text = open(JL_SRC).read()
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


# ---- the Julia binding (Julia is not installed here: these read its text) ----

def _julia_code(path):
    """Source without comments, docstrings and string literals."""
    text = open(path).read()
    text = re.sub(r'"""[\s\S]*?"""', '""', text)
    text = re.sub(r"#=[\s\S]*?=#", "", text)
    text = re.sub(r'"(?:\\.|[^"\\])*"', '""', text)
    text = re.sub(r"'(?:\\.|[^'\\])'", "''", text)
    return re.sub(r"#[^\n]*", "", text)


def _block_balance(code):
    """Block openers minus `end`s, both outside brackets: inside them `for`
    and `if` belong to comprehensions and `end` is an index."""
    openers = {"function", "macro", "module", "baremodule", "struct", "if", "for", "while",
               "begin", "let", "do", "try", "quote"}
    depth_br, bal = 0, 0
    for tok in re.finditer(r"[A-Za-z_]\w*!?|[\[\](){}]", code):
        t = tok.group(0)
        if t in "[({":
            depth_br += 1
        elif t in "])}":
            depth_br -= 1
        elif t == "end" and depth_br == 0:
            bal -= 1
        elif t in openers and depth_br == 0:
            # `mutable struct` opens one block; `abstract type ... end` does not occur here
            bal += 1
    return bal, depth_br


@pytest.mark.parametrize("path", [JL_SRC, JL_EXT])
def test_julia_files_are_balanced(path):
    """Every block opened is closed and every bracket matched (a syntax sanity
    check in the absence of a Julia parser)."""
    code = _julia_code(path)
    bal, br = _block_balance(code)
    assert bal == 0, f"{path}: {bal} unclosed blocks"
    assert br == 0, f"{path}: unbalanced brackets"
    # the check sees a dropped `end`
    k = code.rindex("end")
    assert _block_balance(code[:k] + code[k + 3:])[0] == 1


def test_julia_binding_is_a_dropin_for_mesh3d():
    """The reference's own Mesh3D constructors (src/mesh.jl:95-174, 196-238)
    build MI355X meshes: get_backend, allocate/zeros, fill! and extrema are
    defined for the device arrays, the library mesh lives in the mesh's
    _workspace slot (no global table holding meshes), and the four exported
    calls have MIArray-mesh methods."""
    code = _julia_code(JL_SRC)
    for pat in (r"KA\.get_backend\(::MIArray\)\s*=\s*MI355X\(\)",
                r"KA\.allocate\(::MI355X,",
                r"KA\.zeros\(b::MI355X,",
                r"KA\.synchronize\(::MI355X\)",
                r"function Base\.extrema\(a::MIArray",
                r"function Base\.fill!\(a::MIArray",
                r"struct MI355X <: KA\.GPU end",
                r"mesh\._workspace = \(mi355 = mh,\)",
                r"function deposit!\(mesh::MIMesh",
                r"function solve!\(mesh::MIMesh; at_cathode::Bool=false\)",
                r"function interpolate_field\(mesh::MIMesh",
                r"clear_mesh!\(mesh::MIMesh\)"):
        assert re.search(pat, code), pat
    assert "IdDict" not in code and "HANDLES" not in code
    assert "mi355_mesh" not in code


def test_julia_chainrules_extension():
    """The rrules ship as a package extension on ChainRulesCore, one per
    differentiable function, and the step's passes at_cathode to the solve
    adjoint."""
    proj = open(os.path.join(JL_PKG, "Project.toml")).read()
    assert re.search(r'\[weakdeps\]\s*\nChainRulesCore = "d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"',
                     proj)
    assert re.search(r'\[extensions\]\s*\nSpaceChargeMI355ChainRulesCoreExt = "ChainRulesCore"',
                     proj)
    assert 'SpaceCharge = "886cafed-6ce3-46b8-9147-fe45a2210a4b"' in proj
    code = _julia_code(JL_EXT)
    assert re.search(r"^module SpaceChargeMI355ChainRulesCoreExt", code, re.M)
    for fn in ("space_charge_fields", "deposit_charge", "solve_field", "interpolate_at"):
        assert re.search(r"function ChainRulesCore\.rrule\(::typeof\(%s\)" % fn, code), fn
        assert re.search(r"^function %s\(" % fn, _julia_code(JL_SRC), re.M), fn
    assert code.count("solve_adjoint(mesh, g_ef; at_cathode=at_cathode)") == 2
    for fn in ("deposit_adjoint", "solve_adjoint", "interpolate_adjoint"):
        assert re.search(r"^function %s\(" % fn, _julia_code(JL_SRC), re.M), fn


def test_integration_julia_example_uses_mesh3d():
    """The documented Julia usage is the reference's own: Mesh3D on MIArrays."""
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    i = text.index("## Julia")
    j = text.index("\n## ", i + 5)
    julia = text[i:j]
    assert "mi355_mesh" not in julia
    assert re.search(r"mesh = Mesh3D\(\(128, 128, 128\), x, y, z; T=Float32\)", julia)
    assert "backend=MI355X()" in julia
    assert "rrule" in julia and "space_charge_fields" in julia
