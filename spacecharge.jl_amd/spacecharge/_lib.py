"""ctypes binding of libspacecharge_mi355.so (the C ABI in include/spacecharge_mi355.h).

The shared library is built in-tree by ``spacecharge.jl_amd/build.py``.  If it
is missing this module raises at import time: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_PKG), "libspacecharge_mi355.so")

SC_F32 = 4
SC_F64 = 8
SC_OK = 0

# Every function the header declares, with its ctypes prototype.
_vp, _i, _i64, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
_pd = ctypes.POINTER(ctypes.c_double)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_sz = ctypes.c_size_t
_pvp = ctypes.POINTER(_vp)
_pi = ctypes.POINTER(_i)
_pf = ctypes.POINTER(ctypes.c_float)
PROTOTYPES = {
    "sc_ctx_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "sc_ctx_destroy": (_i, [_vp]),
    "sc_ctx_set_stream": (_i, [_vp, _vp]),
    "sc_ctx_get_stream": (_vp, [_vp]),
    "sc_synchronize": (_i, [_vp]),
    "sc_last_error": (ctypes.c_char_p, [_vp]),
    "sc_device_count": (_i, [ctypes.POINTER(_i)]),
    "sc_alloc": (_i, [_vp, _sz, ctypes.POINTER(_vp)]),
    "sc_free": (_i, [_vp, _vp]),
    "sc_memcpy_h2d": (_i, [_vp, _vp, _vp, _sz]),
    "sc_memcpy_d2h": (_i, [_vp, _vp, _vp, _sz]),
    "sc_memcpy_d2d": (_i, [_vp, _vp, _vp, _sz]),
    "sc_memset": (_i, [_vp, _vp, _i, _sz]),
    "sc_mesh_create": (_i, [_vp, _i, _i64, _i64, _i64, ctypes.POINTER(_vp)]),
    "sc_mesh_create_with_buffers": (_i, [_vp, _i, _i64, _i64, _i64, _vp, _vp, ctypes.POINTER(_vp)]),
    "sc_mesh_destroy": (_i, [_vp]),
    "sc_mesh_rho": (_vp, [_vp]),
    "sc_mesh_efield": (_vp, [_vp]),
    "sc_mesh_workspace_bytes": (_i64, [_vp]),
    "sc_mesh_set_green_reuse": (_i, [_vp, _i]),
    "sc_mesh_force_generic_transforms": (_i, [_vp, _i]),
    "sc_mesh_set_field_layout": (_i, [_vp, _i]),
    "sc_ctx_set_grouped_launch": (_i, [_vp, _i]),
    "sc_mesh_set_deposit_tiling": (_i, [_vp, _i, _i]),
    "sc_mesh_get_deposit_tiling": (_i, [_vp, _pi, _pi]),
    "sc_mesh_autotune_deposit": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i64, _pd, _pd, _i, _pf]),
    "sc_bounds_from_particles": (_i, [_vp, _i, _vp, _vp, _vp, _i64, _pi64, _i, _pd, _pd, _pd]),
    "sc_extrema": (_i, [_vp, _i, _vp, _i64, _pd, _pd]),
    "sc_mesh_auto_bounds": (_i, [_vp, _i, _vp, _vp, _vp, _i64]),
    "sc_mesh_read_bounds": (_i, [_vp, _pd, _pd, _pd]),
    "sc_clear_mesh": (_i, [_vp]),
    "sc_deposit": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i64, _pd, _pd, _i]),
    "sc_solve": (_i, [_vp, _pd, _pd, _pd, _d, _i]),
    "sc_solve_potential": (_i, [_vp, _pd, _d, _vp]),
    "sc_interpolate": (_i, [_vp, _i, _vp, _vp, _vp, _i64, _pd, _pd, _vp, _vp, _vp]),
    "sc_interpolate_noalloc": (_i, [_vp, _i, _vp, _vp, _vp, _i64, _pd, _pd, _vp, _vp, _vp]),
    "sc_sort_permutation": (_i, [_vp, _i, _vp, _vp, _vp, _i64, _pd, _pd, _vp]),
    "sc_deposit_adjoint": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _i64, _pd, _pd, _vp, _vp, _vp, _vp,
                                _vp, _i]),
    "sc_solve_adjoint": (_i, [_vp, _pd, _pd, _pd, _d, _i, _vp, _vp]),
    "sc_interpolate_adjoint": (_i, [_vp, _i, _vp, _vp, _vp, _i64, _pd, _pd, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _i]),
    "sc_ctx_set_pool_size": (_i, [_vp, _i]),
    "sc_deposit_batched": (_i, [_pvp, _i, _i, _pvp, _pvp, _pvp, _pvp, _pi64, _pd, _pd, _i]),
    "sc_solve_batched": (_i, [_pvp, _i, _pd, _pd, _pd, _pd, _pi]),
    "sc_interpolate_batched": (_i, [_pvp, _i, _i, _pvp, _pvp, _pvp, _pi64, _pd, _pd, _pvp, _pvp,
                                    _pvp]),
    "sc_timing_enable": (_i, [_vp, _i]),
    "sc_timing_reset": (_i, [_vp]),
    "sc_kernel_count": (_i, []),
    "sc_kernel_name": (ctypes.c_char_p, [_i]),
    "sc_timing_read": (_i, [_vp, _i, _pd, _pi64]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python spacecharge.jl_amd/build.py` "
            "(the space-charge step has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def d3(v) -> ctypes.Array:
    return (ctypes.c_double * 3)(*[float(a) for a in v])


def i3(v) -> ctypes.Array:
    return (ctypes.c_int64 * 3)(*[int(a) for a in v])
