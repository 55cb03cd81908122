"""ctypes bindings to the in-tree native libraries.

* ``libsunstate.so`` -- host construction of gridT/physT/propT state
  (triangulated-rectangle grids, vertical grid, flux heights, flat field access).
* ``libsuntans.so``  -- the drop-in solvers with the reference's C signatures
  (``CGSolveQ``, ``CGSolve``, ``OperatorQ``, ``Preconditioner``, ``OperatorH``,
  ``HorizontalFaceScalars``, ``UpdateScalars``) over ``libsunhip.so``, the
  gfx950 HIP kernels behind a plain-pointer C ABI.

Nothing here falls back to a CPU implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> C.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() (or python -m suntans_amd.build)")
    return C.CDLL(path)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


_state = None
_drop = None


def state_lib() -> C.CDLL:
    global _state
    if _state is None:
        L = _load("libsunstate.so")
        L.sun_grid_tri.restype = _vp
        L.sun_grid_tri.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double]
        L.sun_grid_set_param.argtypes = [_vp, C.c_char_p, C.c_double]
        L.sun_grid_param.restype = C.c_double
        L.sun_grid_param.argtypes = [_vp, C.c_char_p]
        L.sun_grid_polygons.restype = _vp
        L.sun_grid_polygons.argtypes = [C.c_int, _dp, _dp, C.c_int, C.c_int, _ip, _ip, _dp, _dp]
        L.sun_grid_vertical.argtypes = [_vp, C.c_int, _dp, _dp, C.c_int]
        L.sun_grid_open.argtypes = [_vp, _ip]
        L.sun_grid_load.restype = _vp
        L.sun_grid_load.argtypes = [C.c_char_p]
        L.sun_grid_generate.restype = _vp
        L.sun_grid_generate.argtypes = [C.c_char_p, _vp]
        L.sun_restart_write.argtypes = [C.c_char_p, _vp, _vp, C.c_int, C.c_int]
        L.sun_restart_read.argtypes = [C.c_char_p, _vp, _vp, _ip, C.c_int]
        L.sun_phys_alloc.restype = _vp
        L.sun_phys_alloc.argtypes = [_vp]
        L.sun_prop_default.restype = _vp
        L.sun_update_dz.argtypes = [_vp, _vp, _vp, C.c_int]
        L.sun_set_flux_height.argtypes = [_vp, _vp, _vp]
        L.sun_phys_clone.restype = _vp
        L.sun_phys_clone.argtypes = [_vp, _vp]
        L.sun_phys_copy.argtypes = [_vp, _vp, _vp]
        L.sun_grid_free.argtypes = [_vp]
        L.sun_phys_free.argtypes = [_vp, _vp]
        L.sun_prop_free.argtypes = [_vp]
        for f in ("sun_field_len",):
            getattr(L, f).restype = C.c_long
            getattr(L, f).argtypes = [_vp, C.c_char_p]
        L.sun_field_get.restype = C.c_long
        L.sun_field_get.argtypes = [_vp, _vp, C.c_char_p, _dp]
        L.sun_field_set.restype = C.c_long
        L.sun_field_set.argtypes = [_vp, _vp, C.c_char_p, _dp]
        L.sun_grid_int.restype = C.c_long
        L.sun_grid_int.argtypes = [_vp, C.c_char_p, _ip]
        L.sun_grid_set_int.restype = C.c_long
        L.sun_grid_set_int.argtypes = [_vp, C.c_char_p, _ip]
        L.sun_grid_scalar.restype = C.c_long
        L.sun_grid_scalar.argtypes = [_vp, C.c_char_p]
        L.sun_prop_set.argtypes = [_vp, C.c_char_p, C.c_double]
        L.sun_prop_get.restype = C.c_double
        L.sun_prop_get.argtypes = [_vp, C.c_char_p]
        L.sun_compute_qsource.argtypes = [_vp, _vp, _vp, _vp]
        _state = L
    return _state


def dropin_lib() -> C.CDLL:
    """The drop-in solver library (loads libsunhip.so and the HIP runtime)."""
    global _drop
    if _drop is None:
        _load("libsunhip.so")  # fail loudly with the right name first
        L = _load("libsuntans.so")
        L.CGSolveQ.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.CGSolve.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.OperatorQ.argtypes = [_vp] * 7
        L.Preconditioner.argtypes = [_vp] * 6
        L.OperatorH.argtypes = [_vp] * 7
        L.HorizontalFaceScalars.argtypes = [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.ComputeQSource.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int]
        L.Corrector.argtypes = [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.Continuity.argtypes = [_vp] * 4
        L.UpdateDZ.argtypes = [_vp, _vp, _vp, C.c_int]
        L.SetFluxHeight.argtypes = [_vp, _vp, _vp]
        L.sunhip_UPredictor_begin.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.sunhip_UPredictor_solve.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.sunhip_WPredictor.argtypes = [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int]
        L.UpdateScalars.argtypes = ([_vp] * 7 + [C.c_double, C.c_double, _vp, C.c_double]
                                    + [_vp] * 4 + [C.c_int] * 6)
        L.sunhip_last_solve.argtypes = [C.c_int, _ip, _dp, _dp]
        L.sunhip_last_solve.restype = C.c_int
        L.sunhip_dropin_reset.argtypes = []
        L.sunhip_dropin_set_mode.argtypes = [C.c_int]
        L.sunhip_dropin_set_precision.argtypes = [C.c_int, C.c_double]
        _drop = L
    return _drop
