"""Python view of the thin device ABI (include/sunhip.h) in libsunhip.so.

:class:`DeviceGrid` owns one grid context on the GPU; fields move in and out
as flat numpy arrays.  Used by bench.py, the batched/serving forms and the
tests; the reference-signature drop-ins (libsuntans.so) sit on the same ABI.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import native

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class Topology(C.Structure):
    _fields_ = [("Nc", C.c_int), ("Ne", C.c_int), ("maxfaces", C.c_int), ("Nkmax", C.c_int),
                ("nfaces", _ip), ("neigh", _ip), ("face", _ip), ("normal", _ip), ("grad", _ip),
                ("mark", _ip), ("Nk", _ip), ("Nkc", _ip), ("cellp", _ip), ("celldist", _ip),
                ("edgep", _ip), ("edgedist", _ip), ("Ac", _dp), ("df", _dp), ("dg", _dp),
                ("xv", _dp), ("yv", _dp)]


class Extents(C.Structure):
    _fields_ = [("ctop", _ip), ("ctopold", _ip), ("Nke", _ip), ("etop", _ip)]


class CGParams(C.Structure):
    _fields_ = [("maxiters", C.c_int), ("precond", C.c_int), ("resnorm", C.c_int),
                ("epsilon", C.c_double), ("check_every", C.c_int)]


class CGStats(C.Structure):
    _fields_ = [("niters", C.c_int), ("converged", C.c_int), ("resid", C.c_double), ("eps0", C.c_double),
                ("device_ms", C.c_double), ("setup_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScalarParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("dt", C.c_double), ("kappa", C.c_double), ("TVD", C.c_int),
                ("step", C.c_int), ("alpha_top", C.c_int), ("alpha_bot", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = native._load("libsunhip.so")
        vp = C.c_void_p
        L.sunhip_error.restype = C.c_char_p
        L.sunhip_grid_create.argtypes = [C.POINTER(vp), C.POINTER(Topology), C.c_int]
        L.sunhip_grid_destroy.argtypes = [vp]
        L.sunhip_set_extents.argtypes = [vp, C.POINTER(Extents)]
        L.sunhip_grid_bytes.argtypes = [vp]
        L.sunhip_grid_bytes.restype = C.c_int64
        L.sunhip_grid_sizes.argtypes = [vp, C.POINTER(C.c_int64)]
        L.sunhip_upload.argtypes = [vp, C.c_char_p, _dp]
        L.sunhip_download.argtypes = [vp, C.c_char_p, _dp]
        L.sunhip_copy_field.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_q_solve.argtypes = [vp, C.POINTER(CGParams), C.POINTER(CGStats)]
        L.sunhip_q_solve_fixed.argtypes = [vp, C.POINTER(CGParams), C.c_int, C.POINTER(CGStats)]
        L.sunhip_q_operator.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_q_precondition.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_h_solve.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_double, C.c_double,
                                     C.POINTER(CGStats)]
        L.sunhip_face_scalars.argtypes = [vp, C.POINTER(ScalarParams), C.c_char_p]
        L.sunhip_update_scalars.argtypes = [vp, C.POINTER(ScalarParams), C.c_char_p]
        L.sunhip_sync.argtypes = [vp]
        L.sunhip_profile.argtypes = [vp, C.c_int]
        L.sunhip_kernel_times.argtypes = [vp, _dp, _ip]
        L.sunhip_kernel_times.restype = C.c_int
        L.sunhip_copy_probe.argtypes = [vp, C.c_int64, C.c_int, _dp]
        L.sunhip_q_solve_mixed.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_int,
                                           C.POINTER(CGStats), _ip]
        L.sunhip_h_solve_mixed.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, C.POINTER(CGStats), _ip]
        L.sunhip_set_tuning.argtypes = [vp, C.c_char_p, C.c_int]
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().sunhip_error().decode()} (rc={rc})")


def device_count() -> int:
    return int(lib().sunhip_device_count())


class DeviceGrid:
    """One grid on one GPU (topology from a :class:`State`)."""

    def __init__(self, st, device: int = 0):
        L = lib()
        self._keep = {}
        ints = {n: np.ascontiguousarray(st.ints(n), np.int32) for n in
                ("nfaces", "neigh", "face", "normal", "grad", "mark", "Nk", "Nkc", "cellp", "celldist",
                 "edgep", "edgedist")}
        reals = {n: np.ascontiguousarray(st.get(n)) for n in ("Ac", "df", "dg", "xv", "yv")}
        T = Topology(st.Nc, st.Ne, st.scalar("maxfaces"), st.scalar("Nkmax"))
        for n, a in ints.items():
            setattr(T, n, native.iptr(a))
        for n, a in reals.items():
            setattr(T, n, native.dptr(a))
        self.h = C.c_void_p()
        check(L.sunhip_grid_create(C.byref(self.h), C.byref(T), device), "sunhip_grid_create")
        self.st = st
        self.set_extents(st)

    def close(self):
        if getattr(self, "h", None):
            lib().sunhip_grid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_extents(self, st=None):
        st = st or self.st
        a = {n: np.ascontiguousarray(st.ints(n), np.int32) for n in ("ctop", "ctopold", "Nke", "etop")}
        X = Extents(*(native.iptr(a[n]) for n in ("ctop", "ctopold", "Nke", "etop")))
        check(lib().sunhip_set_extents(self.h, C.byref(X)), "sunhip_set_extents")

    def sizes(self):
        o = (C.c_int64 * 6)()
        lib().sunhip_grid_sizes(self.h, o)
        return dict(zip(("Nflat", "NflatW", "NflatE", "NT", "Nt", "NE"), list(o)))

    def bytes(self) -> int:
        return int(lib().sunhip_grid_bytes(self.h))

    def upload(self, name: str, arr) -> None:
        a = np.ascontiguousarray(arr, np.float64)
        check(lib().sunhip_upload(self.h, name.encode(), native.dptr(a)), f"upload {name}")

    def download(self, name: str, n: int) -> np.ndarray:
        out = np.empty(n, np.float64)
        check(lib().sunhip_download(self.h, name.encode(), native.dptr(out)), f"download {name}")
        return out

    def copy(self, dst: str, src: str) -> None:
        check(lib().sunhip_copy_field(self.h, dst.encode(), src.encode()), f"copy {src}->{dst}")

    def sync(self):
        check(lib().sunhip_sync(self.h), "sync")

    def q_solve(self, maxiters=2000, precond=2, resnorm=0, epsilon=1e-10, check_every=0, fixed=None):
        P = CGParams(maxiters, precond, resnorm, epsilon, check_every)
        S = CGStats()
        if fixed is None:
            check(lib().sunhip_q_solve(self.h, C.byref(P), C.byref(S)), "sunhip_q_solve")
        else:
            check(lib().sunhip_q_solve_fixed(self.h, C.byref(P), int(fixed), C.byref(S)), "sunhip_q_solve_fixed")
        return S.as_dict()

    def q_solve_mixed(self, epsilon=1e-10, inner_tol=1e-5, max_outer=20, maxiters=2000, check_every=0):
        """FP32 inner PCG in FP64 iterative refinement; returns stats with
        niters = refinement steps and inner_iters = FP32 PCG iterations."""
        P = CGParams(maxiters, 2, 0, epsilon, check_every)
        S = CGStats()
        n = C.c_int()
        check(lib().sunhip_q_solve_mixed(self.h, C.byref(P), inner_tol, max_outer, C.byref(S), C.byref(n)),
              "sunhip_q_solve_mixed")
        d = S.as_dict()
        d["inner_iters"] = n.value
        return d

    def tune(self, key: str, value: int) -> None:
        check(lib().sunhip_set_tuning(self.h, key.encode(), int(value)), f"tune {key}")

    def h_solve(self, grav, theta, dt, maxiters=1000, precond=1, resnorm=0, epsilon=1e-10, check_every=0):
        P = CGParams(maxiters, precond, resnorm, epsilon, check_every)
        S = CGStats()
        check(lib().sunhip_h_solve(self.h, C.byref(P), grav, theta, dt, C.byref(S)), "sunhip_h_solve")
        return S.as_dict()

    def h_solve_mixed(self, grav, theta, dt, epsilon=1e-10, inner_tol=1e-5, max_outer=20, maxiters=1000,
                      precond=1, check_every=0):
        """FP32 inner free-surface PCG in FP64 iterative refinement."""
        P = CGParams(maxiters, precond, 0, epsilon, check_every)
        S = CGStats()
        n = C.c_int()
        check(lib().sunhip_h_solve_mixed(self.h, C.byref(P), grav, theta, dt, inner_tol, max_outer, C.byref(S),
                                         C.byref(n)), "sunhip_h_solve_mixed")
        d = S.as_dict()
        d["inner_iters"] = n.value
        return d

    def update_scalars(self, theta, dt, kappa, TVD, step=2, scal="s"):
        S = ScalarParams(theta, dt, kappa, TVD, step, 0, 0)
        check(lib().sunhip_update_scalars(self.h, C.byref(S), scal.encode()), "sunhip_update_scalars")

    def profile(self, on: bool):
        lib().sunhip_profile(self.h, int(on))

    def kernel_times(self):
        """{kernel: (total ms, launches)} accumulated since profiling was enabled."""
        ms = (C.c_double * 8)()
        n = (C.c_int * 8)()
        k = lib().sunhip_kernel_times(self.h, ms, n)
        names = ("qapply", "qprec", "happly", "hprec", "mixed_apply", "mixed_prec", "resid64", "other")
        return {names[i]: (ms[i], n[i]) for i in range(min(k, 8)) if n[i] > 0}

    def copy_probe(self, nbytes: int, reps: int):
        """Device copy kernel of nbytes, reps times; returns average ms."""
        out = C.c_double()
        check(lib().sunhip_copy_probe(self.h, int(nbytes), int(reps), C.byref(out)), "copy_probe")
        return out.value
