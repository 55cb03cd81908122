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
                ("epsilon", C.c_double), ("check_every", C.c_int), ("qc_alias", C.c_int)]


class CGStats(C.Structure):
    _fields_ = [("niters", C.c_int), ("converged", C.c_int), ("resid", C.c_double), ("eps0", C.c_double),
                ("device_ms", C.c_double), ("setup_ms", C.c_double), ("outer", C.c_int), ("inner", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ScalarParams(C.Structure):
    _fields_ = [("theta", C.c_double), ("dt", C.c_double), ("kappa", C.c_double), ("TVD", C.c_int),
                ("step", C.c_int), ("alpha_top", C.c_int), ("alpha_bot", C.c_int)]


class VGridParams(C.Structure):
    _fields_ = [("dz", _dp), ("dzsmall", C.c_double), ("smoothbot", C.c_double), ("stairstep", C.c_int),
                ("fixdzz", C.c_int), ("linearFS", C.c_int)]


class UPredParams(C.Structure):
    """sunhip_upred_params: dt, theta, grav, nu, laxWendroff_Vertical,
    thetaM (< 0: no vertical momentum advection), nonlinear."""
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("grav", C.c_double), ("nu", C.c_double),
                ("lax_vertical", C.c_double), ("thetaM", C.c_double), ("nonlinear", C.c_int)]


class WPredParams(C.Structure):
    """sunhip_wpred_params: dt, theta, nu, nu_H, CdW, laxWendroff_Vertical,
    nonlinear (0, 1, 2, 5), step (prop->n), AB, TVDmomentum,
    conserveMomentum (must be 0)."""
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("nu", C.c_double), ("nu_H", C.c_double),
                ("CdW", C.c_double), ("lax_vertical", C.c_double), ("nonlinear", C.c_int), ("step", C.c_int),
                ("AB", C.c_int), ("TVDmomentum", C.c_int), ("conserveMomentum", C.c_int)]

    @classmethod
    def of(cls, st) -> "WPredParams":
        g = st.get_prop
        return cls(g("dt"), g("theta"), g("nu"), g("nu_H"), g("CdW"), g("laxWendroff_Vertical"),
                   int(g("nonlinear")), int(g("n")), int(g("AB")), int(g("TVDmomentum")),
                   int(g("conserveMomentum")))


def face_distances(st) -> np.ndarray:
    """Distance from each edge's first cell centre to its midpoint
    (main/phys.c:4272-4273), for sunhip_set_face_distances."""
    grad = st.ints("grad").reshape(-1, 2)
    nc1 = np.where(grad[:, 0] == -1, grad[:, 1], grad[:, 0])
    dx = st.get("xv")[nc1] - st.get("xe")
    dy = st.get("yv")[nc1] - st.get("ye")
    return np.sqrt(dx * dx + dy * dy)


class HPhys(C.Structure):
    _fields_ = [("grav", C.c_double), ("theta", C.c_double), ("dt", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        L = native._load("libsunhip.so")
        vp = C.c_void_p
        L.sunhip_error.restype = C.c_char_p
        L.sunhip_grid_create.argtypes = [C.POINTER(vp), C.POINTER(Topology), C.c_int]
        L.sunhip_validate.argtypes = [C.POINTER(Topology), C.POINTER(Extents)]
        L.sunhip_set_dz.argtypes = [vp, _dp]
        L.sunhip_grid_destroy.argtypes = [vp]
        L.sunhip_set_extents.argtypes = [vp, C.POINTER(Extents)]
        L.sunhip_grid_bytes.argtypes = [vp]
        L.sunhip_grid_bytes.restype = C.c_int64
        L.sunhip_grid_sizes.argtypes = [vp, C.POINTER(C.c_int64)]
        L.sunhip_upload.argtypes = [vp, C.c_char_p, _dp]
        L.sunhip_download.argtypes = [vp, C.c_char_p, _dp]
        L.sunhip_copy_field.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_field_scale.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_double]
        L.sunhip_q_solve.argtypes = [vp, C.POINTER(CGParams), C.POINTER(CGStats)]
        L.sunhip_q_solve_fixed.argtypes = [vp, C.POINTER(CGParams), C.c_int, C.POINTER(CGStats)]
        L.sunhip_q_operator.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_q_precondition.argtypes = [vp, C.c_char_p, C.c_char_p]
        L.sunhip_h_solve.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_double, C.c_double,
                                     C.POINTER(CGStats)]
        L.sunhip_h_solve_noalloc.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_double, C.c_double,
                                             C.POINTER(CGStats)]
        L.sunhip_face_scalars.argtypes = [vp, C.POINTER(ScalarParams), C.c_char_p]
        L.sunhip_update_scalars.argtypes = [vp, C.POINTER(ScalarParams), C.c_char_p]
        L.sunhip_sync.argtypes = [vp]
        L.sunhip_profile.argtypes = [vp, C.c_int]
        L.sunhip_kernel_times.argtypes = [vp, _dp, _ip]
        L.sunhip_kernel_times.restype = C.c_int
        L.sunhip_copy_probe.argtypes = [vp, C.c_int64, C.c_int, _dp]
        L.sunhip_pcie_probe.argtypes = [vp, C.c_int64, C.c_int, _dp, _dp]
        L.sunhip_update_scalars_src.argtypes = [vp, vp, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                                C.c_char_p, C.c_char_p]
        L.sunhip_q_solve_mixed.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_int,
                                           C.POINTER(CGStats), _ip]
        L.sunhip_h_solve_mixed.argtypes = [vp, C.POINTER(CGParams), C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_int, C.POINTER(CGStats), _ip]
        L.sunhip_set_tuning.argtypes = [vp, C.c_char_p, C.c_int]
        # batched solves
        pp = C.POINTER(vp)
        L.sunhip_batch_create.argtypes = [C.POINTER(vp), C.c_int]
        L.sunhip_batch_destroy.argtypes = [vp]
        L.sunhip_batch_attach.argtypes = [vp, vp]
        L.sunhip_batch_detach.argtypes = [vp, vp]
        L.sunhip_batch_begin_q.argtypes = [vp, vp, C.POINTER(CGParams), C.c_int64]
        L.sunhip_batch_begin_h.argtypes = [vp, vp, C.POINTER(CGParams), C.POINTER(HPhys), C.c_int64]
        L.sunhip_batch_begin_q_mixed.argtypes = [vp, vp, C.POINTER(CGParams), C.c_double, C.c_int, C.c_int64]
        L.sunhip_batch_advance.argtypes = [vp, C.c_int]
        L.sunhip_batch_poll.argtypes = [vp, C.POINTER(C.c_int64), _ip, C.POINTER(CGStats), C.c_int]
        L.sunhip_batch_active.argtypes = [vp]
        L.sunhip_batch_sync.argtypes = [vp]
        L.sunhip_batch_q_solve.argtypes = [vp, pp, C.c_int, C.POINTER(CGParams), C.POINTER(CGStats)]
        L.sunhip_batch_h_solve.argtypes = [vp, pp, C.c_int, C.POINTER(CGParams), C.POINTER(HPhys),
                                           C.POINTER(CGStats)]
        L.sunhip_batch_q_solve_mixed.argtypes = [vp, pp, C.c_int, C.POINTER(CGParams), C.c_double, C.c_int,
                                                 C.POINTER(CGStats), _ip]
        L.sunhip_batch_update_scalars.argtypes = [vp, pp, C.c_int, C.POINTER(ScalarParams), C.c_char_p]
        L.sunhip_u_predict.argtypes = [vp, C.POINTER(UPredParams)]
        L.sunhip_w_predict.argtypes = [vp, C.POINTER(WPredParams)]
        L.sunhip_set_face_distances.argtypes = [vp, vp]
        L.sunhip_update_dz.argtypes = [vp, C.POINTER(VGridParams), C.c_int]
        L.sunhip_set_flux_height.argtypes = [vp, C.POINTER(VGridParams)]
        L.sunhip_get_extents.argtypes = [vp, _ip, _ip, _ip, _ip]
        L.sunhip_u_solve.argtypes = [vp, C.POINTER(UPredParams), C.POINTER(CGParams), C.POINTER(CGStats)]
        L.sunhip_compute_qsource.argtypes = [vp, C.c_double, C.c_double, C.c_char_p]
        L.sunhip_corrector.argtypes = [vp, C.c_double, C.c_char_p, C.c_char_p]
        L.sunhip_continuity.argtypes = [vp, C.c_double, C.c_char_p]
        L.sunhip_batch_compute_qsource.argtypes = [vp, pp, C.c_int, _dp, _dp, C.c_char_p]
        L.sunhip_batch_correct.argtypes = [vp, pp, C.c_int, _dp, _dp, C.c_char_p, C.c_char_p, C.c_char_p]
        L.sunhip_batch_set_tuning.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.sunhip_batch_tuning.argtypes = [vp, _ip, _ip, _ip]
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what}: {lib().sunhip_error().decode()} (rc={rc})")


def device_count() -> int:
    return int(lib().sunhip_device_count())


def topology(st, override=None):
    """The sunhip_topology of a State, with the arrays it points to (keep
    them alive while the structure is used).  override replaces named
    arrays (tests of the input checks)."""
    ints = {n: np.ascontiguousarray(st.ints(n), np.int32) for n in
            ("nfaces", "neigh", "face", "normal", "grad", "mark", "Nk", "Nkc", "cellp", "celldist",
             "edgep", "edgedist")}
    reals = {n: np.ascontiguousarray(st.get(n)) for n in ("Ac", "df", "dg", "xv", "yv")}
    for n, a in (override or {}).items():
        if n in ints:
            ints[n] = np.ascontiguousarray(a, np.int32)
        else:
            reals[n] = np.ascontiguousarray(a, np.float64)
    T = Topology(st.Nc, st.Ne, st.scalar("maxfaces"), st.scalar("Nkmax"))
    for n, a in ints.items():
        setattr(T, n, native.iptr(a))
    for n, a in reals.items():
        setattr(T, n, native.dptr(a))
    return T, (ints, reals)


def extents(st, override=None):
    """The sunhip_extents of a State and the arrays it points to."""
    a = {n: np.ascontiguousarray(st.ints(n), np.int32) for n in ("ctop", "ctopold", "Nke", "etop")}
    for n, v in (override or {}).items():
        a[n] = np.ascontiguousarray(v, np.int32)
    return Extents(*(native.iptr(a[n]) for n in ("ctop", "ctopold", "Nke", "etop"))), a


def validate(st, with_extents=True, override=None):
    """sunhip_validate on a State's topology (and extents): 0 or raises
    RuntimeError with the library's message.  Needs no device."""
    override = override or {}
    T, keep = topology(st, {k: v for k, v in override.items() if k not in ("ctop", "ctopold", "Nke", "etop")})
    X, keepx = extents(st, {k: v for k, v in override.items() if k in ("ctop", "ctopold", "Nke", "etop")})
    check(lib().sunhip_validate(C.byref(T), C.byref(X) if with_extents else None), "sunhip_validate")
    return 0


class DeviceGrid:
    """One grid on one GPU (topology from a :class:`State`)."""

    def __init__(self, st, device: int = 0):
        L = lib()
        self._keep = {}
        T, keep = topology(st)
        self.h = C.c_void_p()
        check(L.sunhip_grid_create(C.byref(self.h), C.byref(T), device), "sunhip_grid_create")
        self.Nc, self.Ne = st.Nc, st.Ne
        self.st = st
        self.set_extents(st)
        self._dz = np.ascontiguousarray(st.get("dz"), np.float64)  # grid->dz: the z-level windows
        check(L.sunhip_set_dz(self.h, native.dptr(self._dz)), "sunhip_set_dz")

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
        X, keep = extents(st)
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

    def scale(self, dst: str, src: str, a: float) -> None:
        check(lib().sunhip_field_scale(self.h, dst.encode(), src.encode(), a), f"scale {src}->{dst}")

    def sync(self):
        check(lib().sunhip_sync(self.h), "sync")

    def q_solve(self, maxiters=2000, precond=2, resnorm=0, epsilon=1e-10, check_every=0, fixed=None, qc_alias=0):
        """CGSolveQ on the resident fields.  precond 1: qc_alias = 1 divides
        the scaled solution by the final residual at the end, as the
        reference does when its c aliases r (Solve's call)."""
        P = CGParams(maxiters, precond, resnorm, epsilon, check_every, qc_alias)
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

    def h_solve_noalloc(self, grav, theta, dt, maxiters=1000, precond=1, resnorm=0, epsilon=1e-10, check_every=0):
        """sunhip_h_solve with its work arrays in the library's static
        scratch (allocates no device memory)."""
        P = CGParams(maxiters, precond, resnorm, epsilon, check_every)
        S = CGStats()
        check(lib().sunhip_h_solve_noalloc(self.h, C.byref(P), grav, theta, dt, C.byref(S)),
              "sunhip_h_solve_noalloc")
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

    def update_scalars(self, theta, dt, kappa, TVD, step=2, scal="s", src1=None, src2=None, ftop=None, fbot=None,
                       alpha_top=0, alpha_bot=0, bcell=None):
        """One UpdateScalars step of field scal; src1/src2 (cell fields) and
        ftop/fbot (2-D cell fields) name the source terms, None leaves one
        out; bcell names the open-boundary rows of first-order upwind."""
        S = ScalarParams(theta, dt, kappa, TVD, step, alpha_top, alpha_bot)
        enc = lambda n: None if n is None else n.encode()
        check(lib().sunhip_update_scalars_src(self.h, C.byref(S), scal.encode(), enc(src1), enc(src2), enc(ftop),
                                              enc(fbot), enc(bcell)), "sunhip_update_scalars")

    def _vgrid(self, st):
        self._dz = np.ascontiguousarray(st.get("dz"), np.float64)
        return VGridParams(self._dz.ctypes.data_as(_dp), st.grid_param("dzsmall"), st.grid_param("smoothbot"),
                           int(st.grid_param("stairstep")), int(st.grid_param("fixdzz")),
                           int(st.get_prop("linearFS")))

    def update_dz(self, st, option: int = 0):
        """UpdateDZ on the device from the resident h (vertical-grid
        parameters and dz from the State)."""
        V = self._vgrid(st)
        check(lib().sunhip_update_dz(self.h, C.byref(V), int(option)), "sunhip_update_dz")

    def set_flux_height(self, st):
        V = self._vgrid(st)
        check(lib().sunhip_set_flux_height(self.h, C.byref(V)), "sunhip_set_flux_height")

    def extents(self):
        """The device's ctop, ctopold, etop, etopold in reference order."""
        nc, ne = self.Nc, self.Ne
        out = {k: np.zeros(n, np.int32) for k, n in (("ctop", nc), ("ctopold", nc), ("etop", ne), ("etopold", ne))}
        check(lib().sunhip_get_extents(self.h, *[out[k].ctypes.data_as(_ip) for k in ("ctop", "ctopold", "etop",
                                                                                        "etopold")]),
              "sunhip_get_extents")
        return out

    def u_predict(self, P: "UPredParams"):
        """UPredictor before the open-boundary hooks: U*** in utmp, E in ue, D."""
        check(lib().sunhip_u_predict(self.h, C.byref(P)), "sunhip_u_predict")

    def u_solve(self, P: "UPredParams", maxiters=1000, precond=1, resnorm=0, epsilon=1e-10):
        """UPredictor after the hooks: htmp, free-surface PCG, drying, u."""
        prm = CGParams(maxiters, precond, resnorm, epsilon, 0)
        S = CGStats()
        check(lib().sunhip_u_solve(self.h, C.byref(P), C.byref(prm), C.byref(S)), "sunhip_u_solve")
        return S.as_dict()

    def set_face_distances(self, def1):
        d = np.ascontiguousarray(def1, np.float64)
        check(lib().sunhip_set_face_distances(self.h, d.ctypes.data_as(C.c_void_p)), "sunhip_set_face_distances")

    def w_predict(self, P: "WPredParams"):
        """WPredictor: w, wtmp, wc, stmp, Cn_W, Cn_W2 and the face values in ue."""
        check(lib().sunhip_w_predict(self.h, C.byref(P)), "sunhip_w_predict")

    def compute_qsource(self, theta, dt, src="src"):
        """ComputeQSource: src from u, utmp2, w, wtmp2, dzf; D <- df/dg."""
        check(lib().sunhip_compute_qsource(self.h, theta, dt, src.encode()), "sunhip_compute_qsource")

    def corrector(self, dt, qc="q", q="qp"):
        """Corrector: u, w projected with the correction qc; q += qc."""
        check(lib().sunhip_corrector(self.h, dt, qc.encode(), q.encode()), "sunhip_corrector")

    def continuity(self, theta, w="w"):
        """Continuity: w from u, utmp2, wtmp2, dzf."""
        check(lib().sunhip_continuity(self.h, theta, w.encode()), "sunhip_continuity")

    def profile(self, on: bool):
        lib().sunhip_profile(self.h, int(on))

    def kernel_times(self):
        """{kernel: (total ms, launches)} accumulated since profiling was enabled."""
        ms = (C.c_double * 16)()
        n = (C.c_int * 16)()
        k = lib().sunhip_kernel_times(self.h, ms, n)
        names = ("qapply", "qprec", "happly", "hprec", "mixed_apply", "mixed_prec", "resid64", "other",
                 "tvd_sums", "tvd_faces", "tvd_column", "qsource", "corrector", "continuity", "wp_source",
                 "wp_column")
        return {names[i]: (ms[i], n[i]) for i in range(min(k, len(names))) if n[i] > 0}

    def copy_probe(self, nbytes: int, reps: int):
        """Device copy kernel of nbytes, reps times; returns average ms."""
        out = C.c_double()
        check(lib().sunhip_copy_probe(self.h, int(nbytes), int(reps), C.byref(out)), "copy_probe")
        return out.value

    def pcie_probe(self, nbytes: int, reps: int):
        """Pinned host <-> HBM copies of nbytes: (h2d ms, d2h ms) per copy."""
        a, b = C.c_double(), C.c_double()
        check(lib().sunhip_pcie_probe(self.h, int(nbytes), int(reps), C.byref(a), C.byref(b)), "pcie_probe")
        return a.value, b.value


class Batch:
    """Many grids solved together (sunhip_batch_*): attached grids enqueue on
    the batch's stream and the PCG iterations of all active solves share the
    same launches."""

    KIND = {0: "q", 1: "h", 2: "q32", 3: "qmix"}

    def __init__(self, device: int = 0):
        self.h = C.c_void_p()
        check(lib().sunhip_batch_create(C.byref(self.h), device), "sunhip_batch_create")

    def close(self):
        if getattr(self, "h", None):
            lib().sunhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _grids(dgs):
        arr = (C.c_void_p * len(dgs))(*[d.h.value for d in dgs])
        return arr

    def attach(self, dg):
        check(lib().sunhip_batch_attach(self.h, dg.h), "sunhip_batch_attach")

    def detach(self, dg):
        check(lib().sunhip_batch_detach(self.h, dg.h), "sunhip_batch_detach")

    def begin_q(self, dg, tag, epsilon=1e-10, maxiters=2000, resnorm=0):
        P = CGParams(maxiters, 2, resnorm, epsilon, 0)
        check(lib().sunhip_batch_begin_q(self.h, dg.h, C.byref(P), tag), "sunhip_batch_begin_q")

    def begin_q_mixed(self, dg, tag, epsilon=1e-10, inner_tol=1e-4, max_outer=20, maxiters=2000):
        P = CGParams(maxiters, 2, 0, epsilon, 0)
        check(lib().sunhip_batch_begin_q_mixed(self.h, dg.h, C.byref(P), inner_tol, max_outer, tag),
              "sunhip_batch_begin_q_mixed")

    def begin_h(self, dg, tag, grav, theta, dt, epsilon=1e-10, maxiters=1000, precond=1, resnorm=0):
        P = CGParams(maxiters, precond, resnorm, epsilon, 0)
        ph = HPhys(grav, theta, dt)
        check(lib().sunhip_batch_begin_h(self.h, dg.h, C.byref(P), C.byref(ph), tag), "sunhip_batch_begin_h")

    def advance(self, iters: int):
        check(lib().sunhip_batch_advance(self.h, iters), "sunhip_batch_advance")

    def poll(self, maxn: int = 4096):
        tags = (C.c_int64 * maxn)()
        kinds = (C.c_int * maxn)()
        st = (CGStats * maxn)()
        n = lib().sunhip_batch_poll(self.h, tags, kinds, st, maxn)
        if n < 0:
            check(n, "sunhip_batch_poll")
        return [(int(tags[i]), self.KIND[int(kinds[i])], st[i].as_dict()) for i in range(n)]

    def active(self) -> int:
        return int(lib().sunhip_batch_active(self.h))

    def sync(self):
        check(lib().sunhip_batch_sync(self.h), "sunhip_batch_sync")

    def q_solve(self, dgs, epsilon=1e-10, maxiters=2000, resnorm=0):
        P = CGParams(maxiters, 2, resnorm, epsilon, 0)
        st = (CGStats * len(dgs))()
        check(lib().sunhip_batch_q_solve(self.h, self._grids(dgs), len(dgs), C.byref(P), st),
              "sunhip_batch_q_solve")
        return [s.as_dict() for s in st]

    def h_solve(self, dgs, phys, epsilon=1e-10, maxiters=1000, precond=1, resnorm=0):
        P = CGParams(maxiters, precond, resnorm, epsilon, 0)
        ph = (HPhys * len(dgs))(*[HPhys(*p) for p in phys])
        st = (CGStats * len(dgs))()
        check(lib().sunhip_batch_h_solve(self.h, self._grids(dgs), len(dgs), C.byref(P), ph, st),
              "sunhip_batch_h_solve")
        return [s.as_dict() for s in st]

    def q_solve_mixed(self, dgs, epsilon=1e-10, inner_tol=1e-4, max_outer=20, maxiters=2000):
        P = CGParams(maxiters, 2, 0, epsilon, 0)
        st = (CGStats * len(dgs))()
        inner = (C.c_int * len(dgs))()
        check(lib().sunhip_batch_q_solve_mixed(self.h, self._grids(dgs), len(dgs), C.byref(P), inner_tol,
                                               max_outer, st, inner), "sunhip_batch_q_solve_mixed")
        out = [s.as_dict() for s in st]
        for d, n in zip(out, inner):
            d["inner_iters"] = int(n)
        return out

    def update_scalars(self, dgs, params, scal="s"):
        """params: one (theta, dt, kappa, TVD, step) per grid."""
        sp = (ScalarParams * len(dgs))(*[ScalarParams(th, dt, ka, tvd, step, 0, 0)
                                         for th, dt, ka, tvd, step in params])
        check(lib().sunhip_batch_update_scalars(self.h, self._grids(dgs), len(dgs), sp, scal.encode()),
              "sunhip_batch_update_scalars")

    def compute_qsource(self, dgs, theta, dt, src="src"):
        """ComputeQSource of every grid in one launch; theta/dt per grid or scalars."""
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float64), (len(dgs),)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(dgs),)))
        check(lib().sunhip_batch_compute_qsource(self.h, self._grids(dgs), len(dgs), th.ctypes.data_as(_dp),
                                                 d.ctypes.data_as(_dp), src.encode()),
              "sunhip_batch_compute_qsource")

    def correct(self, dgs, theta, dt, qc="q", q="qp", w="w"):
        """Corrector then Continuity of every grid, one launch each."""
        th = np.ascontiguousarray(np.broadcast_to(np.asarray(theta, np.float64), (len(dgs),)))
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (len(dgs),)))
        check(lib().sunhip_batch_correct(self.h, self._grids(dgs), len(dgs), th.ctypes.data_as(_dp),
                                         d.ctypes.data_as(_dp), qc.encode(), q.encode(), w.encode()),
              "sunhip_batch_correct")

    def set_tuning(self, unroll=0, blocks_per_cu=0, chunk=0):
        check(lib().sunhip_batch_set_tuning(self.h, unroll, blocks_per_cu, chunk), "sunhip_batch_set_tuning")

    def tuning(self):
        u, b, c = C.c_int(), C.c_int(), C.c_int()
        n = lib().sunhip_batch_tuning(self.h, C.byref(u), C.byref(b), C.byref(c))
        return {"unroll": u.value, "blocks_per_cu": b.value, "chunk": c.value, "tuned_shapes": n}


def member_upload(dg, st, which=("q", "src", "dzz", "dzf", "D")):
    """Upload a State's pressure-solve inputs to its DeviceGrid."""
    names = {"q": "qc", "src": "stmp", "dzz": "dzz", "dzf": "dzf", "D": "D", "h": "h", "htmp": "htmp"}
    for n in which:
        dg.upload(n, st.get(names[n]))


def member_upload_step(dg, st):
    """Upload a State's inputs of the stencils either side of the pressure
    solve (ComputeQSource, Corrector, Continuity): velocities, flux heights,
    layer thicknesses, D and the nonhydrostatic pressure (as "qp")."""
    for n in ("u", "utmp2", "w", "wtmp2", "dzf", "dzz", "D"):
        dg.upload(n, st.get(n))
    dg.upload("qp", st.get("q"))


def member_upload_tvd(dg, st):
    """Upload a State's scalar-transport inputs (UpdateScalars) to its
    DeviceGrid; every cell active, no open-boundary scalar."""
    for n in ("s", "u", "utmp2", "dzf", "wnew", "wtmp2", "dzz", "dzzold", "kappa_tv"):
        dg.upload(n, st.get(n))
    dg.upload("active", np.ones(st.Nc))
