"""Uniform access to the three implementations of the path for the tests.

* ``oracle()``  -- the FP64 C restatement (oracle/_build/liboracle.so), the
                   checker;
* ``reference()`` -- the real reference compiled from its sources
                   (oracle/_ref/libsunref.so), present only where the
                   reference tree was available at build time;
* ``dropin()``  -- the product: libsuntans.so, whose functions have the
                   reference's signatures and run the HIP kernels.

Each adapter exposes the same methods, taking a :class:`State`.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import tempfile

from suntans_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsunref.so")

_vp = C.c_void_p


class _Cpu:
    def __init__(self, path, prefix):
        self.L = C.CDLL(path)
        self.pre = prefix
        for name, args in {
            "CGSolveQ": [_vp] * 6, "CGSolve": [_vp] * 3, "QCoefficients": [_vp] * 6,
            "OperatorQ": [_vp] * 7, "Preconditioner": [_vp] * 6, "HCoefficients": [_vp] * 5,
            "OperatorH": [_vp] * 7, "ComputeQSource": [_vp] * 4, "ConditionQ": [_vp] * 4,
            "OperatorQC": [_vp] * 8, "Corrector": [_vp] * 4, "Continuity": [_vp] * 4,
            "HorizontalFaceScalars": [_vp] * 5 + [C.c_int],
            "TriSolve": [_vp] * 5 + [C.c_int],
        }.items():
            f = self.fn(name)
            f.argtypes = args
        self.fn("InnerProduct3").argtypes = [_vp] * 3
        self.fn("InnerProduct3").restype = C.c_double

    def fn(self, name):
        return getattr(self.L, self.pre + name)

    # calls mirror the reference call sites
    def cgsolveq(self, st, c="stmp3"):
        """c: the conditioning rows (Solve passes stmp3, the residual rows)."""
        self.fn("CGSolveQ")(st.ptr("qc"), st.ptr("stmp"), st.ptr(c), st.g, st.p, st.prop)

    def cgsolve(self, st):
        self.fn("CGSolve")(st.g, st.p, st.prop)

    def qcoef(self, st):
        self.fn("QCoefficients")(st.ptr("wtmp"), st.ptr("qtmp"), st.ptr("stmp3"), st.g, st.p, st.prop)

    def operatorq(self, st, x="stmp", y="stmp2"):
        """y = OperatorQ(x) with coefficients in wtmp."""
        self.fn("OperatorQ")(st.ptr("wtmp"), st.ptr(x), st.ptr(y), st.ptr("stmp3"), st.g, st.p, st.prop)

    def preconditioner(self, st, x="stmp", xc="uold"):
        self.fn("Preconditioner")(st.ptr(x), st.ptr(xc), st.ptr("wtmp"), st.g, st.p, st.prop)

    def hcoef(self, st):
        self.fn("HCoefficients")(st.ptr("hcoef"), st.ptr("hfcoef"), st.g, st.p, st.prop)

    def operatorh(self, st, x="htmp2", y="htmp3"):
        self.fn("OperatorH")(st.ptr(x), st.ptr(y), st.ptr("hcoef"), st.ptr("hfcoef"), st.g, st.p, st.prop)

    def compute_qsource(self, st, src="stmp"):
        self.fn("ComputeQSource")(st.ptr(src), st.g, st.p, st.prop)

    def corrector(self, st, qc="qc"):
        self.fn("Corrector")(st.ptr(qc), st.g, st.p, st.prop)

    def continuity(self, st, w="w"):
        self.fn("Continuity")(st.ptr(w), st.g, st.p, st.prop)

    def face_scalars(self, st, tvd):
        self.fn("HorizontalFaceScalars")(st.g, st.p, st.prop, st.ptr("s"), st.ptr("boundary_s"), int(tvd))

    def update_dz(self, st, option):
        getattr(self.L, self.pre + "UpdateDZ")(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop),
                                               C.c_int(option))

    def set_flux_height(self, st):
        getattr(self.L, self.pre + "SetFluxHeight")(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def upredictor_begin(self, st):
        return self.L.oracle_UPredictor_begin(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def upredictor_solve(self, st):
        return self.L.oracle_UPredictor_solve(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def upredictor_rhs(self, st):
        self.L.oracle_UPredictor_rhs(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def upredictor_finish(self, st):
        self.L.oracle_UPredictor_finish(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def upredictor(self, st):
        """The whole of UPredictor (reference), or its two device-covered
        parts in order (oracle); returns nonzero where the reference exits."""
        if self.pre == "ref_":
            self.L.ref_UPredictor(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))
            return 0
        rc = self.L.oracle_UPredictor_begin(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))
        return rc or self.L.oracle_UPredictor_solve(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def wpredictor(self, st):
        getattr(self.L, self.pre + "WPredictor")(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def momentum_faces(self, st):
        """Face values of wc into ut, as WPredictor's call makes them."""
        if self.pre == "ref_":
            self.L.ref_GetMomentumFaceValues(C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))
        else:
            self.L.oracle_MomentumFaceValues(*(C.c_void_p(st.ptr(f)) for f in ("ut", "wc", "boundary_w", "utmp2")),
                                             C.c_void_p(st.g), C.c_void_p(st.p), C.c_void_p(st.prop))

    def update_salinity(self, st):
        getattr(self.L, self.pre + "update_salinity")(C.c_void_p(st.g), C.c_void_p(st.p),
                                                      C.c_void_p(st.prop))

    def update_salinity_src(self, st, flags, alpha_top=0, alpha_bot=0):
        """UpdateScalars on salinity with source terms: flags 1 src1 = uold,
        2 src2 = wtmp, 4 Ftop = htmp, 8 Fbot = htmp2."""
        f = getattr(self.L, self.pre + "update_salinity_src")
        f.argtypes = [_vp] * 3 + [C.c_int] * 3
        f(st.g, st.p, st.prop, int(flags), int(alpha_top), int(alpha_bot))

    def inner3(self, st, x, y):
        return self.fn("InnerProduct3")(st.ptr(x), st.ptr(y), st.g)


class OracleImpl(_Cpu):
    name = "oracle"

    def __init__(self):
        super().__init__(ORACLE_SO, "oracle_")
        self.L.oracle_update_salinity.argtypes = [_vp] * 3

    class _Stats(C.Structure):
        _fields_ = [("niters", C.c_int), ("resid", C.c_double), ("eps0", C.c_double)]

    def stats(self, which):
        s = self._Stats.in_dll(self.L, "oracle_last_q" if which == "q" else "oracle_last_h")
        return s.niters, s.resid


class RefImpl(_Cpu):
    name = "reference"

    def __init__(self):
        super().__init__(REF_SO, "ref_")
        self.L.ref_set_verbose.argtypes = [C.c_int]
        self.L.ref_update_salinity.argtypes = [_vp] * 3
        self._last = {}
        self._hist = {}
        self.history = False  # also capture the per-iteration residuals
        self.call_s = 0.0     # wall time of the last solve call alone

    def _captured(self, fn, which):
        """Run fn with the reference's convergence report captured from stdout
        (it prints the iteration count at VERBOSE>2, main/phys.c:2527-2536,
        and every iteration's residual at VERBOSE>3, :2522 and :3200)."""
        import sys
        import time
        sys.stdout.flush()
        self.L.ref_set_verbose(4 if self.history else 3)
        libc = C.CDLL(None)
        fd = os.dup(1)
        with tempfile.TemporaryFile(mode="w+b") as tmp:
            os.dup2(tmp.fileno(), 1)
            try:
                t = time.perf_counter()
                fn()
                self.call_s = time.perf_counter() - t
                libc.fflush(None)
            finally:
                os.dup2(fd, 1)
                os.close(fd)
                self.L.ref_set_verbose(0)
            tmp.seek(0)
            out = tmp.read().decode()
        m = re.search(r"after (\d+) (?:iterations|steps)[^=]*(?:res|RES)=([-+0-9.eE]+|nan|inf)", out)
        self._last[which] = (int(m.group(1)), float(m.group(2))) if m else (None, None)
        self._hist[which] = [float(v) for v in re.findall(r"Iteration: \d+, resid=([-+0-9.eE]+|nan|inf)", out)]
        return out

    def cgsolveq(self, st, c="stmp3"):
        self._captured(lambda: super(RefImpl, self).cgsolveq(st, c), "q")

    def cgsolve(self, st):
        self._captured(lambda: super(RefImpl, self).cgsolve(st), "h")

    def stats(self, which):
        return self._last.get(which, (None, None))

    def residuals(self, which):
        """Residuals sqrt(eps/eps0) the last solve printed after each
        iteration (history mode only)."""
        return list(self._hist.get(which, []))


class DropinImpl:
    """The product, through the reference signatures."""
    name = "dropin"

    def __init__(self):
        self.L = native.dropin_lib()

    def cgsolveq(self, st, c="stmp3"):
        self.L.CGSolveQ(st.ptr("qc"), st.ptr("stmp"), st.ptr(c), st.g, st.p, st.prop, 0, 1, 0)

    def cgsolve(self, st):
        self.L.CGSolve(st.g, st.p, st.prop, 0, 1, 0)

    def operatorq(self, st, x="stmp", y="stmp2"):
        self.L.OperatorQ(st.ptr("wtmp"), st.ptr(x), st.ptr(y), st.ptr("stmp3"), st.g, st.p, st.prop)

    def preconditioner(self, st, x="stmp", xc="uold"):
        self.L.Preconditioner(st.ptr(x), st.ptr(xc), st.ptr("wtmp"), st.g, st.p, st.prop)

    def operatorh(self, st, x="htmp2", y="htmp3"):
        self.L.OperatorH(st.ptr(x), st.ptr(y), st.ptr("hcoef"), st.ptr("hfcoef"), st.g, st.p, st.prop)

    def compute_qsource(self, st, src="stmp"):
        self.L.ComputeQSource(st.ptr(src), st.g, st.p, st.prop, 0, 1)

    def corrector(self, st, qc="qc"):
        self.L.Corrector(st.ptr(qc), st.g, st.p, st.prop, 0, 1, 0)

    def continuity(self, st, w="w"):
        self.L.Continuity(st.ptr(w), st.g, st.p, st.prop)

    def face_scalars(self, st, tvd):
        self.L.HorizontalFaceScalars(st.g, st.p, st.prop, st.ptr("s"), st.ptr("boundary_s"), int(tvd), 0, 0)

    def update_dz(self, st, option):
        self.L.UpdateDZ(st.g, st.p, st.prop, int(option))

    def set_flux_height(self, st):
        self.L.SetFluxHeight(st.g, st.p, st.prop)

    def upredictor_begin(self, st):
        self.L.sunhip_UPredictor_begin(st.g, st.p, st.prop, 0, 1, 0)

    def upredictor_solve(self, st):
        self.L.sunhip_UPredictor_solve(st.g, st.p, st.prop, 0, 1, 0)

    def wpredictor(self, st):
        self.L.sunhip_WPredictor(st.g, st.p, st.prop, 0, 1, 0)

    def update_salinity(self, st):
        kappa = st.get_prop("kappa_s")
        kappaH = st.get_prop("kappa_sH")
        theta = st.get_prop("theta")
        tvd = int(st.get_prop("TVDsalt"))
        self.L.UpdateScalars(st.g, st.p, st.prop, st.ptr("wnew"), st.ptr("s"), st.ptr("boundary_s"),
                             st.ptr("Cn_R"), kappa, kappaH, st.ptr("kappa_tv"), theta, None, None, None, None,
                             0, 0, 0, 0, 1, tvd)

    def update_salinity_src(self, st, flags, alpha_top=0, alpha_bot=0):
        """As the CPU adapters: flags 1 src1 = uold, 2 src2 = wtmp, 4 Ftop =
        htmp, 8 Fbot = htmp2."""
        g = st.get_prop
        self.L.UpdateScalars(st.g, st.p, st.prop, st.ptr("wnew"), st.ptr("s"), st.ptr("boundary_s"),
                             st.ptr("Cn_R"), g("kappa_s"), g("kappa_sH"), st.ptr("kappa_tv"), g("theta"),
                             st.ptr("uold") if flags & 1 else None, st.ptr("wtmp") if flags & 2 else None,
                             st.ptr("htmp") if flags & 4 else None, st.ptr("htmp2") if flags & 8 else None,
                             int(alpha_top), int(alpha_bot), 0, 0, 1, int(g("TVDsalt")))

    def stats(self, which):
        n = C.c_int()
        r = C.c_double()
        ms = C.c_double()
        if self.L.sunhip_last_solve(0 if which == "q" else 1, C.byref(n), C.byref(r), C.byref(ms)):
            return None, None
        return n.value, r.value


_cache = {}


def oracle():
    if "o" not in _cache:
        _cache["o"] = OracleImpl()
    return _cache["o"]


def reference():
    if not os.path.exists(REF_SO):
        return None
    if "r" not in _cache:
        _cache["r"] = RefImpl()
    return _cache["r"]


def dropin():
    if "d" not in _cache:
        _cache["d"] = DropinImpl()
    return _cache["d"]
