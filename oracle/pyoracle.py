"""Python access to the oracle -- TEST INFRASTRUCTURE ONLY.

``Oracle`` drives the C restatement (oracle/_ref/libsmac_oracle.so) through ctypes;
``lusgs`` / ``lrelax`` run its stand-alone solvers on caller arrays.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg use this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from minismac_amd._abi import SmacIterStats, SmacParams
from minismac_amd.case import Case, dptr, fortran

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "libsmac_oracle.so")
SLIPS_PATH = os.path.join(REF_DIR, "libsmac_oracle_slips.so")
FP64_PATH = os.path.join(REF_DIR, "libsmac_oracle_fp64.so")

_lib = None
_slips = None
_fp64 = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load(LIB_PATH)
    return _lib


def slips_lib():
    """The oracle built with planted slips (negative controls of the pinning tests):
    ``slips_lib().orc_set_slip(n)`` turns slip n on, 0 turns them all off."""
    global _slips
    if _slips is None:
        _slips = _load(SLIPS_PATH)
        _slips.orc_set_slip.argtypes = [C.c_int]
        _slips.orc_set_slip.restype = None
    return _slips


def fp64_lib():
    """The oracle with the viscous terms' FP32 temporaries kept in FP64: a smooth residual
    for the FP64 tangent checks of the adjoint (its central differences carry no FP32
    rounding)."""
    global _fp64
    if _fp64 is None:
        _fp64 = _load(FP64_PATH)
    return _fp64


def _load(path):
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} missing: run `make -C oracle oracle`")
    L = C.CDLL(path)
    P = C.POINTER
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [P(SmacParams), C.c_int, C.c_int, P(C.c_double), P(C.c_double),
                             P(C.c_double), P(C.c_double), C.c_int, C.c_void_p]
    L.orc_destroy.argtypes = [C.c_void_p]
    for fn in ("orc_step", "orc_assemble", "orc_finish"):
        getattr(L, fn).argtypes = [C.c_void_p, P(SmacIterStats)]
        getattr(L, fn).restype = C.c_int
    L.orc_sa_assemble.argtypes = [C.c_void_p]
    L.orc_sa_assemble.restype = C.c_int
    L.orc_solve.argtypes = [C.c_void_p]
    L.orc_solve.restype = C.c_int
    L.orc_array.argtypes = [C.c_void_p, C.c_char_p]
    L.orc_array.restype = P(C.c_double)
    L.orc_set_counters.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_double]
    L.orc_metric.argtypes = [C.c_int, C.c_int, P(C.c_double), P(C.c_double), P(C.c_double),
                             P(C.c_double)]
    L.orc_npairs.argtypes = [C.c_void_p]
    L.orc_npairs.restype = C.c_int
    L.orc_pairs.argtypes = [C.c_void_p]
    L.orc_pairs.restype = P(C.c_int)
    L.orc_lrelax.argtypes = [P(SmacParams), C.c_int, C.c_int, C.c_int, P(C.c_int)] + \
        [P(C.c_double)] * 8
    L.orc_lrelax.restype = C.c_int
    L.orc_lusgs.argtypes = [P(SmacParams), C.c_int, C.c_int, C.c_int, C.c_int, P(C.c_int)] + \
        [P(C.c_double)] * 9
    L.orc_lusgs.restype = C.c_int
    L.orc_face_split.argtypes = [C.c_double] * 6 + [P(C.c_double)] * 2
    L.orc_face_split.restype = None
    return L


_PLANES = {"q": 3, "dq": 3, "s": 3, "anut": 1, "vnut": 1, "rtxy": 6, "dj": 1, "x": 1, "y": 1,
           "smin": 1, "sf": 1, "stc": 1, "ds": 1, "scr1": 9, "sjm": 1, "sjp": 1, "skm": 1,
           "skp": 1, "vort": 1, "fr1": 1}


class Oracle:
    """One case in the C restatement (CPU).  Arrays come back as Fortran-ordered copies."""

    def __init__(self, case: Case, L=None):
        L = L or lib()
        self.L = L
        self.case = case
        self.jmax, self.kmax = case.jmax, case.kmax
        self._x, self._y = fortran(case.x), fortran(case.y)
        rtxy, dj = case.metrics()
        self._rtxy, self._dj = fortran(rtxy), fortran(dj)
        self._regs = case.c_regions()
        self.params = case.params
        h = L.orc_create(C.byref(self.params), self.jmax, self.kmax, dptr(self._x), dptr(self._y),
                         dptr(self._rtxy), dptr(self._dj), len(self._regs),
                         C.cast(self._regs, C.c_void_p))
        if not h:
            raise ValueError("orc_create rejected the case")
        self.h = C.c_void_p(h)

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    __del__ = close

    def _call(self, fn):
        st = SmacIterStats()
        rc = fn(self.h, C.byref(st))
        if rc:
            raise RuntimeError(f"oracle call failed with status {rc}")
        return st.as_dict()

    def step(self) -> dict:
        return self._call(self.L.orc_step)

    def assemble(self) -> dict:
        return self._call(self.L.orc_assemble)

    def sa_assemble(self) -> None:
        self.L.orc_sa_assemble(self.h)

    def solve(self) -> None:
        rc = self.L.orc_solve(self.h)
        if rc:
            raise RuntimeError(f"oracle solve failed with status {rc}")

    def finish(self) -> dict:
        return self._call(self.L.orc_finish)

    def array(self, name: str) -> np.ndarray:
        planes = 9 if name in ("btc", "bjm", "bjp", "bkm", "bkp") or name.startswith("lhs_") \
            else _PLANES[name]
        n = self.jmax * self.kmax * planes
        p = self.L.orc_array(self.h, name.encode())
        if not p:
            raise KeyError(name)
        a = np.ctypeslib.as_array(p, shape=(n,)).copy()
        shape = (self.jmax, self.kmax) + ((planes,) if planes > 1 else ())
        return a.reshape(shape, order="F")

    def set_array(self, name: str, values: np.ndarray) -> None:
        planes = _PLANES.get(name, 9)
        n = self.jmax * self.kmax * planes
        p = self.L.orc_array(self.h, name.encode())
        dst = np.ctypeslib.as_array(p, shape=(n,))
        dst[:] = np.asarray(values, dtype=np.float64).ravel(order="F")

    def set_counters(self, nt: int, resmax0: float, clift: float) -> None:
        self.L.orc_set_counters(self.h, nt, resmax0, clift)

    def wake_pairs(self) -> np.ndarray:
        n = self.L.orc_npairs(self.h)
        p = self.L.orc_pairs(self.h)
        return np.array([[p[2 * i], p[2 * i + 1]] for i in range(n)], dtype=np.int32)


def face_split(rt, rx, ry, q2, q3, beta, L=None):
    """A+ and A- of one face (orc_face_split) as 3x3 arrays."""
    ap = np.zeros(9)
    am = np.zeros(9)
    (L or lib()).orc_face_split(rt, rx, ry, q2, q3, beta, dptr(ap), dptr(am))
    return ap.reshape(3, 3, order="F"), am.reshape(3, 3, order="F")


def metric(x: np.ndarray, y: np.ndarray):
    x, y = fortran(x), fortran(y)
    jm, km = x.shape
    rtxy = np.zeros((jm, km, 2, 3), order="F")
    dj = np.zeros((jm, km), order="F")
    lib().orc_metric(jm, km, dptr(x), dptr(y), dptr(rtxy), dptr(dj))
    return rtxy, dj


# ------------------------------------------------------------------------------------------
# the compiled reference


def _pairs(wake_pairs):
    if wake_pairs is None or len(wake_pairs) == 0:
        return 0, None
    w = np.ascontiguousarray(np.asarray(wake_pairs, dtype=np.int32).reshape(-1))
    return len(w) // 2, w


def lusgs(params: SmacParams, neqs: int, q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1,
          wake_pairs=None) -> None:
    """orc_lusgs on Fortran-ordered float64 arrays (modified in place); wake_pairs are the
    1-based (jt, jb) pairs."""
    jmax, kmax = q.shape[0], q.shape[1]
    n, w = _pairs(wake_pairs)
    arrs = [q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1]
    for a in arrs:
        assert a.flags.f_contiguous and a.dtype == np.float64
    rc = lib().orc_lusgs(C.byref(params), jmax, kmax, neqs, n,
                         w.ctypes.data_as(C.POINTER(C.c_int)) if w is not None else None,
                         *[dptr(a) for a in arrs])
    if rc:
        raise RuntimeError(f"orc_lusgs returned {rc}")


def lrelax(params: SmacParams, q, dq, s, btc, bjm, bjp, bkm, bkp, wake_pairs=None) -> None:
    """orc_lrelax on Fortran-ordered float64 arrays (modified in place)."""
    jmax, kmax = q.shape[0], q.shape[1]
    n, w = _pairs(wake_pairs)
    arrs = [q, dq, s, btc, bjm, bjp, bkm, bkp]
    for a in arrs:
        assert a.flags.f_contiguous and a.dtype == np.float64
    rc = lib().orc_lrelax(C.byref(params), jmax, kmax, n,
                          w.ctypes.data_as(C.POINTER(C.c_int)) if w is not None else None,
                          *[dptr(a) for a in arrs])
    if rc:
        raise RuntimeError(f"orc_lrelax returned {rc}")
