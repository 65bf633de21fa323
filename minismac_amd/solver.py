"""Host side of the gfx950 implicit iteration: ctypes binding of libsmac_amd.so.

``Solver`` holds one device-resident case (smac_create/smac_step/smac_run, the replacement
of the reference's solve loop body around ``step``, ref/smac2d.f:1376-1431); ``lrelax`` and
``lusgs`` are the host-array drop-ins for the reference's own call sites
(ref/step.f:538, ref/step.f:552, ref/tursa.f:252).  There is no CPU fallback: if the
library is missing or no gfx950 device is visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._abi import STATUS_NAMES, SmacAdjointStats, SmacIterStats, SmacParams, SmacBcRegion
from .case import Case, dptr, fortran

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libsmac_amd.so")

# every entry point declared in include/smac_amd.h
EXPORTS = [
    "smac_create", "smac_destroy", "smac_set_state", "smac_get_state", "smac_step", "smac_run",
    "smac_get_array", "smac_timings", "smac_set_option", "smac_lrelax", "smac_lusgs",
    "smac_get_stats", "smac_batch_create", "smac_batch_run", "smac_batch_run_until",
    "smac_batch_timings",
    "smac_batch_destroy", "smac_lusgs_noalloc", "smac_adjoint", "smac_adjoint_apply",
    "smac_last_error", "smac_version",
]

_lib = None


class SmacError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        D = P(C.c_double)
        L.smac_create.argtypes = [P(SmacParams), C.c_int, C.c_int, D, D, D, D, C.c_int,
                                  P(SmacBcRegion), P(C.c_void_p)]
        L.smac_destroy.argtypes = [C.c_void_p]
        L.smac_set_state.argtypes = [C.c_void_p, D, D, C.c_int, C.c_double, C.c_double]
        L.smac_get_state.argtypes = [C.c_void_p, D, D, D]
        L.smac_step.argtypes = [C.c_void_p, P(SmacIterStats)]
        L.smac_run.argtypes = [C.c_void_p, C.c_int, C.c_double, P(SmacIterStats), P(C.c_int)]
        L.smac_get_array.argtypes = [C.c_void_p, C.c_char_p, D]
        L.smac_timings.argtypes = [C.c_void_p, D]
        L.smac_set_option.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.smac_lrelax.argtypes = [P(SmacParams), C.c_int, C.c_int, C.c_int, P(C.c_int)] + [D] * 8
        L.smac_lusgs.argtypes = [P(SmacParams), C.c_int, C.c_int, C.c_int, C.c_int,
                                 P(C.c_int)] + [D] * 9
        L.smac_get_stats.argtypes = [C.c_void_p, P(SmacIterStats)]
        L.smac_batch_create.argtypes = [P(C.c_void_p), C.c_int, P(C.c_void_p)]
        L.smac_batch_run.argtypes = [C.c_void_p, C.c_int]
        L.smac_batch_run_until.argtypes = [C.c_void_p, C.c_int, C.c_double, P(C.c_int)]
        L.smac_batch_timings.argtypes = [C.c_void_p, D]
        L.smac_batch_destroy.argtypes = [C.c_void_p]
        L.smac_adjoint.argtypes = [C.c_void_p, C.c_int, D, D, C.c_int, C.c_double, D, D,
                                   P(SmacAdjointStats)]
        L.smac_adjoint_apply.argtypes = [C.c_void_p, C.c_int, D, D]
        L.smac_lusgs_noalloc.argtypes = [P(SmacParams), C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_void_p] + [C.c_void_p] * 11
        for fn in EXPORTS[:-2]:
            getattr(L, fn).restype = C.c_int
        L.smac_last_error.restype = C.c_char_p
        L.smac_version.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc: int, where: str) -> None:
    if rc != 0:
        raise SmacError(rc, where, lib().smac_last_error().decode(errors="replace"))


_PLANES = {"q": 3, "dq": 3, "s": 3, "anut": 1, "vnut": 1, "smin": 1, "sf": 1, "stc": 1,
           "ds": 1, "vort": 1, "btc": 9, "bjm": 9, "bjp": 9, "bkm": 9, "bkp": 9,
           "adj_jac": 81, "adj_lam_cl": 3, "adj_lam_cd": 3, "adj_gcl": 3, "adj_gcd": 3}


class Solver:
    """One case resident on the current HIP device."""

    def __init__(self, case: Case, keep_lhs: bool = False, timing: bool = False,
                 options: dict | None = None):
        L = lib()
        self.case = case
        self.jmax, self.kmax = case.jmax, case.kmax
        self.params = case.params
        x, y = fortran(case.x), fortran(case.y)
        rtxy, dj = case.metrics()
        rtxy, dj = fortran(rtxy), fortran(dj)
        self._regs = case.c_regions()
        h = C.c_void_p()
        _check(L.smac_create(C.byref(self.params), self.jmax, self.kmax, dptr(x), dptr(y),
                             dptr(rtxy), dptr(dj), len(self._regs), self._regs, C.byref(h)),
               "smac_create")
        self.h = h
        if keep_lhs:
            _check(L.smac_set_option(self.h, b"keep_lhs", 1), "smac_set_option")
        if timing:
            _check(L.smac_set_option(self.h, b"timing", 1), "smac_set_option")
        for name, value in (options or {}).items():
            self.set_option(name, value)

    def set_option(self, name: str, value: int) -> None:
        _check(lib().smac_set_option(self.h, name.encode(), int(value)), "smac_set_option")

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().smac_destroy(self.h)
            self.h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self) -> dict:
        st = SmacIterStats()
        _check(lib().smac_step(self.h, C.byref(st)), "smac_step")
        return st.as_dict()

    def run(self, niter: int, drop: float = 0.0):
        st = SmacIterStats()
        done = C.c_int(0)
        _check(lib().smac_run(self.h, niter, drop, C.byref(st), C.byref(done)), "smac_run")
        return st.as_dict(), done.value

    def stats(self) -> dict:
        """Stats of the last pass made on this case (by step, run or a Batch)."""
        st = SmacIterStats()
        _check(lib().smac_get_stats(self.h, C.byref(st)), "smac_get_stats")
        return st.as_dict()

    def timings(self) -> dict:
        out = (C.c_double * 7)()
        _check(lib().smac_timings(self.h, out), "smac_timings")
        keys = ["sa", "assemble", "resid", "solve", "update_force", "band_fwd", "pass"]
        return dict(zip(keys, list(out)))

    def array(self, name: str) -> np.ndarray:
        planes = _PLANES[name]
        out = np.zeros((self.jmax, self.kmax, planes) if planes > 1 else (self.jmax, self.kmax),
                       order="F")
        _check(lib().smac_get_array(self.h, name.encode(), dptr(out)), "smac_get_array")
        return out

    def adjoint(self, shape_dirs=(), maxit: int = 2000, tol: float = 1e-12, check: bool = True):
        """Gradients of CL and CD at the current (steady) state (smac_adjoint): returns
        (dcl, dcd, stats), dcl[0]/dcd[0] per degree of alpha, then one entry per shape
        direction, each a (dx, dy) pair of (jmax, kmax) grid displacements.  check=False
        returns unconverged solves too (SMAC_ERR_NOCONV), e.g. to build the operators."""
        n = len(shape_dirs)
        npts = self.jmax * self.kmax
        dx = np.zeros((max(n, 1), npts))
        dy = np.zeros((max(n, 1), npts))
        for i, (ax, ay) in enumerate(shape_dirs):
            dx[i] = np.asarray(ax, dtype=np.float64).ravel(order="F")
            dy[i] = np.asarray(ay, dtype=np.float64).ravel(order="F")
        dcl = np.zeros(1 + n)
        dcd = np.zeros(1 + n)
        st = SmacAdjointStats()
        rc = lib().smac_adjoint(self.h, n, dptr(dx.ravel()), dptr(dy.ravel()), maxit, tol,
                                dptr(dcl), dptr(dcd), C.byref(st))
        if not (rc == 11 and not check):
            _check(rc, "smac_adjoint")
        return dcl, dcd, st.as_dict()

    def adjoint_apply(self, op: int, v: np.ndarray) -> np.ndarray:
        """An operator of the last adjoint call on a (jmax, kmax, 3) field: 0 F_q^T, 1 the
        adjoint operator, 2 P^T, 3 P (smac_adjoint_apply)."""
        vin = np.asfortranarray(np.asarray(v, dtype=np.float64).reshape(self.jmax, self.kmax, 3,
                                                                          order="F"))
        out = np.zeros_like(vin, order="F")
        _check(lib().smac_adjoint_apply(self.h, op, dptr(vin), dptr(out)), "smac_adjoint_apply")
        return out

    def set_state(self, q=None, anut=None, nt=0, resmax0=1.0, clift=0.0) -> None:
        qa = fortran(q) if q is not None else None
        aa = fortran(anut) if anut is not None else None
        _check(lib().smac_set_state(self.h, dptr(qa) if qa is not None else None,
                                    dptr(aa) if aa is not None else None, nt, resmax0, clift),
               "smac_set_state")

    def get_state(self):
        q = np.zeros((self.jmax, self.kmax, 3), order="F")
        a = np.zeros((self.jmax, self.kmax), order="F")
        v = np.zeros((self.jmax, self.kmax), order="F")
        _check(lib().smac_get_state(self.h, dptr(q), dptr(a), dptr(v)), "smac_get_state")
        return q, a, v


class Batch:
    """Several Solvers advanced together on one stream, the LU-SGS band sweeps of all of them
    in one launch per direction (smac_batch_run)."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        b = C.c_void_p()
        _check(lib().smac_batch_create(arr, len(self.solvers), C.byref(b)), "smac_batch_create")
        self.b = b

    def run(self, npass: int) -> None:
        _check(lib().smac_batch_run(self.b, int(npass)), "smac_batch_run")

    def run_until(self, max_pass: int, drop: float) -> list:
        """Up to max_pass passes, each case stopping on the device after the pass that drops
        its residual norm to `drop` times its first pass's (smac_run's criterion); returns
        the passes each case made."""
        out = (C.c_int * len(self.solvers))()
        _check(lib().smac_batch_run_until(self.b, int(max_pass), float(drop), out),
               "smac_batch_run_until")
        return list(out)

    def timings(self) -> dict:
        """The forward mean-flow band launch of the last pass: device ms, grid points it
        covered and its precision (0 FP64, 1 FP32-LHS)."""
        out = (C.c_double * 3)()
        _check(lib().smac_batch_timings(self.b, out), "smac_batch_timings")
        return {"band_fwd_ms": out[0], "points": int(out[1]), "precision": int(out[2])}

    def close(self) -> None:
        if getattr(self, "b", None):
            lib().smac_batch_destroy(self.b)
            self.b = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _pairs_arg(pairs):
    if pairs is None or len(pairs) == 0:
        return 0, None, None
    arr = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1))
    return arr.size // 2, arr, arr.ctypes.data_as(C.POINTER(C.c_int))


def lrelax(params: SmacParams, q, dq, s, btc, bjm, bjp, bkm, bkp, wake_pairs=None) -> None:
    """Drop-in for lrelax(jmax,kmax,q,dq,s,btc,bjm,bjp,bkm,bkp) (ref/lrelax.f:51): writes
    dq and the wake-cut rows of s in place (Fortran-ordered float64 arrays)."""
    jmax, kmax = q.shape[:2]
    n, keep, ptr = _pairs_arg(wake_pairs)
    for a in (q, dq, s, btc, bjm, bjp, bkm, bkp):
        if not (a.dtype == np.float64 and a.flags.f_contiguous):
            raise TypeError("arrays must be Fortran-ordered float64")
    _check(lib().smac_lrelax(C.byref(params), jmax, kmax, n, ptr, dptr(q), dptr(dq), dptr(s),
                             dptr(btc), dptr(bjm), dptr(bjp), dptr(bkm), dptr(bkp)), "smac_lrelax")
    del keep


def lusgs(params: SmacParams, neqs: int, q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1,
          wake_pairs=None) -> None:
    """Drop-in for lusgs(jmax,kmax,neqs,q,dq,rhs,btc,bjm,bjp,bkm,bkp,scr1)
    (ref/lusgs.f:51): writes dq, rhs and scr1 in place."""
    jmax, kmax = q.shape[:2]
    n, keep, ptr = _pairs_arg(wake_pairs)
    for a in (q, dq, rhs, btc, bjm, bjp, bkm, bkp, scr1):
        if not (a.dtype == np.float64 and a.flags.f_contiguous):
            raise TypeError("arrays must be Fortran-ordered float64")
    _check(lib().smac_lusgs(C.byref(params), jmax, kmax, neqs, n, ptr, dptr(q), dptr(dq),
                            dptr(rhs), dptr(btc), dptr(bjm), dptr(bjp), dptr(bkm), dptr(bkp),
                            dptr(scr1)), "smac_lusgs")
    del keep


def lusgs_noalloc(params: SmacParams, jmax: int, kmax: int, neqs: int, npair: int, wake_ptr: int,
                  q_ptr: int, dq_ptr: int, rhs_ptr: int, btc_ptr: int, bjm_ptr: int, bjp_ptr: int,
                  bkm_ptr: int, bkp_ptr: int, scr1_ptr: int, stall_ptr: int = 0,
                  stream: int = 0) -> None:
    """smac_lusgs_noalloc on device addresses (e.g. torch tensors' data_ptr()): enqueued on
    `stream` and not waited for."""
    _check(lib().smac_lusgs_noalloc(C.byref(params), jmax, kmax, neqs, npair, wake_ptr or None,
                                    q_ptr, dq_ptr, rhs_ptr, btc_ptr, bjm_ptr, bjp_ptr, bkm_ptr,
                                    bkp_ptr, scr1_ptr, stall_ptr or None, stream or None),
           "smac_lusgs_noalloc")
