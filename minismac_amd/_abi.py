"""ctypes mirror of the structures in include/smac_amd.h (layout must match exactly)."""
from __future__ import annotations

import ctypes as C

SMAC_OK = 0
STATUS_NAMES = {
    0: "SMAC_OK", 1: "SMAC_ERR_ARG", 2: "SMAC_ERR_IMPSCH", 3: "SMAC_ERR_IFLXO",
    4: "SMAC_ERR_NEQS", 5: "SMAC_ERR_SIZE", 6: "SMAC_ERR_BC", 7: "SMAC_ERR_NONFINITE",
    8: "SMAC_ERR_HIP", 9: "SMAC_ERR_NOGPU", 10: "SMAC_ERR_STALL", 11: "SMAC_ERR_NOCONV",
}
SMAC_FP64, SMAC_FP32_LHS = 0, 1
SMAC_ORDER_REFERENCE, SMAC_ORDER_ZEBRA = 0, 1


class SmacParams(C.Structure):
    _fields_ = [
        ("alpha", C.c_double), ("beta", C.c_double), ("reynum", C.c_double),
        ("dtau", C.c_double), ("dt", C.c_double), ("underr", C.c_double),
        ("cdis", C.c_double), ("pin", C.c_double), ("pout", C.c_double),
        ("chord", C.c_double), ("xmom", C.c_double), ("ymom", C.c_double),
        ("impsch", C.c_int), ("ivis", C.c_int), ("iturb", C.c_int), ("iflxo", C.c_int),
        ("njsp", C.c_int), ("nksp", C.c_int), ("nbcimp", C.c_int), ("ntjsp", C.c_int),
        ("ntksp", C.c_int), ("precision", C.c_int), ("line_order", C.c_int),
    ]


class SmacBcRegion(C.Structure):
    _fields_ = [("ibcval", C.c_int), ("jbeg", C.c_int), ("jend", C.c_int),
                ("kbeg", C.c_int), ("kend", C.c_int)]


class SmacIterStats(C.Structure):
    _fields_ = [
        ("nt", C.c_int), ("jres", C.c_int), ("kres", C.c_int), ("nonfinite", C.c_int),
        ("resmax", C.c_double), ("resmax0", C.c_double), ("divmax", C.c_double),
        ("turres", C.c_double), ("rhsnorm", C.c_double), ("cl", C.c_double),
        ("cd", C.c_double), ("cm", C.c_double), ("clp", C.c_double), ("cdp", C.c_double),
        ("fxp", C.c_double), ("fyp", C.c_double), ("fxs", C.c_double), ("fys", C.c_double),
        ("fmom", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


def make_params(**kw) -> SmacParams:
    """Reference defaults (ref/initia.f:104-112, 418-437; ref/smac2d.in; ref/geom.f:72-76)
    overridden by keyword."""
    d = dict(alpha=10.0, beta=5.0, reynum=1.5e6, dtau=0.1, dt=1.0e12, underr=0.01, cdis=1.0,
             pin=1.5, pout=1.0, chord=1.0, xmom=0.25, ymom=0.0, impsch=1, ivis=1, iturb=1,
             iflxo=3, njsp=2, nksp=2, nbcimp=1, ntjsp=2, ntksp=0, precision=SMAC_FP64,
             line_order=SMAC_ORDER_REFERENCE)
    unknown = set(kw) - set(d)
    if unknown:
        raise TypeError(f"unknown parameter(s): {sorted(unknown)}")
    d.update(kw)
    p = SmacParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def params_dict(p: SmacParams) -> dict:
    return {name: getattr(p, name) for name, _ in p._fields_}


class SmacAdjointStats(C.Structure):
    _fields_ = [("iters", C.c_int * 2), ("resid", C.c_double * 2), ("rhsnorm", C.c_double),
                ("seconds", C.c_double)]

    def as_dict(self) -> dict:
        return {"iters": list(self.iters), "resid": list(self.resid), "rhsnorm": self.rhsnorm,
                "seconds": self.seconds}
