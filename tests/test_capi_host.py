"""The C ABI library without a GPU: it loads, exports every entry point declared in
include/smac_amd.h, and turns the reference's configuration errors into status codes
before touching a device (the reference stops or aborts, ref/step.f:557-560,
ref/conup.f:200-209, ref/lusgs.f:108-115, ref/initia.f:205-247)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from minismac_amd import solver
from minismac_amd._abi import make_params
from minismac_amd.case import config_case, dptr, fortran

from conftest import ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "smac_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smac_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = solver.lib()
    names = declared_functions()
    assert set(names) == set(solver.EXPORTS)
    for n in names:
        assert hasattr(L, n), n
    assert b"gfx950" in L.smac_version()


def _create(**kw):
    case = config_case(0)
    p = make_params(**{**dict(alpha=4.0, beta=10.0, reynum=1e6), **kw})
    x, y = fortran(case.x), fortran(case.y)
    rtxy, dj = case.metrics()
    rtxy, dj = fortran(rtxy), fortran(dj)
    regs = case.c_regions()
    h = C.c_void_p()
    rc = solver.lib().smac_create(C.byref(p), case.jmax, case.kmax, dptr(x), dptr(y), dptr(rtxy),
                                  dptr(dj), len(regs), regs, C.byref(h))
    return rc, solver.lib().smac_last_error().decode()


@pytest.mark.parametrize("kw,code", [
    (dict(impsch=3), 2), (dict(iflxo=5), 3), (dict(underr=0.0), 1), (dict(underr=1.5), 1),
    (dict(ivis=2), 1), (dict(ntksp=1), 1), (dict(dtau=-1.0), 1),
    (dict(line_order=2), 1), (dict(line_order=1, impsch=2), 1),
])
def test_bad_settings_return_status(kw, code):
    rc, msg = _create(**kw)
    assert rc == code and msg


def test_lusgs_rejects_bad_neqs_without_gpu():
    p = make_params()
    z = np.zeros((5, 5, 9), order="F")
    v = np.zeros((5, 5, 3), order="F")
    rc = solver.lib().smac_lusgs(C.byref(p), 5, 5, 2, 0, None, dptr(v), dptr(v), dptr(v),
                                 dptr(z), dptr(z), dptr(z), dptr(z), dptr(z), dptr(z))
    assert rc == 4


def test_oversize_grid_is_a_status_not_a_crash():
    p = make_params()
    h = C.c_void_p()
    x = np.zeros(4, order="F")
    rc = solver.lib().smac_create(C.byref(p), 9000, 4, dptr(x), dptr(x), dptr(x), dptr(x), 1,
                                  (solver.SmacBcRegion * 1)(), C.byref(h))
    assert rc == 5
