"""The reference's own Fortran driver, unchanged, with lrelax/lusgs replaced by the
ISO_C_BINDING drop-ins of minismac_amd/fortran (oracle/_ref/smac2d_gpu): its trajectory on
the shipped case is the reference's, bit for bit.  Only the solver moved to the GPU, and
it has no transcendental functions, so SA agrees exactly too."""
import os

import numpy as np
import pytest

import pyoracle
from minismac_amd.case import shipped_case

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

EXE = os.path.join(pyoracle.REF_DIR, "smac2d_gpu")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in build (make -C oracle gpu-dropin) absent")
@pytest.mark.parametrize("impsch,tag", [(1, "shipped_lr"), (2, "shipped_lusgs")])
def test_reference_driver_with_gpu_solver(tmp_path, impsch, tag):
    ref = np.load(os.path.join(GOLDEN, f"ref_{tag}.npz"))
    case = shipped_case(os.path.join(GOLDEN, "shipped_xy_119x31.dat"),
                        os.path.join(GOLDEN, "shipped_bcmain_119x31.dat"), impsch=impsch)
    n = len(ref["nt"])
    old = pyoracle.REF_EXE
    try:
        pyoracle.REF_EXE = EXE
        tr = pyoracle.run_reference(case, n, workdir=str(tmp_path), timeout=600)
    finally:
        pyoracle.REF_EXE = old
    assert "failed" not in tr["stdout"]
    for k in ("resmax", "jres", "kres", "divmax", "turres", "fxp", "fyp", "fxs", "fys", "fmom"):
        a, b = tr[k], ref[k]
        bad = np.nonzero(a != b)[0]
        assert bad.size == 0, f"{k} differs first at nt={bad[0] + 1}: {a[bad[0]]!r} vs {b[bad[0]]!r}"
