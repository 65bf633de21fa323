"""Write the golden fixtures of the compiled reference (run here, where /root/reference and
oracle/_ref/smac2d_ref exist; the fixtures are committed so the GPU box needs neither).

Each fixture ref_<name>.npz holds, per iteration nt = 1..n, what the reference computed at
full precision through the instrumented seams of oracle/ref_seams.F: resmax0, resmax, jres,
kres, divmax, turres, rhsnorm (= ||s||_2 entering the solver), the wall force components and
CL/CD/CM formed from them, and the final q and anut.

The shipped 119x31 case uses copies of the reference's own input files, kept here as
shipped_xy_119x31.dat / shipped_bcmain_119x31.dat (input data of the reference's only test
case).  The generated cases use minismac_amd.case.config_case.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from minismac_amd.case import config_case, shipped_case  # noqa: E402

REF_INPUTS = "/root/reference/ref"
SHIPPED_XY = os.path.join(HERE, "shipped_xy_119x31.dat")
SHIPPED_BC = os.path.join(HERE, "shipped_bcmain_119x31.dat")


def cases():
    return {
        "shipped_lr": (lambda: shipped_case(SHIPPED_XY, SHIPPED_BC, impsch=1), 100),
        "shipped_lusgs": (lambda: shipped_case(SHIPPED_XY, SHIPPED_BC, impsch=2), 100),
        "c1_119x31_lr": (lambda: config_case(0, impsch=1), 500),
        "c1_119x31_lusgs": (lambda: config_case(0, impsch=2), 500),
    }


def main() -> None:
    for src, dst in ((os.path.join(REF_INPUTS, "xy.dat_original_119x31"), SHIPPED_XY),
                     (os.path.join(REF_INPUTS, "bcmain.dat_original_119x31"), SHIPPED_BC)):
        if not os.path.exists(dst):
            shutil.copyfile(src, dst)
    for name, (mk, n) in cases().items():
        case = mk()
        d = tempfile.mkdtemp(prefix=f"golden_{name}_")
        tr = pyoracle.run_reference(case, n, workdir=d, dump_at=(n,))
        qo = pyoracle.load_dump(d, "qout", n, case.jmax, case.kmax)
        arrays = {k: np.asarray(v) for k, v in tr.items()
                  if isinstance(v, np.ndarray) and k != "nt"}
        arrays["nt"] = tr["nt"]
        arrays["q_final"] = qo["q"]
        arrays["anut_final"] = qo["anut"]
        out = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(out, **arrays)
        shutil.rmtree(d)
        print(f"{out}: {n} iterations, final cl={arrays['cl'][-1]:.10f} cd={arrays['cd'][-1]:.10f}")


if __name__ == "__main__":
    main()
