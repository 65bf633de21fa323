"""Writes lusgs_945x241_divergence.json: the oracle's pass-by-pass record of the 945x241
workload case (NACA 4412, Re = 3e6, alpha = 8, LU-SGS, FP64) from the free stream until the
state goes non-finite.  The reference's mean flow is laminar in its viscous terms and this
case runs away within two hundred passes; the GPU path must follow the same trajectory bit
for bit and fail at the same pass.

    python tests/golden/make_lusgs_945x241_divergence.py      (about 40 s on one core)
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import pyoracle  # noqa: E402
from minismac_amd.case import config_case  # noqa: E402

KEYS = ["resmax", "jres", "kres", "divmax", "rhsnorm", "cl", "cd", "cm"]


def main():
    case = config_case(1, impsch=2)
    o = pyoracle.Oracle(case)
    rows = []
    for n in range(1, 400):
        st = o.step()
        finite = all(math.isfinite(st[k]) for k in ("rhsnorm", "cl", "cd")) and \
            bool(np.all(np.isfinite(o.array("q"))))
        if not finite:
            break
        rows.append([st[k] for k in KEYS])
    out = {"case": case.name, "impsch": 2, "keys": KEYS, "passes": rows,
           "first_nonfinite_pass": len(rows) + 1}
    with open(os.path.join(HERE, "lusgs_945x241_divergence.json"), "w") as f:
        json.dump(out, f)
    print(f"{len(rows)} finite passes; non-finite at pass {len(rows) + 1}")


if __name__ == "__main__":
    main()
