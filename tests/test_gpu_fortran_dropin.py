"""The Fortran side of the drop-in boundary, executed: tests/fortran/smac_harness.F, linked with
minismac_amd/fortran/smac_dropin.F (the lrelax/lusgs a maintainer links in place of the
reference's) and libsmac_amd.so, runs both routines on the GPU with the reference's COMMON
blocks filled as the reference fills them -- the settings from /timm/ /visc/ /impli/ ...,
the wake-cut regions in /bcwake/ as ref/rbc.f lays them out -- and returns dq (and, for
lusgs, the LU factors and the destroyed rhs) bit for bit equal to the oracle's.

The harness is built by __graft_entry__.build() where the reference's include files are
(they declare the COMMON layout); without it the test is skipped.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
from minismac_amd.case import config_case

from conftest import ROOT

pytestmark = pytest.mark.gpu

HARNESS = os.path.join(ROOT, "oracle", "_ref", "dropin", "smac_harness")


def _bcwake(case):
    """/bcwake/ as ref/rbc.f builds it for a pair of cut segments on k = 1: region i's
    points pair, in reverse order, with those of the other region."""
    segs = [r for r in case.resolved_regions() if r[0] == 60]
    out = []
    for i, (_, jb, je, _, _) in enumerate(segs):
        other = segs[1 - i]
        out.append((jb, je, other[2]))
    return out


@pytest.mark.skipif(not os.path.exists(HARNESS), reason="Fortran harness not built "
                    "(needs the reference's include files at build time)")
@pytest.mark.parametrize("impsch,neqs", [(1, 3), (2, 3), (2, 1)])
def test_fortran_dropin_matches_oracle(tmp_path, impsch, neqs):
    case = config_case(0, impsch=impsch)
    o = pyoracle.Oracle(case)
    for _ in range(3):
        o.step()
    o.assemble()
    q, s = o.array("q"), o.array("s")
    blocks = [o.array("lhs_" + n) for n in ("btc", "bjm", "bjp", "bkm", "bkp")]
    if neqs == 1:
        # the scalar system: the (1,1) entries and the first equation
        q, s = q[:, :, :1], s[:, :, :1]
        blocks = [b[:, :, :1] for b in blocks]
    p = case.params
    wake = _bcwake(case)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([case.jmax, case.kmax, neqs, impsch], dtype=np.int32).tofile(f)
        np.array([p.iflxo, p.ivis, p.iturb, p.njsp, p.nksp, p.nbcimp, p.ntjsp, p.ntksp],
                 dtype=np.int32).tofile(f)
        np.array([p.alpha, p.beta, p.reynum, p.dtau, p.dt, p.underr, p.cdis],
                 dtype=np.float64).tofile(f)
        np.array([p.pin, p.pout], dtype=np.float64).tofile(f)
        np.array([len(wake)], dtype=np.int32).tofile(f)
        for w in wake:
            np.array(w, dtype=np.int32).tofile(f)
        for a in [q, s] + blocks:
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(f)
    print(f"Fortran harness: impsch={impsch} neqs={neqs}", flush=True)
    r = subprocess.run([HARNESS, str(fin), str(fout)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    nel = case.npts * neqs
    out = np.fromfile(fout, dtype=np.float64)
    dq = out[:nel].reshape(q.shape, order="F")
    rhs = out[nel:2 * nel].reshape(q.shape, order="F")
    # the oracle on the same system, through its stand-alone solvers
    dq_o = np.zeros_like(q, order="F")
    rhs_o = np.asfortranarray(s.copy())
    bl = [np.asfortranarray(b.copy()) for b in blocks]
    pairs = case.wake_pairs()
    if impsch == 1:
        pyoracle.lrelax(p, np.asfortranarray(q), dq_o, rhs_o, *bl, wake_pairs=pairs)
    else:
        scr1 = np.zeros_like(bl[0], order="F")
        pyoracle.lusgs(p, neqs, np.asfortranarray(q), dq_o, rhs_o, *bl, scr1, wake_pairs=pairs)
        assert np.array_equal(out[2 * nel:].reshape(bl[0].shape, order="F"), scr1)
    assert np.array_equal(dq, dq_o)
    assert np.array_equal(rhs, rhs_o)
