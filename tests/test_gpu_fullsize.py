"""The gfx950 path against the oracle at the workload sizes themselves.

945x241 (LU-SGS and line relaxation) and 1889x481 (LU-SGS) passes are compared with the
oracle bit for bit, as on the small grids: resmax and its location, divmax, the five wall
force sums, CL/CD/CM and q.  ||s||_2 is a parallel sum of 3*jmax*kmax squares in another
order than the oracle's serial one; its rounding grows like sqrt(3N) ulps (about 1e-13
relative at 945x241), so it is held to 1e-12 relative here.  The
1889x481 FP32-LHS case is held to what the precision split fixes: the first pass's residual
(FP64 arithmetic on the same state) is the FP64 path's bit for bit, and its solve differs
from the FP64 solve only at FP32 rounding.  The 64-case seeded FP32-LHS batch runs whole:
every case finite, a sample bit-identical to the same cases stepped alone, and one FP64
case of the batch bit-identical to the oracle.  The 945x241 workload case itself (Re = 3e6,
LU-SGS) runs away from the free stream -- the reference's mean flow is laminar -- and the GPU
follows the oracle's recorded trajectory (tests/golden/lusgs_945x241_divergence.json) bit
for bit through all 181 finite passes and fails at the same pass.
"""
import json
import os

import numpy as np
import pytest

import pyoracle
from minismac_amd._abi import SMAC_FP32_LHS
from minismac_amd.case import batch_cases, config_case
from minismac_amd.solver import Batch, SmacError, Solver

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXACT = ["resmax", "jres", "kres", "divmax", "cl", "cd", "cm", "fxp", "fyp", "fxs", "fys", "fmom"]


def _compare(g, o, tag):
    for k in EXACT:
        assert g[k] == o[k], f"{tag}: {k} gpu={g[k]!r} oracle={o[k]!r}"
    assert g["rhsnorm"] == pytest.approx(o["rhsnorm"], rel=1e-12, abs=0), f"{tag}: rhsnorm"


@pytest.mark.parametrize("cfg,impsch,npass", [(1, 2, 3), (1, 1, 2), (2, 2, 1)])
def test_workload_size_passes_match_oracle(cfg, impsch, npass):
    case = config_case(cfg, impsch=impsch)
    print(f"{case.name} impsch={impsch}: {npass} passes against the oracle", flush=True)
    o = pyoracle.Oracle(case)
    with Solver(case) as g:
        for it in range(npass):
            _compare(g.step(), o.step(), f"{case.name} pass {it + 1}")
        q, _, _ = g.get_state()
    assert np.array_equal(q, o.array("q"))


def test_1889x481_fp32lhs_first_pass():
    """Same state, one pass: the residual is FP64 arithmetic in both modes (bit-identical),
    and the FP32 solve changes dq only at FP32 rounding."""
    c64 = config_case(2, impsch=2)
    c32 = config_case(2, impsch=2, precision=SMAC_FP32_LHS)
    print("1889x481 FP64 vs FP32-LHS, one pass", flush=True)
    with Solver(c64) as a, Solver(c32) as b:
        sa, sb = a.step(), b.step()
        for k in ("resmax", "jres", "kres", "divmax", "rhsnorm"):
            assert sa[k] == sb[k], k
        dqa, dqb = a.array("dq"), b.array("dq")
    assert np.all(np.isfinite(dqb))
    err = np.abs(dqb - dqa).max() / np.abs(dqa).max()
    print(f"max |dq32 - dq64| / max |dq64| = {err:.3e}", flush=True)
    assert err < 1e-4


def test_batch_64_cases_fp32lhs():
    cases = batch_cases(64, seed=0, impsch=2, precision=SMAC_FP32_LHS)
    # one FP64 copy of the first case rides along for the oracle comparison
    c0 = batch_cases(1, seed=0, impsch=2)[0]
    print(f"64-case FP32-LHS batch + 1 FP64 case ({c0.name})", flush=True)
    sol = [Solver(c) for c in cases + [c0]]
    npass = 3
    try:
        with Batch(sol) as b:
            b.run(npass)
        for s in sol:
            st = s.stats()
            assert st["nt"] == npass and not st["nonfinite"], s.case.name
            assert np.isfinite(st["rhsnorm"]) and np.isfinite(st["cl"]), s.case.name
        sizes = [(c.jmax, c.kmax) for c in cases]
        sample = sorted({sizes.index(sz) for sz in set(sizes)})   # first case of each size
        for i in sample:
            with Solver(cases[i]) as e:
                e.run(npass)
                qe, ae, _ = e.get_state()
                se = e.stats()
            qb, ab, _ = sol[i].get_state()
            assert np.array_equal(qe, qb) and np.array_equal(ae, ab), cases[i].name
            assert se["rhsnorm"] == sol[i].stats()["rhsnorm"]
        o = pyoracle.Oracle(c0)
        for _ in range(npass):
            ost = o.step()
        _compare(sol[-1].stats(), ost, f"batch FP64 case {c0.name}")
        assert np.array_equal(sol[-1].get_state()[0], o.array("q"))
    finally:
        for s in sol:
            s.close()


def test_945x241_re3e6_lusgs_runs_away_exactly_as_the_oracle():
    with open(os.path.join(ROOT, "tests", "golden", "lusgs_945x241_divergence.json")) as f:
        gold = json.load(f)
    case = config_case(1, impsch=2)
    keys = gold["keys"]
    print(f"{case.name} LU-SGS: {len(gold['passes'])} passes against the oracle's record",
          flush=True)
    with Solver(case) as g:
        for n, row in enumerate(gold["passes"], 1):
            st = g.step()
            for k, v in zip(keys, row):
                if k == "rhsnorm":
                    assert st[k] == pytest.approx(v, rel=1e-12, abs=0), (n, k)
                else:
                    assert st[k] == v, (n, k, st[k], v)
        with pytest.raises(SmacError) as ei:
            g.step()
        assert ei.value.status == 7          # SMAC_ERR_NONFINITE
    assert gold["first_nonfinite_pass"] == len(gold["passes"]) + 1
