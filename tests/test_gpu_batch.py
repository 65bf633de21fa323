"""Batches: several cases advanced together on one stream, with the LU-SGS band sweeps of all
cases in one launch per direction (smac_batch_run).  A batched case must follow exactly the
trajectory the same case follows when it is stepped on its own (bit for bit: the same kernels
on the same operands), across several run calls with host reads, option changes and state
changes between them.
"""
import numpy as np
import pytest

from minismac_amd._abi import SMAC_FP32_LHS
from minismac_amd.case import batch_cases, config_case, naca_case
from minismac_amd.solver import Batch, SmacError, Solver

pytestmark = pytest.mark.gpu


def _mixed_cases():
    return [
        naca_case(239, 61, m=0.02, p=0.3, t=0.10, alpha=6.0, reynum=2.0e5, beta=10.0, impsch=2),
        config_case(0, impsch=1),
        naca_case(239, 61, m=0.05, p=0.5, t=0.15, alpha=-2.0, reynum=5.0e6, beta=10.0,
                  impsch=2, precision=SMAC_FP32_LHS),
        naca_case(477, 121, m=0.0, p=0.4, t=0.09, alpha=12.0, reynum=1.0e6, beta=10.0, impsch=2),
    ]


def _same(a: Solver, b: Solver, tag: str):
    qa, aa, va = a.get_state()
    qb, ab, vb = b.get_state()
    assert np.array_equal(qa, qb), f"{tag}: q differs ({a.case.name})"
    assert np.array_equal(aa, ab), f"{tag}: nu~ differs ({a.case.name})"
    assert np.array_equal(va, vb), f"{tag}: nu_t differs ({a.case.name})"
    sa, sb = a.stats(), b.stats()
    for k in ("nt", "rhsnorm", "resmax", "cl", "cd", "cm"):
        assert sa[k] == sb[k], f"{tag}: {k} {sa[k]!r} vs {sb[k]!r} ({a.case.name})"


def test_multi_case_launches_match_eager_runs():
    """Batches of LU-SGS cases run each kernel of a pass once for all cases after the first
    pass (k_multi): mixed sizes, both precisions, cases with and without the SA model and an
    inviscid one, band splits changed between calls (the kernel tables are rebuilt) and a
    state change that sends a case back to its first pass."""
    cases = [
        naca_case(239, 61, m=0.02, p=0.3, t=0.10, alpha=6.0, reynum=2.0e5, beta=10.0, impsch=2),
        naca_case(477, 121, m=0.04, p=0.4, t=0.12, alpha=4.0, reynum=1.0e6, beta=10.0,
                  impsch=2, precision=SMAC_FP32_LHS),
        naca_case(239, 61, m=0.05, p=0.5, t=0.15, alpha=-2.0, reynum=5.0e6, beta=10.0,
                  impsch=2, iturb=0),
        config_case(0, impsch=2, ivis=0, iturb=0),
        config_case(0, impsch=2, precision=SMAC_FP32_LHS),
    ]
    eager = [Solver(c) for c in cases]
    batched = [Solver(c) for c in cases]
    try:
        with Batch(batched) as b:
            for call, npass in enumerate((2, 5, 3)):
                print(f"multi-case batch call {call}: {npass} passes", flush=True)
                if call == 1:
                    for s in (eager[1], batched[1]):
                        s.set_option("lusgs_rows", 16)
                b.run(npass)
                for s in eager:
                    s.run(npass)
                for se, sb in zip(eager, batched):
                    _same(se, sb, f"multi call {call}")
            # one case back to its first pass: that pass runs per case, then k_multi again
            q0, a0, _ = eager[3].get_state()
            for s in (eager[3], batched[3]):
                s.set_state(q=q0, anut=a0, nt=0, resmax0=0.0, clift=0.0)
            b.run(3)
            for s in eager:
                s.run(3)
            for se, sb in zip(eager, batched):
                _same(se, sb, "after reset")
    finally:
        for s in eager + batched:
            s.close()


def test_batch_runs_match_eager_runs():
    cases = _mixed_cases()
    eager = [Solver(c) for c in cases]
    batched = [Solver(c) for c in cases]
    try:
        with Batch(batched) as b:
            for call, npass in enumerate((1, 3, 2, 4)):
                print(f"batch call {call}: {npass} passes of {len(cases)} cases", flush=True)
                if call == 2:
                    # a new band split on one case rebuilds the batch's band tables
                    for s in (eager[0], batched[0]):
                        s.set_option("lusgs_rows", 24)
                b.run(npass)
                for s in eager:
                    s.run(npass)
                for se, sb in zip(eager, batched):
                    _same(se, sb, f"call {call}")
            # host-side state changes between runs reach the batched cases
            for se, sb in zip(eager, batched):
                q, a, _ = se.get_state()
                q = q * (1.0 + 1e-3)
                st = se.stats()
                for s in (se, sb):
                    s.set_state(q=q, anut=a, nt=st["nt"], resmax0=1.0, clift=st["cl"])
            b.run(3)
            for s in eager:
                s.run(3)
            for se, sb in zip(eager, batched):
                _same(se, sb, "after set_state")
    finally:
        for s in eager + batched:
            s.close()


def test_batch_reset_to_first_pass():
    """Resetting a case to nt=0 between calls makes its next pass a first pass again
    (resmax0 latch), in a batch as on its own."""
    cases = _mixed_cases()[:2]
    eager = [Solver(c) for c in cases]
    batched = [Solver(c) for c in cases]
    try:
        with Batch(batched) as b:
            b.run(3)
            for s in eager:
                s.run(3)
            for s in (eager[0], batched[0]):
                s.set_state(q=None, anut=None, nt=0, resmax0=1.0, clift=0.0)
            b.run(3)
            for s in eager:
                s.run(3)
            for se, sb in zip(eager, batched):
                _same(se, sb, "after reset")
    finally:
        for s in eager + batched:
            s.close()


def test_batch_rejects_timing_and_duplicates():
    c = config_case(0, impsch=2)
    with Solver(c, timing=True) as t, Solver(c) as s:
        with pytest.raises(SmacError):
            Batch([t, s])
        with pytest.raises(SmacError):
            Batch([s, s])
        with Batch([s]) as b:
            b.run(0)
        assert s.stats()["nt"] == 0


def test_batch_workload_generator_cases_step():
    """A few of the batched workload's generated cases (seeded, LU-SGS) take finite passes."""
    cs = batch_cases(64, seed=0, impsch=2)
    assert len({(c.jmax, c.kmax) for c in cs}) == 3
    pick = [c for c in cs if c.jmax == 239][:2] + [c for c in cs if c.jmax == 477][:1]
    sol = [Solver(c) for c in pick]
    try:
        with Batch(sol) as b:
            b.run(5)
            t = b.timings()
        for s in sol:
            st = s.stats()
            assert st["nt"] == 5 and np.isfinite(st["rhsnorm"]) and np.isfinite(st["cl"])
        # the timed launch: the FP64 mean-flow bands of all three cases
        assert t["precision"] == 0 and t["points"] == sum(c.npts for c in pick)
        assert 0.0 < t["band_fwd_ms"] < 100.0
    finally:
        for s in sol:
            s.close()


def test_device_side_stop_matches_smac_run():
    """smac_batch_run_until: each case stops on the device after the pass that drops its
    residual norm to `drop` times its first pass's, so its pass count and state equal
    smac_run(max_pass, drop) on it alone -- cases that stop early, one that runs every pass,
    both precisions and two sizes.  A second call continues from the stopped states with the
    same reference norms, and a plain run afterwards advances every case again."""
    def steady(jm, km, **kw):
        return naca_case(jm, km, reynum=1000.0, beta=10.0, impsch=2, **kw)
    cases = [steady(119, 31, alpha=4.0, dtau=10.0),
             steady(119, 31, m=0.02, alpha=8.0, dtau=5.0),
             config_case(0, impsch=2),
             steady(119, 31, alpha=4.0, dtau=10.0, precision=SMAC_FP32_LHS),
             steady(239, 61, alpha=2.0, dtau=10.0)]
    max_pass, drop = 150, 5e-4
    print(f"device-side stop: {len(cases)} cases, max_pass={max_pass}, drop={drop}", flush=True)
    alone = [Solver(c) for c in cases]
    sol = [Solver(c) for c in cases]
    try:
        done = [s.run(max_pass, drop)[1] for s in alone]
        with Batch(sol) as b:
            passes = b.run_until(max_pass, drop)
            assert passes == done, (passes, done)
            assert max(done) == max_pass and min(done) < max_pass, done
            for s0, s1 in zip(alone, sol):
                assert all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
                assert s0.stats()["rhsnorm"] == s1.stats()["rhsnorm"]
            # again from the stopped states: the same reference norms as smac_run keeps
            done2 = [s.run(max_pass, drop)[1] for s in alone]
            assert b.run_until(max_pass, drop) == done2
            # a plain batch run advances every case
            for s in alone:
                s.run(3)
            b.run(3)
        for s0, s1 in zip(alone, sol):
            assert all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
    finally:
        for s in alone + sol:
            s.close()


def test_stop_with_line_relaxation_cases_matches_smac_run():
    """smac_batch_run_until on a batch whose passes take the per-case form (line-relaxation
    cases beside LU-SGS ones): the host reads the stop flags after every pass and gives a
    stopped case no further kernels (its LU-SGS bands leave the multi-case sweeps on the
    device flags); every case's pass count and state equal smac_run on it alone."""
    from minismac_amd._abi import SMAC_ORDER_ZEBRA
    steady = dict(reynum=1000.0, dtau=10.0, beta=10.0, alpha=4.0)
    cases = [naca_case(119, 31, impsch=1, line_order=SMAC_ORDER_ZEBRA, underr=1.0, **steady),
             naca_case(119, 31, impsch=1, line_order=SMAC_ORDER_ZEBRA, underr=1.0,
                       precision=SMAC_FP32_LHS, **steady),
             naca_case(119, 31, impsch=2, **steady),
             config_case(0, impsch=1)]
    max_pass, drop = 150, 5e-4
    print(f"host-side stop: {len(cases)} cases, max_pass={max_pass}, drop={drop}", flush=True)
    alone = [Solver(c) for c in cases]
    sol = [Solver(c) for c in cases]
    try:
        done = [s.run(max_pass, drop)[1] for s in alone]
        with Batch(sol) as b:
            passes = b.run_until(max_pass, drop)
        assert passes == done, (passes, done)
        assert max(done) == max_pass and min(done) < max_pass, done
        for s0, s1 in zip(alone, sol):
            assert all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
            assert s0.stats()["rhsnorm"] == s1.stats()["rhsnorm"]
    finally:
        for s in alone + sol:
            s.close()


def test_batch_with_lagged_line_operands_matches_eager_runs():
    """Option lr_refactor in a batch: each line-relaxation case keeps its own lagged operands
    (factored every N of its own passes), so its batched passes equal its passes alone."""
    from minismac_amd._abi import SMAC_ORDER_ZEBRA
    steady = dict(reynum=1000.0, dtau=10.0, beta=10.0, alpha=4.0, underr=1.0, impsch=1,
                  line_order=SMAC_ORDER_ZEBRA)
    cases = [naca_case(119, 31, **steady), naca_case(239, 61, precision=SMAC_FP32_LHS, **steady)]
    opts = [{"lr_refactor": 3}, {"lr_refactor": 4}]
    alone = [Solver(c, options=o) for c, o in zip(cases, opts)]
    sol = [Solver(c, options=o) for c, o in zip(cases, opts)]
    try:
        for s in alone:
            s.run(4)
            s.run(7)
        with Batch(sol) as b:
            b.run(4)
            b.run(7)
        for s0, s1 in zip(alone, sol):
            assert all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
    finally:
        for s in alone + sol:
            s.close()


def test_device_side_stop_at_full_size():
    """The device-side stop at the workload grid's size: 945x241 and 477x121 steady cases
    (FP32-LHS and FP64) batched to a residual drop equal their smac_run alone."""
    steady = dict(reynum=1000.0, dtau=10.0, beta=10.0, impsch=2)
    cases = [naca_case(945, 241, m=0.04, alpha=8.0, wall_spacing=2e-6,
                       precision=SMAC_FP32_LHS, **steady),
             naca_case(477, 121, m=0.02, alpha=4.0, wall_spacing=5e-6, **steady)]
    max_pass, drop = 80, 0.05
    alone = [Solver(c) for c in cases]
    sol = [Solver(c) for c in cases]
    try:
        done = [s.run(max_pass, drop)[1] for s in alone]
        with Batch(sol) as b:
            passes = b.run_until(max_pass, drop)
        print(f"full-size stop: passes {passes}", flush=True)
        assert passes == done, (passes, done)
        for s0, s1 in zip(alone, sol):
            assert all(np.array_equal(a, c) for a, c in zip(s0.get_state(), s1.get_state()))
    finally:
        for s in alone + sol:
            s.close()
