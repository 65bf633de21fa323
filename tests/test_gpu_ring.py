"""The LDS-ring form of the 3x3 LU-SGS band sweeps (option lusgs_ring): operands streamed into
LDS by loader wavefronts.  It performs the same per-point arithmetic in the same order as the
register-prefetch form (lusgs_ring=0), so the results must be bit-identical to it, for FP64
and FP32-LHS, one band and several, alone and in a batch."""
import numpy as np
import pytest

from minismac_amd._abi import SMAC_FP32_LHS
from minismac_amd.case import config_case, naca_case
from minismac_amd.solver import Batch, SmacError, Solver

pytestmark = pytest.mark.gpu


def _run(case, opts, n):
    with Solver(case, options=opts) as s:
        s.run(n)
        q, a, v = s.get_state()
        return q, a, v, s.stats()


@pytest.mark.parametrize("precision", [0, SMAC_FP32_LHS])
@pytest.mark.parametrize("shape,rows", [((945, 241), 64), ((239, 61), 16), ((239, 61), 64)])
def test_ring_sweeps_match_register_sweeps(precision, shape, rows):
    jm, km = shape
    case = (config_case(1, impsch=2, precision=precision) if jm == 945 else
            naca_case(jm, km, alpha=5.0, reynum=1e6, beta=10.0, impsch=2, precision=precision))
    print(f"ring vs registers: {jm}x{km} rows={rows} precision={precision}", flush=True)
    q0, a0, v0, s0 = _run(case, {"lusgs_rows": rows, "lusgs_ring": 0}, 6)
    q1, a1, v1, s1 = _run(case, {"lusgs_rows": rows, "lusgs_ring": 1}, 6)
    assert np.array_equal(q0, q1) and np.array_equal(a0, a1) and np.array_equal(v0, v1)
    assert s0["rhsnorm"] == s1["rhsnorm"] and s0["cl"] == s1["cl"]


def test_ring_in_a_batch_matches_register_sweeps_alone():
    cases = [naca_case(239, 61, m=0.03, alpha=3.0, reynum=5e5, beta=10.0, impsch=2),
             naca_case(477, 121, m=0.02, alpha=8.0, reynum=2e6, beta=10.0, impsch=2,
                       precision=SMAC_FP32_LHS),
             naca_case(239, 61, m=0.05, alpha=-2.0, reynum=3e6, beta=10.0, impsch=2,
                       precision=SMAC_FP32_LHS)]
    ref = [_run(c, {"lusgs_ring": 0}, 5) for c in cases]
    sol = [Solver(c, options={"lusgs_ring": 1}) for c in cases]
    try:
        with Batch(sol) as b:
            b.run(5)
        for (q0, a0, v0, s0), s in zip(ref, sol):
            q1, a1, v1 = s.get_state()
            assert np.array_equal(q0, q1) and np.array_equal(a0, a1) and np.array_equal(v0, v1)
    finally:
        for s in sol:
            s.close()


@pytest.mark.parametrize("first_ring", [0, 1])
def test_batch_register_launch_matches_ring_alone(first_ring):
    """A multi-case band launch takes its sweep form from its first case.  With that case on
    the register-prefetch form (lusgs_ring=0) the FP32-LHS batch runs k_band_multi, with it on
    the ring form k_band_ring_multi; either way every case equals its ring-form run alone."""
    cases = [naca_case(945, 241, m=0.04, alpha=8.0, reynum=3e6, beta=10.0, impsch=2,
                       wall_spacing=2e-6, precision=SMAC_FP32_LHS),
             naca_case(239, 61, m=0.02, alpha=4.0, reynum=1e6, beta=10.0, impsch=2,
                       precision=SMAC_FP32_LHS)]
    print(f"batch with first case lusgs_ring={first_ring}", flush=True)
    ref = [_run(c, {"lusgs_ring": 1}, 4) for c in cases]
    sol = [Solver(cases[0], options={"lusgs_ring": first_ring}),
           Solver(cases[1], options={"lusgs_ring": 1})]
    try:
        with Batch(sol) as b:
            b.run(4)
        for (q0, a0, v0, s0), s in zip(ref, sol):
            q1, a1, v1 = s.get_state()
            assert np.array_equal(q0, q1) and np.array_equal(a0, a1) and np.array_equal(v0, v1)
    finally:
        for s in sol:
            s.close()


@pytest.mark.parametrize("ring", [1, 0])
def test_fp32_sweeps_repeat_bit_identical(ring):
    """The FP32-LHS band sweeps, run many times over, give the same bits every time.  A
    16-byte store of two hand-over words with an SGPR offset, whose data registers the next
    instruction overwrote, used to put the third word in the first now and then (about one
    run in a hundred of three passes at 945x241, both sweep forms); the slots are now written
    one 8-byte word per store.  A run that goes non-finite fails with the pass and the
    places of the non-finite values: a slot word read half written could carry the sentinel's
    NaN high half, which the sweeps now treat as not yet written."""
    case = naca_case(945, 241, m=0.04, alpha=8.0, reynum=3e6, beta=10.0, impsch=2,
                     wall_spacing=2e-6, precision=SMAC_FP32_LHS)
    print(f"repeated FP32-LHS sweeps, lusgs_ring={ring}", flush=True)

    def three_passes():
        with Solver(case, options={"lusgs_ring": ring}) as s:
            for p in range(3):
                try:
                    s.run(1)
                except SmacError as e:   # say where the values went non-finite
                    where = {}
                    for name in ("q", "dq", "s"):
                        j, k, m = np.nonzero(~np.isfinite(s.array(name)))   # (jmax, kmax, 3)
                        where[name] = (len(j), sorted(set(k.tolist()))[:8],
                                       list(zip(j[:4].tolist(), k[:4].tolist(), m[:4].tolist())))
                    raise AssertionError(f"pass {p + 1}: {e}; non-finite (count, rows, "
                                         f"first (j,k,m)): {where}") from e
            return s.get_state()[0]

    ref = three_passes()
    differ = [rep for rep in range(150) if not np.array_equal(ref, three_passes())]
    assert not differ, f"runs {differ} differ from the first"
