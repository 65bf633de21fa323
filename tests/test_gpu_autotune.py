"""Autotuner on the device: it returns settings from its search space, caches them, and the
settings it picks leave the results bit-identical to the library defaults."""
import numpy as np
import pytest

from minismac_amd import autotune
from minismac_amd.case import config_case
from minismac_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _state_after(case, opts, n=5):
    with Solver(case, options=opts) as s:
        s.run(n)
        return s.get_state(), s.stats()


@pytest.mark.parametrize("impsch", [2, 1])
def test_tuned_settings_are_valid_cached_and_exact(tmp_path, impsch):
    case = config_case(0, impsch=impsch)
    path = str(tmp_path / "tune.json")
    print(f"autotune 119x31 impsch={impsch}", flush=True)
    opts = autotune.tune(case, passes=3, repeats=1, sweeps=1, path=path)
    space = dict(autotune._space(case))
    assert set(opts) == set(space)
    for k, v in opts.items():
        assert v in space[k]
    assert autotune._load(path)[autotune.key(case)]["options"] == opts
    # any setting of the knobs gives the default results bit for bit
    extreme = {k: min(v) for k, v in space.items()}
    (q0, a0, v0), s0 = _state_after(case, {})
    for o in (opts, extreme):
        (q1, a1, v1), s1 = _state_after(case, o)
        assert np.array_equal(q0, q1) and np.array_equal(a0, a1) and np.array_equal(v0, v1), o
        assert s0["rhsnorm"] == s1["rhsnorm"] and s0["cl"] == s1["cl"], o
