"""The committed record of the 945x241 Re = 3e6 LU-SGS run-away (tests/golden/, made by
make_lusgs_945x241_divergence.py from the oracle) is the oracle's: its first passes
reproduce exactly, and the record ends where the state first goes non-finite."""
import json
import os

import pyoracle
from minismac_amd.case import config_case

from conftest import GOLDEN


def test_divergence_record_is_the_oracles():
    with open(os.path.join(GOLDEN, "lusgs_945x241_divergence.json")) as f:
        gold = json.load(f)
    assert gold["first_nonfinite_pass"] == len(gold["passes"]) + 1 == 182
    o = pyoracle.Oracle(config_case(1, impsch=2))
    for row in gold["passes"][:3]:
        st = o.step()
        assert [st[k] for k in gold["keys"]] == row
    # the run-away: the residual grows by orders of magnitude before the end
    rhs = [r[gold["keys"].index("rhsnorm")] for r in gold["passes"]]
    assert rhs[-1] > 1e6 * rhs[0]
