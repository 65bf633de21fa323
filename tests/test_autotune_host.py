"""Autotuner host logic (no GPU): the search space per scheme, the cache key and a cache
hit that needs no device."""
import json

from minismac_amd import autotune
from minismac_amd._abi import SMAC_FP32_LHS
from minismac_amd.case import config_case


def test_search_space_follows_the_scheme():
    lu = dict(autotune._space(config_case(0, impsch=2)))
    assert set(lu) == {"lusgs_ring", "lusgs_rows", "lusgs_prefetch", "lusgs_prefetch_sa"}
    assert all(1 <= r <= 64 for r in lu["lusgs_rows"])
    assert set(lu["lusgs_prefetch"]) <= {1, 2, 3, 4, 6, 8}
    laminar = dict(autotune._space(config_case(0, impsch=2, iturb=0)))
    assert set(laminar) == {"lusgs_ring", "lusgs_rows", "lusgs_prefetch"}
    assert set(dict(autotune._space(config_case(0, impsch=1)))) == {"lr_lines"}


def test_key_separates_size_scheme_and_precision():
    a = autotune.key(config_case(0, impsch=2))
    assert a != autotune.key(config_case(0, impsch=1))
    assert a != autotune.key(config_case(0, impsch=2, precision=SMAC_FP32_LHS))
    assert "119x31" in a


def test_cached_options_are_returned_without_tuning(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    case = config_case(0, impsch=2)
    path.write_text(json.dumps({autotune.key(case): {"options": {"lusgs_rows": 32,
                                                                 "lusgs_prefetch": 6}}}))

    def no_probe(*a, **k):
        raise AssertionError("a cache hit must not start a tuning run")

    monkeypatch.setattr(autotune, "_Probe", no_probe)
    assert autotune.tune(case, path=str(path)) == {"lusgs_rows": 32, "lusgs_prefetch": 6}
    # tune_cases tunes each distinct shape once and hands every case its options
    opts = autotune.tune_cases([case, config_case(0, impsch=2, alpha=2.0)], path=str(path))
    assert opts == [{"lusgs_rows": 32, "lusgs_prefetch": 6}] * 2


def test_store_is_atomic_json(tmp_path):
    path = tmp_path / "sub" / "t.json"
    autotune._store(str(path), {"k": {"options": {"lr_lines": 16}}})
    assert autotune._load(str(path)) == {"k": {"options": {"lr_lines": 16}}}
    assert autotune._load(str(tmp_path / "absent.json")) == {}


def test_batch_case_mix_is_seeded_and_in_range():
    """The batched workload's case mix: reproducible for a seed, parameters inside their
    ranges, sizes from the three grids."""
    import numpy as np
    from minismac_amd.case import BATCH_SIZES, batch_cases
    a = batch_cases(8, seed=3)
    b = batch_cases(8, seed=3)
    c = batch_cases(8, seed=4)
    assert [x.name for x in a] == [x.name for x in b]
    assert [x.name for x in a] != [x.name for x in c]
    sizes = {(jm, km) for jm, km, _ in BATCH_SIZES}
    for x, y in zip(a, b):
        assert (x.jmax, x.kmax) in sizes
        assert -4.0 <= x.params.alpha <= 16.0
        assert 1e5 <= x.params.reynum <= 1e7
        assert np.array_equal(x.x, y.x) and np.array_equal(x.y, y.y)
