"""Autotuner logic on the CPU: a stand-in workload whose run time depends on
the tunables; the tuner must find its minimum, apply it, and cache it.  The
tuner's logic runs against an in-memory stand-in for the library's tunables,
since the library accepts only values that a GPU parity test covers."""
import json
import math
import time

import pytest

from gcdt_amd import GcdtError, get_tunable, set_tunable
from gcdt_amd import tune


@pytest.fixture()
def fake_knobs(monkeypatch):
    """tune.get / tune.apply over a dict instead of the library."""
    knobs = dict(tune.DEFAULTS)
    monkeypatch.setattr(tune, "get", lambda name: knobs[name])
    monkeypatch.setattr(tune, "apply", lambda choice: knobs.update({k: int(v) for k, v in choice.items()}))
    return knobs


def _fake_run(calls, knobs):
    def run():
        calls.append((knobs["tail_threshold"], knobs["grid_cap"]))
        tt, gc = calls[-1]
        cost = abs(math.log2(tt) - 13) + abs(math.log2(gc) - 14)   # best at 8192 / 16384
        time.sleep(0.002 * cost + 0.001)
    return run


def test_tuner_finds_minimum_and_caches(tmp_path, fake_knobs):
    path = str(tmp_path / "tune.json")
    calls = []
    cand = {"tail_threshold": [1024, 2048, 4096, 8192, 16384], "grid_cap": [2048, 4096, 8192, 16384]}
    res = tune.tune("unit", _fake_run(calls, fake_knobs), candidates=cand, repeats=1, key="cpu-test", path=path)
    assert res["choice"] == {"tail_threshold": 8192, "grid_cap": 16384}
    assert fake_knobs["tail_threshold"] == 8192 and fake_knobs["grid_cap"] == 16384
    data = json.load(open(path))
    assert data["cpu-test"]["unit"]["choice"] == res["choice"]
    # a later process applies the cached choice without timing anything
    tune.apply({"tail_threshold": 4096, "grid_cap": 4096})
    calls2 = []
    hit = tune.load_or_tune("unit", _fake_run(calls2, fake_knobs), key="cpu-test", path=path)
    assert hit["cached"] and not calls2
    assert fake_knobs["tail_threshold"] == 8192


def test_tunable_ranges_are_checked():
    with pytest.raises(GcdtError, match="grid_cap"):
        set_tunable("grid_cap", 10)
    with pytest.raises(GcdtError, match="unknown"):
        set_tunable("no_such_knob", 1)
    with pytest.raises(GcdtError, match="1, 2 or 4"):
        set_tunable("flip_items", 3)


def test_only_gpu_tested_values_are_accepted():
    """Values without a passing GPU parity test (tests/test_gpu_tunables.py
    runs every candidate) are refused; the covered ones are accepted, and
    every candidate the tuner tries is covered."""
    saved = {k: get_tunable(k) for k in tune.DEFAULTS}
    try:
        for k, v in tune.CLEARED.items():
            for x in v:
                set_tunable(k, x)
                assert get_tunable(k) == x
        for k, v in tune.CANDIDATES.items():
            assert set(v) <= set(tune.CLEARED[k]), k
        for k, x in (("tail_threshold", 3000), ("grid_cap", 1000), ("grid_cap", 65536), ("tail_threshold", 64)):
            with pytest.raises(GcdtError, match="not covered"):
                set_tunable(k, x)
    finally:
        for k, v in saved.items():
            set_tunable(k, v)


def test_a_gain_that_does_not_hold_is_not_kept(tmp_path, fake_knobs):
    """A value that timed faster during the descent but not when timed again
    against the starting values (in alternation) is dropped: the tuner keeps
    the start."""
    calls = [0]
    cand = {"tail_threshold": [1024, 4096, 8192]}

    def run():
        calls[0] += 1
        lucky = calls[0] <= 16 and fake_knobs["tail_threshold"] == 8192  # only while descending
        time.sleep(0.001 if lucky else 0.004)

    res = tune.tune("unit", run, candidates=cand, repeats=3, key="cpu-test", path=str(tmp_path / "t.json"))
    assert res["choice"] == {"tail_threshold": 4096} and not res["confirmed"]
    assert fake_knobs["tail_threshold"] == 4096
