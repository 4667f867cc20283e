"""Autotuner logic on the CPU: a stand-in workload whose run time depends on
the tunables; the tuner must find its minimum, apply it, and cache it."""
import json
import math
import time

import pytest

from gcdt_amd import GcdtError, get_tunable, set_tunable
from gcdt_amd import tune


@pytest.fixture(autouse=True)
def _restore():
    saved = {k: get_tunable(k) for k in tune.CANDIDATES}
    yield
    tune.apply(saved)


def _fake_run(calls):
    def run():
        calls.append((get_tunable("tail_threshold"), get_tunable("grid_cap")))
        tt, gc = calls[-1]
        cost = abs(math.log2(tt) - 13) + abs(math.log2(gc) - 14)   # best at 8192 / 16384
        time.sleep(0.002 * cost + 0.001)
    return run


def test_tuner_finds_minimum_and_caches(tmp_path):
    path = str(tmp_path / "tune.json")
    calls = []
    knobs = {"tail_threshold": tune.CANDIDATES["tail_threshold"], "grid_cap": tune.CANDIDATES["grid_cap"]}
    res = tune.tune("unit", _fake_run(calls), candidates=knobs, repeats=1, key="cpu-test", path=path)
    assert res["choice"] == {"tail_threshold": 8192, "grid_cap": 16384}
    assert get_tunable("tail_threshold") == 8192 and get_tunable("grid_cap") == 16384
    data = json.load(open(path))
    assert data["cpu-test"]["unit"]["choice"] == res["choice"]
    # a later process applies the cached choice without timing anything
    tune.apply({"tail_threshold": 4096, "grid_cap": 4096})
    calls2 = []
    hit = tune.load_or_tune("unit", _fake_run(calls2), key="cpu-test", path=path)
    assert hit["cached"] and not calls2
    assert get_tunable("tail_threshold") == 8192


def test_tunable_ranges_are_checked():
    with pytest.raises(GcdtError, match="grid_cap"):
        set_tunable("grid_cap", 10)
    with pytest.raises(GcdtError, match="unknown"):
        set_tunable("no_such_knob", 1)
    set_tunable("grid_cap", 8192)
    assert get_tunable("grid_cap") == 8192
    with pytest.raises(GcdtError, match="1, 2 or 4"):
        set_tunable("flip_items", 3)
    set_tunable("flip_items", 2)
    assert get_tunable("flip_items") == 2
