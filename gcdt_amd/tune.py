"""Autotuner for the solve's queue and launch sizes.

Two knobs decide how the round loops map onto the GPU:

* ``tail_threshold`` -- list length at or below which the remaining rounds
  run in one persistent 1024-thread workgroup instead of one launch per phase;
* ``grid_cap`` -- workgroups of a grid-stride launch;
* ``flip_items`` / ``ins_items`` -- list items per thread of the kernels that
  append to a list once per block (more items amortise the append, fewer keep
  more waves in flight);
* ``pivot`` -- which pending point of a triangle an insertion round inserts
  (0: pseudo-random, 1: the one nearest the centroid class first);
* ``single_tester`` -- whether an edge between two active triangles is tested
  by the lower one only.

``tune`` times a workload over a small grid of values by coordinate descent
(one knob at a time, in the order above), keeps the fastest, applies it with
gcdt_set_tunable and caches it per device and workload class in a JSON file
(``$GCDT_TUNE_CACHE`` or ``~/.cache/gcdt_amd/tune.json``); ``load_or_tune``
applies a cached choice without timing.
"""
from __future__ import annotations

import ctypes
import json
import os
import sys
import time

from .lib import check, lib

CANDIDATES = {
    "tail_threshold": [1024, 2048, 4096, 8192, 16384],
    "grid_cap": [2048, 4096, 8192, 16384],
    "flip_items": [1, 2, 4],
    "ins_items": [1, 2, 4],
    "pivot": [0, 1],
    "single_tester": [0, 1],
}
DEFAULTS = {"tail_threshold": 4096, "grid_cap": 4096, "flip_items": 1, "ins_items": 1, "pivot": 1, "single_tester": 1}
# Values the library accepts: those a GPU parity test covers
# (gcdt_amd/csrc/mesh.hip tunable_cleared).  A sweep over the others is
# refused by gcdt_set_tunable with GCDT_ERR_INVALID.
CLEARED = {"tail_threshold": [0, 1024, 2048, 4096, 8192, 16384, 1 << 20], "grid_cap": [2048, 4096, 8192, 16384],
           "flip_items": [1, 2, 4], "ins_items": [1, 2, 4], "pivot": [0, 1], "single_tester": [0, 1]}


def cache_path() -> str:
    p = os.environ.get("GCDT_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "gcdt_amd", "tune.json")


def device_key() -> str:
    buf = ctypes.create_string_buffer(256)
    check(lib().gcdt_device_name(buf, 256), "gcdt_device_name")
    return buf.value.decode()


def get(name: str) -> int:
    v = ctypes.c_int(0)
    check(lib().gcdt_get_tunable(name.encode(), ctypes.byref(v)), "gcdt_get_tunable")
    return v.value


def apply(choice: dict):
    for k, v in choice.items():
        check(lib().gcdt_set_tunable(k.encode(), int(v)), "gcdt_set_tunable")


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def _time(run, repeats):
    """Median of `repeats` timed runs: the flip phase's last rounds vary from
    run to run (a few hundred tail rounds more or less), and a minimum would
    rank settings by their luckiest run."""
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        run()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2]


def _confirm(run, start: dict, choice: dict, repeats: int, min_gain: float):
    """The descent's end point against its starting point, timed in
    alternation (A B A B ...) so that drift and bimodal run times weigh on both
    alike; the choice stands only if its median is at least min_gain below the
    start's.  Returns (kept choice, its median seconds, the start's)."""
    ta, tb = [], []
    for _ in range(repeats):
        for cfg, ts in ((start, ta), (choice, tb)):
            apply(cfg)
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
    ta.sort()
    tb.sort()
    ma, mb = ta[len(ta) // 2], tb[len(tb) // 2]
    return (choice, mb, ma) if mb < (1.0 - min_gain) * ma else (start, ma, ma)


def tune(kind: str, run, candidates=None, repeats: int = 5, key: str | None = None, path: str | None = None,
         min_gain: float = 0.02, confirm_repeats: int = 9):
    """Time `run()` (one complete solve, synchronised) over the candidates and
    apply the fastest; a knob leaves its current value only for a candidate
    at least `min_gain` faster (run-to-run noise is ~1-2%).  The end point
    is then timed against the starting values in alternation and kept only if
    it is still at least `min_gain` faster (single trials of a solve whose
    flip tail is bimodal can look faster by chance).  Returns
    {"choice": {...}, "ms": best, "trials": [...], "confirmed": bool}."""
    candidates = candidates or CANDIDATES
    choice = {k: get(k) for k in candidates}
    start = dict(choice)
    trials = []
    run()  # warm code objects and the pool at the current setting
    for name, values in candidates.items():
        times = {}
        for v in values:
            trial = dict(choice, **{name: v})
            apply(trial)
            # named before it runs, so a failing trial is identified by the
            # last line a run printed
            print(f"# tune {kind}: {trial}", file=sys.stderr, flush=True)
            t = _time(run, repeats)
            trials.append({**trial, "ms": round(t * 1000.0, 3)})
            times[v] = t
        cur = choice[name]
        if cur not in times:
            apply(choice)
            times[cur] = _time(run, repeats)
            trials.append({**choice, "ms": round(times[cur] * 1000.0, 3)})
        best_v = min(times, key=times.get)
        if times[best_v] < (1.0 - min_gain) * times[cur]:
            choice[name] = best_v
    confirmed = True
    if choice != start and confirm_repeats > 0:
        kept, t_kept, t_start = _confirm(run, start, choice, confirm_repeats, min_gain)
        confirmed = kept == choice
        trials.append({"confirm": True, "start_ms": round(t_start * 1000.0, 3), "choice_ms": round(t_kept * 1000.0, 3)
                       if confirmed else None})
        choice = dict(kept)
    apply(choice)
    best_ms = min((t["ms"] for t in trials if "ms" in t and all(t[k] == choice[k] for k in choice)), default=None)
    res = {"choice": choice, "ms": best_ms, "trials": trials, "confirmed": confirmed}
    path = path or cache_path()
    data = _load(path)
    data.setdefault(key or device_key(), {})[kind] = {"choice": choice, "ms": best_ms}
    _store(path, data)
    return res


def load_or_tune(kind: str, run, force: bool = False, **kw):
    """Apply the cached choice for (device, kind), or tune and cache one."""
    path = kw.pop("path", None) or cache_path()
    key = kw.pop("key", None) or device_key()
    if not force:
        hit = _load(path).get(key, {}).get(kind)
        if hit:
            apply(hit["choice"])
            return {"choice": hit["choice"], "ms": hit.get("ms"), "cached": True}
    res = tune(kind, run, key=key, path=path, **kw)
    res["cached"] = False
    return res
