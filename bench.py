#!/usr/bin/env python3
"""Benchmark: full constrained Delaunay triangulation on one MI355X.

Workload (one step): the 10M-point plain CDT -- 10,000,000 int32 draws in
[0,2^20)^2, 80% uniform and 20% Gaussian clusters (sigma 2^12, 64 centres),
rounded and deduplicated, plus 1,000 disjoint circles, each a 100-gon with
vertices snapped to the grid (100,000 constraint segments).  Inputs are
generated once (seeded SplitMix64), mapped exactly onto the 2^30 grid and
uploaded; each timed step is one gcdt_solve from device-resident inputs to a
device-resident triangle list.

Prints one JSON line: value = Mpts/s (points / solve wall time), roofline of
the dominant kernel (live hipEvent timing inside this process), the exact
validity of the last solve, and the CPU baseline (the serial exact-predicate
oracle on the same workload, one host core).

The oracle runs only in the CPU leg (skipped with --no-cpu): it is timed as
cpu_baseline, and its exact validator checks GPU results after their timed
regions (skipped with --no-validate).  No timed region calls it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def _checker(args):
    """The CPU leg's exact validator (the oracle module), or None when the CPU
    leg or validation is off.  Called after timed regions only."""
    if args.no_cpu or args.no_validate:
        return None
    from oracle import oracle as O
    return O

# Algorithmic bytes per processed list item of the instrumented kernels.
#   flip_test : tri 12 + adj 12 + cflag 4 + own 3 points 24
#               + 3 x (neighbour tri 12 + opposite point 8) + cand 4 + 2 claims 16
#   ins_reloc : loc 4 + sons 4 + list write 4 + 2 children tri 24 + 3 points 24 + loc write 4
#               + next round's claim: key write 8 + claim atomic 8 + claimed-triangle append 4
#   ins_claim : loc 4 + point id 4 + 1.x claim atomics 8 (first round of a run only)
#   ins_split : per claimed triangle: list 4 + claim 8 + winner loc 4 + key 8 + tri/adj 24
#               + 3 children x (tri 12 + adj 12) written + sons 4 + loc 4
#   flip_apply: per proposal meta 4 + own claim 8; a winning proposal (92.2% of them on
#               the 10M workload) adds its own vertices, neighbours and coordinates 48,
#               the neighbour record's 60, both records rewritten 2 x 56 and 4 list
#               entries 16; a losing one the winner's claim 8, its mark 4 and a list
#               entry 4: 12 + 0.922 x 236 + 0.078 x 16 = 231
#   locate_bitmap: point 8 + isv 4 + 2 offsets 8 + ~1.5 x (tri id 4 + tri 12 + 3 points 24) + loc 4 + list 4
# With the packed records, each flip_test launch also settles the previous
# round's flips (listed as flip_test_fix items): per settled triangle, 2 x
# (neighbour adj 12 + stamp 4 + meta 4) read + 2 own slots and 2 back-pointers
# written (16); its other reads are the test's own.
BYTES_PER_FIX_IN_TEST = 56
BYTES_PER_ITEM = {
    "flip_test": 132,
    "ins_reloc": 84,
    "ins_claim": 16,
    "flip_apply": 231,
    "locate_bitmap": 88,
    "ins_split": 128,
}


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "cdt_10M_clustered_hbm_traffic.md")


def _traffic_per_item():
    """Measured HBM bytes per list item (rocprofv3 FETCH_SIZE + WRITE_SIZE,
    scaled by a 1 GiB copy) from the committed profile summary."""
    out = {}
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_item "):
                    _, k, v = line.split()
                    out[k] = float(v)
    except OSError:
        pass
    return out


def _prof(lib, name):
    import ctypes
    ms = ctypes.c_double(0)
    launches = ctypes.c_int64(0)
    items = ctypes.c_int64(0)
    if lib.gcdt_prof_read(name.encode(), ctypes.byref(ms), ctypes.byref(launches), ctypes.byref(items)) != 0:
        return None
    return ms.value, launches.value, items.value


def main(argv=None):
    import faulthandler
    faulthandler.enable()  # a crash prints every thread's Python stack
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--circles", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--batch-count", type=int, default=4096, help="instances in the batched workload (0 = skip)")
    ap.add_argument("--per-instance", type=int, default=256,
                    help="instances timed with one plain solve each, for the batched comparison")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="before the solves, copy MIB MiB device-to-device 3 times (a plain copy kernel of "
                         "known size, to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings)")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tunables (no autotuner sweep)")
    ap.add_argument("--tune", action="store_true",
                    help="run the autotuner sweep even when a cached choice exists for this device")
    ap.add_argument("--no-degenerate", action="store_true", help="skip the 4096^2 lattice validity run")
    ap.add_argument("--serve-count", type=int, default=256, help="requests in the serving run (0 = skip)")
    ap.add_argument("--serve-qps", type=float, default=50.0, help="Poisson arrival rate of the serving run")
    ap.add_argument("--serve-workers", type=int, default=3, help="worker streams of the serving run")
    ap.add_argument("--serve-small-n", type=int, default=65536,
                    help="requests up to this many points are merged into batched solves")
    args = ap.parse_args(argv)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X only: this CDT path has no multi-GPU code (the sharding design "
              "is documented, not built). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from gcdt_amd import DeviceBuffer, gen, lib, solve_device
    from gcdt_amd.lib import check

    L = lib()
    t0 = time.time()
    pts, cons = gen.clustered_cdt(n=args.points, circles=args.circles)
    g = gen.to_grid(pts)
    n, c = len(g), len(cons)
    print(f"# workload: {n} points, {c} segments (generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)

    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        with DeviceBuffer(nb) as a, DeviceBuffer(nb) as b:
            check(L.gcdt_memset(a.ptr, 1, nb), "memset")
            for _ in range(3):
                check(L.gcdt_memcpy_d2d(b.ptr, a.ptr, nb), "copy")
            check(L.gcdt_synchronize(), "sync")
    dp = DeviceBuffer.from_array(g)
    dc = DeviceBuffer.from_array(cons)
    dt = DeviceBuffer(12 * (2 * n + 2))
    da = DeviceBuffer(12 * (2 * n + 2))
    tuned = None
    if not args.no_tune:
        from gcdt_amd import tune

        def run_large():
            solve_device(dp, n, dc, c, dt, da)

        tuned = tune.load_or_tune("large", run_large, force=args.tune)
        print(f"# tunables (large): {tuned}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        T, st = solve_device(dp, n, dc, c, dt, da)
        print(f"# warmup {i}: {st['ms_total']:.1f} ms", file=sys.stderr, flush=True)
    check(L.gcdt_synchronize(), "sync")
    import ctypes
    # one solve with every instrumented kernel timed: the per-kernel table and
    # the dominant kernel (among those with a per-item byte model)
    check(L.gcdt_prof_reset(), "prof")
    check(L.gcdt_prof_only(None), "prof")
    check(L.gcdt_prof_enable(1), "prof")
    solve_device(dp, n, dc, c, dt, da)
    check(L.gcdt_synchronize(), "sync")
    check(L.gcdt_prof_enable(0), "prof")
    buf = ctypes.create_string_buffer(4096)
    L.gcdt_prof_names(buf, 4096)
    kern = {}
    for name in [x for x in buf.value.decode().split(";") if x]:
        r = _prof(L, name)
        if r:
            kern[name] = r
    kernels = {k: {"ms": round(v[0], 3), "launches": v[1], "items": v[2]} for k, v in kern.items()}
    modeled = [k for k in kern if k in BYTES_PER_ITEM]
    dom = max(modeled, key=lambda k: kern[k][0]) if modeled else None
    # timed steps: only the dominant kernel carries events (few markers in the
    # stream), so `value` is the plain solve rate and the kernel's average
    # launch time is live from the same timed region
    check(L.gcdt_prof_reset(), "prof")
    if dom:
        check(L.gcdt_prof_only(dom.encode()), "prof")
        check(L.gcdt_prof_enable(1), "prof")
    steps = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        T, st = solve_device(dp, n, dc, c, dt, da)
        steps.append(st)
        print(f"# step {i}: {st['ms_total']:.1f} ms (insert {st['ms_insert']:.1f}, flip {st['ms_flip']:.1f}, "
              f"cons {st['ms_constraints']:.1f}, sample {st['ms_sample']:.1f}, bitmap {st['ms_bitmap']:.1f}; "
              f"rounds ins {st['insert_rounds']} flip {st['flip_rounds']} cons {st['constraint_rounds']})",
              file=sys.stderr, flush=True)
    check(L.gcdt_synchronize(), "sync")
    t_end = time.perf_counter()
    check(L.gcdt_prof_enable(0), "prof")
    ms_per_step = (t_end - t_start) * 1000.0 / args.steps
    mpts = n / (ms_per_step / 1000.0) / 1e6

    roof = None
    if dom:
        ms, launches, items = _prof(L, dom)
        check(L.gcdt_prof_only(None), "prof")
        avg_ms = ms / max(launches, 1)
        fix = _prof(L, "flip_test_fix") if dom == "flip_test" else None
        fix_items = fix[2] if fix else 0
        bytes_per_launch = (BYTES_PER_ITEM[dom] * items + BYTES_PER_FIX_IN_TEST * fix_items) / max(launches, 1)
        achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9 if avg_ms > 0 else 0.0
        tpi = _traffic_per_item().get(dom)
        traffic = round(tpi * items / max(launches, 1)) if tpi else None
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "traffic_GBs": round(traffic / (avg_ms / 1000.0) / 1e9, 1) if traffic and avg_ms > 0 else None,
                "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT) if traffic else None,
                "avg_launch_us": round(avg_ms * 1000.0, 3), "launches_per_step": launches // args.steps,
                "kernel_ms_per_step": round(ms / args.steps, 3),
                "bytes_per_item": BYTES_PER_ITEM[dom], "items_per_launch": items // max(launches, 1),
                "settled_flips_per_launch": fix_items // max(launches, 1),
                "bytes_per_settled_flip": BYTES_PER_FIX_IN_TEST if fix_items else None}
    # whole-solve algorithmic bytes: points in, segments in, triangles out
    b_alg = 8 * n + 8 * c + 12 * T
    solve_roof = {"bytes": b_alg, "achieved_GBs": round(b_alg / (ms_per_step / 1000.0) / 1e9, 3),
                  "frac": round(b_alg / (ms_per_step / 1000.0) / 1e9 / HBM_PEAK_GBS, 6)}

    import ctypes as _ct
    in_use, cached, peak = _ct.c_size_t(0), _ct.c_size_t(0), _ct.c_size_t(0)
    check(L.gcdt_pool_stats(_ct.byref(in_use), _ct.byref(cached), _ct.byref(peak)), "pool")
    pool_peak = int(peak.value)
    # the same solve from host arrays (PCIe in and out): reported, never `value`
    from gcdt_amd import solve as solve_host
    th = time.perf_counter()
    solve_host(g, cons)
    ms_host = (time.perf_counter() - th) * 1000.0
    pcie = {"ms": round(ms_host, 3), "mpts_per_s": round(n / (ms_host / 1000.0) / 1e6, 3),
            "note": "host points/segments in, host triangles out (gcdt_amd.solve)"}

    valid = None
    O = _checker(args)
    if O is not None:
        tris = dt.to_array(np.int32, (T, 3))
        v = O.validate(g, tris, cons)
        valid = {"valid": v["valid"], "not_delaunay": v["not_delaunay"],
                 "missing_constraints": v["missing_constraints"], "cocircular_edges": v["cocircular"]}
        print(f"# validation: {v}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        tc = time.perf_counter()
        O.cdt(g, cons)
        dtc = time.perf_counter() - tc
        cpu = {"value": round(n / dtc / 1e6, 4), "unit": "Mpts/s", "cores": 1, "kind": "port", "host": _host_cpus(1),
               "sample": f"the full workload ({n} points, {c} segments), serial exact-predicate CDT, "
                         f"one host core, {dtc:.2f} s"}

    degenerate = None if args.no_degenerate else bench_degenerate(args)
    batched = bench_batched(args) if args.batch_count > 0 else None
    serving = bench_serving(args) if args.serve_count > 0 else None

    out = {
        "metric": "Mpts/s CDT at 10M pts; batched CDTs/s vs per-instance launch; p99 latency",
        "value": round(mpts, 4),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64/int128 (exact integer predicates)",
        "data": "synthetic (seeded SplitMix64 generator)",
        "config": {"workload": "plain CDT, large: 10M int32 points (80% uniform, 20% Gaussian clusters "
                               "sigma=2^12, 64 centres) + 1000 disjoint 100-gon circles (100k segments)",
                   "points": n, "segments": c, "triangles": T, "parallelism": "single GPU",
                   "tunables": tuned["choice"] if tuned else "defaults"},
        "roofline": roof,
        "solve_roofline": solve_roof,
        "pcie_inclusive": pcie,
        "pool_peak_bytes": pool_peak,
        "kernels_one_profiled_solve": kernels,
        "validity": valid,
        "last_step": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in steps[-1].items()},
        "cpu_baseline": cpu,
        "degenerate": degenerate,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    for b in (dp, dc, dt, da):
        b.free()


def bench_batched(args):
    """Batched workload: `batch_count` star-polygon instances (point counts
    log-uniform in [64, 65536], 0-4 holes) solved in one gcdt_solve_batch,
    against one plain gcdt_solve per instance on the first `per_instance`."""
    import numpy as np

    from gcdt_amd import DeviceBuffer, gen, pack_instances, solve_batch_device, solve_device
    from gcdt_amd.lib import check, lib

    L = lib()
    t0 = time.time()
    raw = gen.batched_instances(count=args.batch_count)
    inst = [(gen.to_grid(p), c) for p, c in raw]
    flat_p, poff, flat_c, coff, toff = pack_instances(inst)
    print(f"# batched workload: {len(inst)} instances, {len(flat_p)} points, {len(flat_c)} segments "
          f"(generated in {time.time() - t0:.1f}s)", file=sys.stderr, flush=True)
    dp = DeviceBuffer.from_array(flat_p)
    dc = DeviceBuffer.from_array(flat_c)
    dt = DeviceBuffer(12 * int(toff[-1]))
    tuned_b = tuned_s = None
    if not args.no_tune:
        from gcdt_amd import tune

        tuned_b = tune.load_or_tune("batched", lambda: solve_batch_device(dp, poff, dc, coff, dt), force=args.tune)
        print(f"# tunables (batched): {tuned_b}", file=sys.stderr, flush=True)
    # the pool's peak over one solve, this workload's inputs and output included
    check(L.gcdt_synchronize(), "sync")
    check(L.gcdt_pool_reset_peak(), "pool")
    st = solve_batch_device(dp, poff, dc, coff, dt)  # warmup
    import ctypes as _ct
    _u, _c, _pk = _ct.c_size_t(0), _ct.c_size_t(0), _ct.c_size_t(0)
    check(L.gcdt_pool_stats(_ct.byref(_u), _ct.byref(_c), _ct.byref(_pk)), "pool")
    pool_peak = int(_pk.value)
    reps = 2
    check(L.gcdt_synchronize(), "sync")
    tb = time.perf_counter()
    for _ in range(reps):
        st = solve_batch_device(dp, poff, dc, coff, dt)
    check(L.gcdt_synchronize(), "sync")
    ms_batch = (time.perf_counter() - tb) * 1000.0 / reps
    print(f"# batched solve: {ms_batch:.1f} ms ({st})", file=sys.stderr, flush=True)
    allt = dt.to_array(np.int32, (int(toff[-1]), 3))
    results = [allt[toff[i]:toff[i + 1]] for i in range(len(inst))]
    for b in (dp, dc, dt):
        b.free()
    # one plain solve per instance
    m = min(args.per_instance, len(inst))
    bufs = []
    for g, c in inst[:m]:
        bufs.append((DeviceBuffer.from_array(g), DeviceBuffer.from_array(c) if len(c) else None,
                     DeviceBuffer(12 * (2 * len(g) + 2)), DeviceBuffer(12 * (2 * len(g) + 2)), len(g), len(c)))
    if not args.no_tune:
        from gcdt_amd import tune

        def run_small():
            for dpi, dci, dti, dai, nn, cc in bufs[:16]:
                solve_device(dpi, nn, dci, cc, dti, dai)

        tuned_s = tune.load_or_tune("per_instance", run_small, force=args.tune)
        print(f"# tunables (per-instance): {tuned_s}", file=sys.stderr, flush=True)
    solve_device(bufs[0][0], bufs[0][4], bufs[0][1], bufs[0][5], bufs[0][2], bufs[0][3])  # warmup
    check(L.gcdt_synchronize(), "sync")
    ts = time.perf_counter()
    for dpi, dci, dti, dai, n, c in bufs:
        solve_device(dpi, n, dci, c, dti, dai)
    check(L.gcdt_synchronize(), "sync")
    ms_single = (time.perf_counter() - ts) * 1000.0
    pts_single = sum(b[4] for b in bufs)
    for b in bufs:
        for x in b[:4]:
            if x is not None:
                x.free()
    batched_cdts = len(inst) / (ms_batch / 1000.0)
    single_cdts = m / (ms_single / 1000.0)
    cpu = None
    refs = None
    if not args.no_cpu:
        rate, thr, secs, refs = _cpu_pool_cdts(inst)
        cpu = {"value": round(rate, 1), "unit": "CDTs/s", "cores": thr, "kind": "port", "host": _host_cpus(thr),
               "sample": f"all {len(inst)} instances, one per host thread, serial exact-predicate CDT, {secs:.2f} s"}
    exact = _check_all(_checker(args), inst, results, refs)
    # the per-instance rate is measured on the first m instances; scale by point
    # counts too so an unrepresentative subset is visible
    return {"instances": len(inst), "points": int(len(flat_p)), "segments": int(len(flat_c)),
            "ms_batched": round(ms_batch, 3), "batched_cdts_per_s": round(batched_cdts, 1),
            "batched_mpts_per_s": round(len(flat_p) / (ms_batch / 1000.0) / 1e6, 3),
            "per_instance_subset": m, "per_instance_ms": round(ms_single, 3),
            "per_instance_cdts_per_s": round(single_cdts, 1),
            "per_instance_mpts_per_s": round(pts_single / (ms_single / 1000.0) / 1e6, 3),
            "speedup_cdts": round(batched_cdts / single_cdts, 2),
            "rounds": {"insert": st["insert_rounds"], "flip": st["flip_rounds"],
                       "constraints": st["constraint_rounds"]},
            "pool_peak_bytes": pool_peak, "pool_peak_bytes_per_point": round(pool_peak / max(len(flat_p), 1), 1),
            "exactness": exact, "cpu_baseline": cpu,
            "tunables": {"batched": tuned_b["choice"] if tuned_b else "defaults",
                         "per_instance": tuned_s["choice"] if tuned_s else "defaults"}}


# Host threads of the CPU legs: the host cores this run may use, at most 16
# (the share of a one-GPU machine's host; nproc there may report more).
CPU_THREADS = 16


def _host_cpus(used):
    """nproc, the CPU model and the threads the CPU leg used."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model, "threads_used": used}


def _cpu_pool_cdts(instances, max_threads=CPU_THREADS):
    """The serial oracle over every instance, one instance per host thread
    (the ctypes call releases the GIL).  Returns (CDTs/s, threads, seconds,
    the oracle's triangle lists)."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    threads = max(1, min(max_threads, os.cpu_count() or 1))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        refs = [r[0] for r in ex.map(lambda gc: O.cdt(gc[0], gc[1]), instances)]
    dt = time.perf_counter() - t0
    return len(instances) / dt, threads, dt, refs


def _check_all(O, instances, results, refs=None, max_threads=CPU_THREADS):
    """Every GPU result through the oracle's exact validator, and compared
    with the oracle's canonical triangle set where the CDT is unique (refs:
    the oracle's triangles per instance, when the CPU leg computed them).
    Runs after the timed regions."""
    if O is None:
        return None
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    def one(i):
        (g, c), tris = instances[i], results[i]
        if tris.shape != (2 * len(g) + 2, 3):
            return "invalid", False
        v = O.validate(g, tris, c)
        if not v["valid"]:
            return "invalid", False
        if not v["unique"]:
            return "valid", False
        ref = refs[i] if refs is not None else O.cdt(g, c)[0]
        return ("equal" if np.array_equal(O.canonical(tris), O.canonical(ref)) else "differs"), True

    threads = max(1, min(max_threads, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(one, range(len(instances))))
    return {"checked": len(res), "invalid": sum(r == "invalid" for r, _ in res),
            "unique": sum(u for _, u in res), "equal_to_oracle": sum(r == "equal" for r, _ in res),
            "differ_from_oracle": sum(r == "differs" for r, _ in res)}


def bench_degenerate(args):
    """Degenerate stress: the 4096 x 4096 lattice (16.7M points, spacing 256)
    plus 2,000 diagonal segments through collinear lattice points; one warmup
    and one timed solve, then the exact validator (not unique: cocircular
    everywhere, so validity is the check)."""
    import numpy as np

    from gcdt_amd import DeviceBuffer, gen, solve_device

    O = _checker(args)
    t0 = time.time()
    pts, cons = gen.lattice_cdt()
    g = gen.to_grid(pts)
    n, c = len(g), len(cons)
    print(f"# degenerate workload: {n} points, {c} segments (generated in {time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    dp = DeviceBuffer.from_array(g)
    dc = DeviceBuffer.from_array(cons)
    dt = DeviceBuffer(12 * (2 * n + 2))
    da = DeviceBuffer(12 * (2 * n + 2))
    tuned = None
    try:
        if not args.no_tune:
            from gcdt_amd import tune
            tuned = tune.load_or_tune("degenerate", lambda: solve_device(dp, n, dc, c, dt, da), repeats=3,
                                      force=args.tune)
            print(f"# tunables (degenerate): {tuned}", file=sys.stderr, flush=True)
        solve_device(dp, n, dc, c, dt, da)
        T, st = solve_device(dp, n, dc, c, dt, da)
        tris = dt.to_array(np.int32, (T, 3))
    finally:
        for b in (dp, dc, dt, da):
            b.free()
    print(f"# degenerate solve: {st['ms_total']:.1f} ms ({st})", file=sys.stderr, flush=True)
    v = O.validate(g, tris, cons) if O is not None else {k: None for k in ("valid", "not_delaunay",
                                                                            "missing_constraints", "cocircular")}
    print(f"# degenerate validation: {v}", file=sys.stderr, flush=True)
    return {"workload": "4096x4096 lattice (spacing 256) + 2000 diagonal segments", "points": n, "segments": c,
            "ms": round(st["ms_total"], 3), "mpts_per_s": round(n / (st["ms_total"] / 1000.0) / 1e6, 3),
            "segments_after_split": st["segments"], "flips": st["flips"],
            "rounds": {"insert": st["insert_rounds"], "flip": st["flip_rounds"], "constraints": st["constraint_rounds"]},
            "valid": v["valid"], "not_delaunay": v["not_delaunay"], "missing_constraints": v["missing_constraints"],
            "cocircular_edges": v["cocircular"], "tunables": tuned["choice"] if tuned else "defaults"}


def bench_serving(args):
    """Serving workload: `serve_count` requests (log-uniform 1e3..1e6 points,
    polygon boundaries of 1-5% of the points) arriving as a Poisson process at
    `serve_qps`; worker streams merge small requests into batched solves.
    Reports throughput and p50/p99 latency from arrival to result on device."""
    import numpy as np

    from gcdt_amd import gen
    from gcdt_amd import serve as S

    # the library's default tunables: serving crashed the host process once
    # (SIGSEGV, no GPU fault) while the per-instance comparison's tuned values
    # (tail_threshold 2048, grid_cap 16384) were applied; the cause is open
    from gcdt_amd import tune
    tune.apply(tune.DEFAULTS)
    t0 = time.time()
    raw = gen.serving_requests(count=args.serve_count, qps=args.serve_qps)
    reqs = [S.Request(i, t, gen.to_grid(p), c) for i, (t, p, c) in enumerate(raw)]
    npts = sum(r.n for r in reqs)
    print(f"# serving workload: {len(reqs)} requests, {npts} points (generated in {time.time() - t0:.1f}s)",
          file=sys.stderr, flush=True)
    # warm the pool and the code objects on a few requests, outside the timed run
    S.Server(workers=args.serve_workers).run([S.Request(-1 - i, 0.0, r.pts, r.cons) for i, r in
                                              enumerate(reqs[:4])], realtime=False)
    srv = S.Server(workers=args.serve_workers, small_n=args.serve_small_n)
    st = srv.run(reqs)
    summ = st.summary(args.serve_qps)
    for r in reqs:
        r.fetch()
    summ.update({"points": int(npts), "workers": args.serve_workers, "small_n": srv.small_n,
                 "mpts_per_s": round(npts / max(st.makespan_s, 1e-9) / 1e6, 3),
                 "latency": "arrival to triangles resident on the device", "tunables": "defaults"})
    # saturation: every request waiting at t=0; continuous batching on the
    # worker streams against one plain solve at a time on one stream
    def fresh():
        return [S.Request(r.rid, 0.0, r.pts, r.cons) for r in reqs]

    sat_reqs = fresh()
    sat = S.Server(workers=args.serve_workers, small_n=args.serve_small_n).run(sat_reqs, realtime=False)
    for r in sat_reqs:
        r.fetch()
    one_reqs = fresh()
    one = S.Server(workers=1, small_n=0).run(one_reqs, realtime=False)
    for r in one_reqs:
        r.fetch()
    sat_rate = sat.completed / max(sat.makespan_s, 1e-9)
    one_rate = one.completed / max(one.makespan_s, 1e-9)
    summ["saturated"] = {"cdts_per_s": round(sat_rate, 2), "mean_batch": round(float(np.mean(sat.batch_sizes)), 2),
                         "batches": sat.batches, "p50_ms": sat.summary(0)["p50_ms"],
                         "p99_ms": sat.summary(0)["p99_ms"],
                         "one_at_a_time_cdts_per_s": round(one_rate, 2),
                         "speedup_cdts": round(sat_rate / max(one_rate, 1e-9), 2)}
    # latency against load: the same requests, their Poisson arrival times
    # rescaled to 10%, 40% and 80% of the saturated rate just measured
    levels = []
    for frac in (0.1, 0.4, 0.8):
        qps = frac * sat_rate
        lv = [S.Request(r.rid, r.arrival * args.serve_qps / qps, r.pts, r.cons) for r in reqs]
        st_l = S.Server(workers=args.serve_workers, small_n=args.serve_small_n).run(lv)
        for r in lv:
            r.fetch()
        sl = st_l.summary(round(qps, 1))
        sl["fraction_of_saturated"] = frac
        levels.append({k: sl[k] for k in ("fraction_of_saturated", "qps_offered", "cdts_per_s", "p50_ms", "p99_ms",
                                          "mean_batch", "completed", "failed")})
        print(f"# serving at {qps:.0f} QPS: {levels[-1]}", file=sys.stderr, flush=True)
    summ["load_levels"] = levels
    refs = None
    if not args.no_cpu:
        rate, thr, secs, refs = _cpu_pool_cdts([(r.pts, r.cons) for r in reqs])
        summ["cpu_baseline"] = {"value": round(rate, 2), "unit": "CDTs/s", "cores": thr, "kind": "port",
                                "host": _host_cpus(thr),
                                "sample": f"all {len(reqs)} requests, one per host thread, serial exact-predicate "
                                          f"CDT, {secs:.2f} s"}
    # every request of the timed run (merged or alone) against the oracle
    summ["exactness"] = _check_all(_checker(args), [(r.pts, r.cons) for r in reqs], [r.fetch() for r in reqs],
                                   refs)
    print(f"# serving: {summ}", file=sys.stderr, flush=True)
    return summ


if __name__ == "__main__":
    main()
