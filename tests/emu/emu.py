"""ctypes binding of the CPU SIMT emulator build (tests/emu/_build/libgcdt_emu.so)
and a case runner for it (test infrastructure).

The emulator build exports the library's C ABI; these helpers mirror
gcdt_amd.api's solve / solve_batch over it.  Cases run in a child process
(``python -m tests.emu.emu <case json>``), because a fault inside the
emulator ends the process with status 86 after printing the kernel, the
work-item and the allocation involved.
"""
from __future__ import annotations

import ctypes
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import importlib  # noqa: E402

L = importlib.import_module("gcdt_amd.lib")  # signature table only; libgcdt.so is not loaded

LIB = os.path.join(HERE, "_build", "libgcdt_emu.so")
# variant with small caps, so the paths for large cases run on small meshes:
# the per-cell reorder ranks buckets of up to 3 elements lane by lane, sorts
# 4-8 in a workgroup and merge-sorts larger ones (the defaults are 256 and
# 2048), and cavities of more than 4 crossed triangles are retriangulated by a
# whole wave (the default is 64)
CELLCAP = (("GCDT_CELL_RANK_MAX=3", "GCDT_CELL_BLOCK_MAX=8", "GCDT_CAV_WAVE_K=4"),
          os.path.join(HERE, "_build", "libgcdt_emu_cellcap.so"))
_lib = None


def build(variant=None):
    """Builds the emulator library (or a variant: (defines, path)) unless it
    is newer than every source (serialised by a lock file, so parallel test
    workers build it once)."""
    import fcntl
    from tests.emu import build_emu
    defines, lib = variant if variant else ((), LIB)
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    with open(os.path.join(HERE, "_build", ".lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        srcs = [os.path.join(build_emu.CSRC, f) for f in build_emu.SOURCES + build_emu.HEADERS]
        srcs += [os.path.join(HERE, f) for f in ("emu_rt.cpp", "build_emu.py", os.path.join("hip", "hip_runtime.h"))]
        if os.path.exists(lib) and os.path.getmtime(lib) >= max(os.path.getmtime(f) for f in srcs):
            return lib
        return build_emu.build(defines=defines, lib=lib)


def load():
    global _lib
    if _lib is None:
        path = os.environ.get("EMU_LIB") or LIB
        if not os.path.exists(path):
            build()
        d = ctypes.CDLL(path)
        for name, (args, res) in L.SIGNATURES.items():
            f = getattr(d, name, None)
            if f is None:
                continue
            f.argtypes = args
            f.restype = res
        d.emu_set_tunable.argtypes = [ctypes.c_char_p, ctypes.c_int]
        d.emu_set_tunable.restype = ctypes.c_int
        d.emu_launches.restype = ctypes.c_ulonglong
        _lib = d
    return _lib


def check(rc, where):
    if rc != 0:
        msg = load().gcdt_last_error()
        raise L.GcdtError(rc, where, msg.decode() if msg else "")


def set_knob(name: str, value: int):
    if load().emu_set_tunable(name.encode(), int(value)) != 0:
        raise ValueError(name)


class Buf:
    def __init__(self, nbytes):
        self.nbytes = max(int(nbytes), 1)
        p = ctypes.c_void_p()
        check(load().gcdt_alloc(ctypes.byref(p), self.nbytes), "gcdt_alloc")
        self.ptr = p.value

    @classmethod
    def of(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(load().gcdt_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes), "h2d")
        return b

    def get(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            check(load().gcdt_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            load().gcdt_release(self.ptr, self.nbytes)
            self.ptr = None


def solve(g, cons, entry="gcdt_solve"):
    """entry: gcdt_solve or gcdt_solve_noalloc (same arguments)."""
    g = np.ascontiguousarray(g, np.int32).reshape(-1, 2)
    c = np.ascontiguousarray(cons, np.int32).reshape(-1, 2)
    n = len(g)
    dp, dt, da = Buf.of(g), Buf(12 * (2 * n + 2)), Buf(12 * (2 * n + 2))
    dc = Buf.of(c) if len(c) else None
    nt = ctypes.c_int(0)
    st = L.Stats()
    check(getattr(load(), entry)(dp.ptr, n, dc.ptr if dc else None, len(c), dt.ptr, da.ptr, ctypes.byref(nt),
                                 ctypes.byref(st), None), entry)
    tris = dt.get(np.int32, (nt.value, 3))
    adj = da.get(np.int32, (nt.value, 3))
    for b in (dp, dt, da, dc):
        if b:
            b.free()
    return tris, adj, st.as_dict()


def solve_steiner(g, cons, max_steiner=None):
    """gcdt_solve_steiner: (tris, [(X, Y, W)], stats)."""
    from gcdt_amd.api import HPOINT_DTYPE, hpoint_coords
    g = np.ascontiguousarray(g, np.int32).reshape(-1, 2)
    c = np.ascontiguousarray(cons, np.int32).reshape(-1, 2)
    n = len(g)
    ms = 4 * len(c) + 64 if max_steiner is None else max_steiner
    cap = 2 * (n + ms) + 2
    dp, dt, da, df = Buf.of(g), Buf(12 * cap), Buf(12 * cap), Buf(4 * cap)
    dh = Buf(HPOINT_DTYPE.itemsize * max(ms, 1))
    dc = Buf.of(c) if len(c) else None
    nt, ns = ctypes.c_int(0), ctypes.c_int(0)
    st = L.Stats()
    try:
        check(load().gcdt_solve_steiner(dp.ptr, n, dc.ptr if dc else None, len(c), ms, dt.ptr, da.ptr, df.ptr,
                                        dh.ptr, ctypes.byref(nt), ctypes.byref(ns), ctypes.byref(st), None),
              "gcdt_solve_steiner")
        tris = dt.get(np.int32, (nt.value, 3))
        raw = dh.get(HPOINT_DTYPE, (ns.value,))
    finally:
        for b in (dp, dt, da, df, dh, dc):
            if b:
                b.free()
    return tris, [hpoint_coords(h) for h in raw], st.as_dict()


def solve_batch(instances):
    from gcdt_amd.api import pack_instances
    flat_p, poff, flat_c, coff, toff = pack_instances(instances)
    dp = Buf.of(flat_p)
    dc = Buf.of(flat_c) if len(flat_c) else None
    dt = Buf(12 * int(toff[-1]))
    st = L.BatchStats()
    P64 = ctypes.POINTER(ctypes.c_int64)
    check(load().gcdt_solve_batch(dp.ptr, poff.ctypes.data_as(P64), len(instances), dc.ptr if dc else None,
                                  coff.ctypes.data_as(P64) if dc else None, dt.ptr, ctypes.byref(st), None, None),
          "gcdt_solve_batch")
    allt = dt.get(np.int32, (int(toff[-1]), 3))
    for b in (dp, dt, dc):
        if b:
            b.free()
    return [allt[toff[i]:toff[i + 1]] for i in range(len(instances))], st.as_dict()


# ------------------------------------------------------------------ cases

def _instances(spec):
    from gcdt_amd import gen
    if "count" in spec:  # the batched workload's star instances
        inst = gen.batched_instances(count=spec["count"], nmin=spec.get("nmin", 64), nmax=spec.get("nmax", 65536),
                                     seed=spec.get("seed", 6))
        return [(gen.to_grid(p), c) for p, c in inst]
    s = gen.Stream(spec.get("seed", 7))
    out = []
    for n in spec["sizes"]:
        p, c = gen.polygon_instance(s, int(n))
        out.append((gen.to_grid(p), c))
    return out


def run_case(case: dict) -> dict:
    """One emulated solve with the given knobs, validated by the oracle."""
    from gcdt_amd import gen
    from oracle import oracle as O
    for k, v in case.get("knobs", {}).items():
        set_knob(k, v)
    res = {"valid": True, "unique_mismatch": 0}
    kind = case["kind"]
    if kind == "noalloc":
        # the allocation-free solve against gcdt_solve on a sequence of inputs
        # (each call runs over the previous call's leftovers in the array)
        if case.get("init", True):
            check(load().gcdt_pool_init(), "gcdt_pool_init")
        res["status"] = []
        for spec in case["inputs"]:
            if spec[0] == "uniform":
                p, c = gen.uniform_cdt(n=spec[1], c=spec[2], seed_pts=spec[3], seed_cons=spec[3] + 1)
            elif spec[0] == "star":
                p, c = gen.star_instance(gen.Stream(spec[2]), spec[1])
            else:
                p, c = gen.lattice_cdt(k=spec[1], c=spec[2], max_len=spec[3])
            g = gen.to_grid(p)
            try:
                t1, a1, _ = solve(g, c, entry="gcdt_solve_noalloc")
            except L.GcdtError as e:
                res["status"].append(e.status)
                continue
            res["status"].append(0)
            t0, _, _ = solve(g, c)
            v = O.validate(g, t1, c)
            res["valid"] &= bool(v["valid"]) and bool(O.validate(g, t0, c)["valid"])
            if v["valid"] and v["unique"]:
                res["unique_mismatch"] += int(not np.array_equal(O.canonical(t0), O.canonical(t1)))
        return res
    if kind == "check":
        # gcdt_check_mesh on a solved mesh and on copies with one planted
        # mistake each; returns the failed-check bits per copy
        p, c = gen.uniform_cdt(n=case["n"], c=case.get("c", 0), seed_pts=3, seed_cons=4)
        g = gen.to_grid(p)
        tris, adj, _ = solve(g, c)
        R = (1 << 30) - 1
        pc = np.concatenate([g, np.array([[0, 0], [R, 0], [R, R], [0, R]], np.int32)])

        def bits(t, a, delaunay=1):
            bufs = [Buf.of(pc), Buf.of(np.ascontiguousarray(t)), Buf.of(np.ascontiguousarray(a))]
            f, first = ctypes.c_int(0), ctypes.c_int(0)
            check(load().gcdt_check_mesh(bufs[0].ptr, len(pc), bufs[1].ptr, bufs[2].ptr, None, len(t), delaunay,
                                         ctypes.byref(f), ctypes.byref(first), None), "gcdt_check_mesh")
            for b in bufs:
                b.free()
            return f.value, first.value

        res["clean"] = bits(tris, adj)
        t2 = tris.copy()
        t2[5, [1, 2]] = t2[5, [2, 1]]  # clockwise, and its neighbours no longer match
        res["swapped"] = bits(t2, adj)
        a2 = adj.copy()
        k = int(np.argmax(a2[7] >= 0))
        a2[7, k] = (a2[7, k] + 1) % len(tris)  # a neighbour that does not name 7 back
        res["adjacency"] = bits(tris, a2)
        a3 = adj.copy()
        a3[9, 0] = len(tris) + 5  # out of range
        res["range"] = bits(tris, a3)
        # flip an interior Delaunay edge that is not cocircular: both new
        # triangles stay valid and counter-clockwise but the edge is not
        # locally Delaunay
        from oracle.oracle import with_corners
        P = with_corners(g)
        done = False
        for t in range(len(tris)):
            for k in range(3):
                n = int(adj[t, k])
                if n <= t:
                    continue
                u, w = int(tris[t, (k + 1) % 3]), int(tris[t, (k + 2) % 3])
                pv = int(tris[t, k])
                o = [int(x) for x in tris[n] if x != u and x != w][0]
                def orient(a, b, cc):
                    return int((P[b][0] - P[a][0]) * (P[cc][1] - P[a][1]) - (P[b][1] - P[a][1]) * (P[cc][0] - P[a][0]))
                if orient(pv, u, o) <= 0 or orient(pv, o, w) <= 0:
                    continue  # not a convex quad
                t3 = tris.copy()
                t3[t] = (pv, u, o)
                t3[n] = (pv, o, w)
                # the flipped pair's own adjacency, rebuilt by brute force
                a4 = adj.copy()
                for x in (t, n):
                    for kk in range(3):
                        e = {int(t3[x, (kk + 1) % 3]), int(t3[x, (kk + 2) % 3])}
                        m = [y for y in range(len(t3)) if y != x and e <= set(int(z) for z in t3[y])]
                        a4[x, kk] = m[0] if m else -1
                        for y in m:
                            for jj in range(3):
                                if {int(t3[y, (jj + 1) % 3]), int(t3[y, (jj + 2) % 3])} == e:
                                    a4[y, jj] = x
                f, first = bits(t3, a4)
                if f == 16:
                    res["nondelaunay"] = (f, first)
                    res["nondelaunay_without_check"] = bits(t3, a4, 0)
                    done = True
                    break
            if done:
                break
        return res
    if kind == "steiner":
        from tests import steiner_cases as SC
        spec = case["gen"]
        if spec[0] == "random":
            g, c = SC.random_crossing(*spec[1:])
        elif spec[0] == "star":
            g, c = SC.star(*spec[1:])
        else:
            g, c = SC.hd_family(*spec[1:])
        try:
            tris, hp, st = solve_steiner(g, c, case.get("max_steiner"))
        except L.GcdtError as e:
            res["status"] = e.status
            return res
        res["steiner"] = len(hp)
        if case.get("digest"):
            import hashlib
            res["digest"] = hashlib.sha256(np.ascontiguousarray(tris).tobytes() + repr(hp).encode()).hexdigest()
        try:
            v = SC.check(g, c, tris, hp, brute=case.get("brute", False), augmented=case.get("augmented", False))
            res["unique"] = bool(v["unique"])
        except AssertionError as e:
            res["valid"] = False
            res["error"] = str(e)[:2000]
        return res
    if kind == "batch":
        inst = _instances(case)
        tris, st = solve_batch(inst)
        if case.get("digest"):
            import hashlib
            res["digest"] = hashlib.sha256(b"".join(t.tobytes() for t in tris)).hexdigest()
        bad = []
        for i, ((g, c), t) in enumerate(zip(inst, tris)):
            v = O.validate(g, t, c)
            if not v["valid"]:
                bad.append(i)
        res["valid"] = not bad
        res["bad"] = bad
        res["stats"] = {k: st[k] for k in ("insert_rounds", "flip_rounds", "constraint_rounds", "num_triangles")}
    else:
        if kind == "uniform":
            p, c = gen.uniform_cdt(n=case["n"], c=case.get("c", 0), seed_pts=case.get("seed", 1),
                                   seed_cons=case.get("seed", 1) + 1)
        elif kind == "lattice":
            p, c = gen.lattice_cdt(k=case["k"], c=case.get("c", 0), max_len=case.get("max_len", 8))
        elif kind == "clustered":
            p, c = gen.clustered_cdt(n=case["n"], circles=case.get("circles", 0))
        else:
            raise ValueError(kind)
        g = gen.to_grid(p)
        tris, adj, st = solve(g, c)
        if case.get("digest"):
            import hashlib
            res["digest"] = hashlib.sha256(tris.tobytes() + adj.tobytes()).hexdigest()
        v = O.validate(g, tris, c)
        res["valid"] = bool(v["valid"])
        if v["valid"] and v.get("unique"):
            ref, _ = O.cdt(g, c)
            res["unique_mismatch"] = int(not np.array_equal(O.canonical(tris), O.canonical(ref)))
        res["stats"] = {k: st[k] for k in ("insert_rounds", "flip_rounds", "constraint_rounds", "num_triangles")}
    res["launches"] = int(load().emu_launches())
    return res


def run_in_child(case: dict, seed: int = 1, align_low: int = 0, conc: int = 1, yield_atomics: int = 0,
                 timeout: float = 600.0, yield_mem: int = 0, lib: str | None = None):
    """Runs `case` in a child process; returns (exit status, parsed result or
    None, stderr tail).  seed: order of workgroups and work-items and the
    fill of new allocations; align_low: allocations against their lower guard
    page; conc: workgroups interleaved at once; yield_atomics: switch
    work-items at every atomic; yield_mem: also between the 16-byte parts
    of a record load or store."""
    env = dict(os.environ, EMU_SEED=str(seed), EMU_ALIGN_LOW=str(align_low), EMU_CONC=str(conc),
               EMU_YIELD_ATOMICS=str(yield_atomics), EMU_YIELD_MEM=str(yield_mem))
    if lib:
        env["EMU_LIB"] = lib
    r = subprocess.run([sys.executable, "-m", "tests.emu.emu", json.dumps(case)], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=timeout)
    out = None
    for line in r.stdout.splitlines():
        if line.startswith("EMU_RESULT "):
            out = json.loads(line[len("EMU_RESULT "):])
    return r.returncode, out, r.stderr[-3000:]


if __name__ == "__main__":
    case = json.loads(sys.argv[1])
    print("EMU_RESULT " + json.dumps(run_case(case)), flush=True)
