"""Host-side convenience layer over the C ABI: device buffers and solve().

Inputs are integer points already on the 2^30 grid (gen.to_grid maps
[0,2^20) coordinates exactly) and constraint vertex-id pairs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import BatchStats, Stats, check, lib


class DeviceBuffer:
    """Device memory from the library's MemCache pool."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 1)
        p = ctypes.c_void_p()
        check(lib().gcdt_alloc(ctypes.byref(p), self.nbytes), "gcdt_alloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().gcdt_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes), "gcdt_memcpy_h2d")
        return b

    def to_array(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes:
            check(lib().gcdt_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "gcdt_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            lib().gcdt_release(self.ptr, self.nbytes)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def solve_device(d_pts: DeviceBuffer, n: int, d_cons: DeviceBuffer | None, c: int, d_tris: DeviceBuffer,
                 d_adj: DeviceBuffer | None = None, stream=None):
    """gcdt_solve on device buffers; returns (num_tris, stats dict)."""
    nt = ctypes.c_int(0)
    st = Stats()
    rc = lib().gcdt_solve(d_pts.ptr, n, d_cons.ptr if d_cons is not None and c > 0 else None, c, d_tris.ptr,
                          d_adj.ptr if d_adj is not None else None, ctypes.byref(nt), ctypes.byref(st), stream)
    check(rc, "gcdt_solve")
    return nt.value, st.as_dict()


def solve(pts_grid: np.ndarray, cons: np.ndarray | None = None, with_adj: bool = False):
    """Full CDT on the GPU.  Returns (tris (2n+2,3) int32, stats[, adj])."""
    p = np.ascontiguousarray(pts_grid, dtype=np.int32).reshape(-1, 2)
    n = len(p)
    c = np.zeros((0, 2), np.int32) if cons is None else np.ascontiguousarray(cons, dtype=np.int32).reshape(-1, 2)
    dp = DeviceBuffer.from_array(p)
    dc = DeviceBuffer.from_array(c) if len(c) else None
    dt = DeviceBuffer(12 * (2 * n + 2))
    da = DeviceBuffer(12 * (2 * n + 2))
    try:
        T, st = solve_device(dp, n, dc, len(c), dt, da)
        tris = dt.to_array(np.int32, (T, 3))
        adj = da.to_array(np.int32, (T, 3)) if with_adj else None
    finally:
        for b in (dp, dc, dt, da):
            if b is not None:
                b.free()
    return (tris, st, adj) if with_adj else (tris, st)


def hpoint_coords(h) -> tuple[int, int, int]:
    """(X, Y, W) Python integers of a gcdt_hpoint: the point (X / W, Y / W)."""
    x = (int(h["x_hi"]) << 64) | int(h["x_lo"])
    y = (int(h["y_hi"]) << 64) | int(h["y_lo"])
    return x, y, int(h["w"])


HPOINT_DTYPE = np.dtype([("x_hi", "<i8"), ("x_lo", "<u8"), ("y_hi", "<i8"), ("y_lo", "<u8"), ("w", "<i8"),
                         ("seg", "<i4", (4,))])


def solve_steiner(pts_grid: np.ndarray, cons: np.ndarray, max_steiner: int | None = None):
    """CDT with crossing constraints resolved by exact crossing vertices.
    Returns (tris (T,3) int32, crossings [(X, Y, W)] for vertex ids n+4, n+5,
    ..., their pieces (S,4) int32, cflag (T,) int32, stats).  max_steiner
    defaults to 4 * len(cons) + 64."""
    p = np.ascontiguousarray(pts_grid, dtype=np.int32).reshape(-1, 2)
    n = len(p)
    c = np.ascontiguousarray(cons, dtype=np.int32).reshape(-1, 2)
    ms = 4 * len(c) + 64 if max_steiner is None else int(max_steiner)
    cap = 2 * (n + ms) + 2
    dp = DeviceBuffer.from_array(p)
    dc = DeviceBuffer.from_array(c) if len(c) else None
    dt, da, df = DeviceBuffer(12 * cap), DeviceBuffer(12 * cap), DeviceBuffer(4 * cap)
    dh = DeviceBuffer(HPOINT_DTYPE.itemsize * max(ms, 1))
    try:
        nt, ns = ctypes.c_int(0), ctypes.c_int(0)
        st = Stats()
        check(lib().gcdt_solve_steiner(dp.ptr, n, dc.ptr if dc is not None else None, len(c), ms, dt.ptr, da.ptr,
                                       df.ptr, dh.ptr, ctypes.byref(nt), ctypes.byref(ns), ctypes.byref(st), None),
              "gcdt_solve_steiner")
        tris = dt.to_array(np.int32, (nt.value, 3))
        cflag = df.to_array(np.int32, (nt.value,))
        raw = np.zeros(ns.value, HPOINT_DTYPE)
        if ns.value:
            lib().gcdt_memcpy_d2h(raw.ctypes.data, dh.ptr, raw.nbytes)
        pts = [hpoint_coords(h) for h in raw]
        segs = raw["seg"].astype(np.int32).reshape(-1, 4)
    finally:
        for b in (dp, dc, dt, da, df, dh):
            if b is not None:
                b.free()
    d = st.as_dict()
    d["steiner"] = ns.value
    return tris, pts, segs, cflag, d


def solve_carved(pts_grid: np.ndarray, cons: np.ndarray):
    """CDT plus even-odd carving.  Returns (all triangles (2n+2,3), constraint
    bits (2n+2,), inside mask (2n+2,) bool, stats)."""
    p = np.ascontiguousarray(pts_grid, dtype=np.int32).reshape(-1, 2)
    n = len(p)
    c = np.ascontiguousarray(cons, dtype=np.int32).reshape(-1, 2)
    dp = DeviceBuffer.from_array(p)
    dc = DeviceBuffer.from_array(c) if len(c) else None
    T = 2 * n + 2
    dt, da, df, di = DeviceBuffer(12 * T), DeviceBuffer(12 * T), DeviceBuffer(4 * T), DeviceBuffer(4 * T)
    try:
        nt = ctypes.c_int(0)
        st = Stats()
        check(lib().gcdt_solve_flags(dp.ptr, n, dc.ptr if dc is not None else None, len(c), dt.ptr, da.ptr, df.ptr,
                                     ctypes.byref(nt), ctypes.byref(st), None), "gcdt_solve_flags")
        ni = ctypes.c_int(0)
        check(lib().gcdt_carve(dt.ptr, da.ptr, df.ptr, nt.value, n, di.ptr, ctypes.byref(ni), None), "gcdt_carve")
        tris = dt.to_array(np.int32, (nt.value, 3))
        cflag = df.to_array(np.int32, (nt.value,))
        inside = di.to_array(np.int32, (nt.value,)).astype(bool)
    finally:
        for b in (dp, dc, dt, da, df, di):
            if b is not None:
                b.free()
    return tris, cflag, inside, st.as_dict()


CHECK_BITS = {1: "vertex ids", 2: "orientation", 4: "adjacency", 8: "constraint bits", 16: "local Delaunay"}


def check_mesh(pts_grid: np.ndarray, tris: np.ndarray, adj: np.ndarray, cflag: np.ndarray | None = None,
               delaunay: bool = True):
    """gcdt_check_mesh on host arrays in the caller's ids (pts_grid without the
    four corners, which are appended as vertices n..n+3).  Returns (failed
    check bits, first failing triangle or -1)."""
    p = np.ascontiguousarray(pts_grid, dtype=np.int32).reshape(-1, 2)
    R = (1 << 30) - 1
    pc = np.concatenate([p, np.array([[0, 0], [R, 0], [R, R], [0, R]], np.int32)])
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    a = np.ascontiguousarray(adj, dtype=np.int32).reshape(-1, 3)
    bufs = [DeviceBuffer.from_array(pc), DeviceBuffer.from_array(t), DeviceBuffer.from_array(a)]
    if cflag is not None:
        bufs.append(DeviceBuffer.from_array(np.ascontiguousarray(cflag, dtype=np.int32)))
    try:
        f, first = ctypes.c_int(0), ctypes.c_int(0)
        check(lib().gcdt_check_mesh(bufs[0].ptr, len(pc), bufs[1].ptr, bufs[2].ptr,
                                    bufs[3].ptr if cflag is not None else None, len(t), int(delaunay),
                                    ctypes.byref(f), ctypes.byref(first), None), "gcdt_check_mesh")
    finally:
        for b in bufs:
            b.free()
    return f.value, first.value


def pack_instances(instances):
    """[(pts_grid (n_i,2), cons (c_i,2) local ids)] -> flat arrays + offsets."""
    pts = [np.ascontiguousarray(p, dtype=np.int32).reshape(-1, 2) for p, _ in instances]
    cons = [np.ascontiguousarray(c if c is not None else np.zeros((0, 2)), dtype=np.int32).reshape(-1, 2)
            for _, c in instances]
    poff = np.zeros(len(instances) + 1, np.int64)
    coff = np.zeros(len(instances) + 1, np.int64)
    poff[1:] = np.cumsum([len(p) for p in pts])
    coff[1:] = np.cumsum([len(c) for c in cons])
    flat_p = np.concatenate(pts) if pts else np.zeros((0, 2), np.int32)
    flat_c = np.concatenate(cons) if cons else np.zeros((0, 2), np.int32)
    toff = np.zeros(len(instances) + 1, np.int64)
    toff[1:] = np.cumsum([2 * len(p) + 2 for p in pts])
    return flat_p, poff, flat_c, coff, toff


def solve_batch_device(d_pts: DeviceBuffer, poff: np.ndarray, d_cons, coff: np.ndarray, d_tris: DeviceBuffer,
                       stream=None, inst_status: np.ndarray | None = None):
    """gcdt_solve_batch on device buffers.  inst_status (optional int32 array
    of K): per-instance status, so that an instance with a bad input fails
    alone instead of failing the call."""
    K = len(poff) - 1
    st = BatchStats()
    poff = np.ascontiguousarray(poff, dtype=np.int64)
    coff = np.ascontiguousarray(coff, dtype=np.int64)
    P64 = ctypes.POINTER(ctypes.c_int64)
    if inst_status is not None:
        assert inst_status.dtype == np.int32 and inst_status.flags.c_contiguous and len(inst_status) == K
    rc = lib().gcdt_solve_batch(d_pts.ptr, poff.ctypes.data_as(P64), K,
                                d_cons.ptr if d_cons is not None else None,
                                coff.ctypes.data_as(P64) if d_cons is not None else None,
                                d_tris.ptr, ctypes.byref(st),
                                inst_status.ctypes.data if inst_status is not None else None, stream)
    check(rc, "gcdt_solve_batch")
    return st.as_dict()


def solve_batch(instances, isolate: bool = False):
    """K independent CDTs in one batched solve.  Returns (list of per-instance
    triangle arrays with local vertex ids, stats); with isolate, stats also
    holds "status", the per-instance GCDT_* status (an instance with a bad
    input fails alone and its triangles are not meaningful)."""
    flat_p, poff, flat_c, coff, toff = pack_instances(instances)
    dp = DeviceBuffer.from_array(flat_p)
    dc = DeviceBuffer.from_array(flat_c) if len(flat_c) else None
    dt = DeviceBuffer(12 * int(toff[-1]))
    status = np.zeros(len(instances), np.int32) if isolate else None
    try:
        st = solve_batch_device(dp, poff, dc, coff, dt, inst_status=status)
        allt = dt.to_array(np.int32, (int(toff[-1]), 3))
    finally:
        for b in (dp, dc, dt):
            if b is not None:
                b.free()
    if isolate:
        st["status"] = status.tolist()
    return [allt[toff[i]:toff[i + 1]] for i in range(len(instances))], st


def set_tunable(name: str, value: int):
    check(lib().gcdt_set_tunable(name.encode(), int(value)), "gcdt_set_tunable")


def get_tunable(name: str) -> int:
    v = ctypes.c_int(0)
    check(lib().gcdt_get_tunable(name.encode(), ctypes.byref(v)), "gcdt_get_tunable")
    return v.value
