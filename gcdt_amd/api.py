"""Host-side convenience layer over the C ABI: device buffers and solve().

Inputs are integer points already on the 2^30 grid (gen.to_grid maps
[0,2^20) coordinates exactly) and constraint vertex-id pairs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from .lib import Stats, check, lib


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
