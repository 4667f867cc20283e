"""Input/output formats over the C ABI.

read_file    -- the reference's text point/constraint file (kernel.cu:5159-5291)
trans_points -- its double -> 2^30-grid map, on the GPU (kernel.cu:171-188)
canonical    -- canonical triangle list (smallest id first, sorted)
write_mesh   -- points and triangles as text
"""
from __future__ import annotations

import ctypes

import numpy as np

from .api import DeviceBuffer
from .lib import check, lib

GRID_RES = 1 << 30  # include/math_base.cuh:19 logPointRes = 30


def read_file(path: str, shuffle: bool = False):
    """Returns (points (n,2) float64 in [0.005,0.995]^2, constraints (c,2) int32)."""
    L = lib()
    P = ctypes.POINTER(ctypes.c_double)()
    C = ctypes.c_void_p()
    n = ctypes.c_int(0)
    c = ctypes.c_int(0)
    check(L.gcdt_read_file(str(path).encode(), int(bool(shuffle)), ctypes.byref(P), ctypes.byref(n),
                           ctypes.byref(C), ctypes.byref(c)), "gcdt_read_file")
    try:
        pts = np.ctypeslib.as_array(P, shape=(2 * n.value,)).reshape(-1, 2).copy() if n.value else \
            np.zeros((0, 2), np.float64)
        cons = np.frombuffer((ctypes.c_int32 * (2 * c.value)).from_address(C.value), np.int32).reshape(-1, 2).copy() \
            if c.value else np.zeros((0, 2), np.int32)
    finally:
        L.gcdt_free_host(ctypes.cast(P, ctypes.c_void_p))
        L.gcdt_free_host(C)
    return pts, cons


def trans_points(fpoints: np.ndarray, res: int = GRID_RES) -> np.ndarray:
    """Map [0,1) doubles onto the integer grid on the device; returns (n,2) int32."""
    f = np.ascontiguousarray(fpoints, dtype=np.float64).reshape(-1, 2)
    n = len(f)
    if n == 0:
        return np.zeros((0, 2), np.int32)
    with DeviceBuffer.from_array(f) as df, DeviceBuffer(8 * n) as dp:
        check(lib().gcdt_trans_points(df.ptr, n, int(res), dp.ptr, None), "gcdt_trans_points")
        return dp.to_array(np.int32, (n, 2))


def canonical(tris: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3).copy()
    check(lib().gcdt_canonicalize_host(t.ctypes.data, len(t)), "gcdt_canonicalize_host")
    return t


def write_mesh(path: str, points: np.ndarray, tris: np.ndarray):
    p = np.ascontiguousarray(points, dtype=np.int32).reshape(-1, 2)
    t = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    check(lib().gcdt_write_mesh(str(path).encode(), p.ctypes.data, len(p), t.ctypes.data, len(t)), "gcdt_write_mesh")


def write_input(path: str, points: np.ndarray, cons: np.ndarray | None = None):
    """Writes points (any numeric type) and constraints in the readFile layout."""
    with open(path, "w") as f:
        f.write(f"{len(points)}\n")
        pts = np.asarray(points).reshape(-1, 2)
        as_float = pts.dtype.kind == "f"
        for i, (x, y) in enumerate(pts):
            # repr of a Python float round-trips exactly through strtod
            f.write(f"{i} {float(x)!r} {float(y)!r}\n" if as_float else f"{i} {int(x)} {int(y)}\n")
        cons = np.zeros((0, 2), np.int64) if cons is None else np.asarray(cons).reshape(-1, 2)
        f.write(f"{len(cons)}\n")
        for i, (a, b) in enumerate(cons):
            f.write(f"{i} {int(a)} {int(b)}\n")
