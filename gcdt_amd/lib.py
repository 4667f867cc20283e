"""ctypes binding of libgcdt.so (the C ABI declared in include/gcdt.h)."""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgcdt.so")

STATUS = {
    0: "ok", 1: "out of device memory", 2: "invalid argument", 3: "not converged", 4: "HIP device error",
    5: "crossing constraints", 6: "scratch too small", 7: "point location failed / duplicate point",
    8: "triangle capacity exceeded", 9: "internal consistency check failed",
}


class GcdtError(RuntimeError):
    def __init__(self, status: int, where: str, msg: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)} ({msg})")


class Int2(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32)]


class Int3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_sample", ctypes.c_double), ("ms_bitmap", ctypes.c_double),
        ("ms_insert", ctypes.c_double), ("ms_flip", ctypes.c_double), ("ms_constraints", ctypes.c_double),
        ("insert_rounds", ctypes.c_int), ("flip_rounds", ctypes.c_int), ("constraint_rounds", ctypes.c_int),
        ("num_triangles", ctypes.c_int), ("flips", ctypes.c_int64), ("segments", ctypes.c_int64),
        ("splits", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", ctypes.c_double), ("ms_insert", ctypes.c_double), ("ms_flip", ctypes.c_double),
        ("ms_constraints", ctypes.c_double), ("insert_rounds", ctypes.c_int), ("flip_rounds", ctypes.c_int),
        ("constraint_rounds", ctypes.c_int), ("num_instances", ctypes.c_int), ("num_points", ctypes.c_int64),
        ("num_triangles", ctypes.c_int64), ("flips", ctypes.c_int64), ("segments", ctypes.c_int64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class HPoint(ctypes.Structure):
    """gcdt_hpoint: a crossing point (X / w, Y / w), X and Y 128-bit."""
    _fields_ = [("x_hi", ctypes.c_int64), ("x_lo", ctypes.c_uint64), ("y_hi", ctypes.c_int64),
                ("y_lo", ctypes.c_uint64), ("w", ctypes.c_int64), ("seg", ctypes.c_int32 * 4)]


vp = ctypes.c_void_p
pvp = ctypes.POINTER(ctypes.c_void_p)
pint = ctypes.POINTER(ctypes.c_int)
pdbl = ctypes.POINTER(ctypes.c_double)
sz = ctypes.c_size_t
i32 = ctypes.c_int

# name -> (argtypes, restype); the full list declared by include/gcdt.h
SIGNATURES = {
    "gcdt_last_error": ([], ctypes.c_char_p),
    "gcdt_device_count": ([pint], i32),
    "gcdt_synchronize": ([], i32),
    "gcdt_device_name": ([ctypes.c_char_p, i32], i32),
    "gcdt_alloc": ([pvp, sz], i32),
    "gcdt_release": ([vp, sz], i32),
    "gcdt_pool_init": ([], i32),
    "gcdt_pool_stats": ([ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(sz)], i32),
    "gcdt_pool_set_budget": ([sz], i32),
    "gcdt_pool_reset_peak": ([], i32),
    "gcdt_pool_trim": ([], i32),
    "gcdt_memcpy_h2d": ([vp, vp, sz], i32),
    "gcdt_memcpy_d2h": ([vp, vp, sz], i32),
    "gcdt_memcpy_d2d": ([vp, vp, sz], i32),
    "gcdt_memset": ([vp, i32, sz], i32),
    "gcdt_scratch_bytes": ([ctypes.c_int64], sz),
    "gcdt_stream_create": ([pvp], i32),
    "gcdt_stream_destroy": ([vp], i32),
    "gcdt_stream_synchronize": ([vp], i32),
    "gcdt_memcpy_h2d_async": ([vp, vp, sz, vp], i32),
    "gcdt_memcpy_d2h_async": ([vp, vp, sz, vp], i32),
    "gcdt_read_file": ([ctypes.c_char_p, i32, ctypes.POINTER(ctypes.POINTER(ctypes.c_double)), pint, pvp, pint], i32),
    "gcdt_free_host": ([vp], None),
    "gcdt_trans_points": ([vp, i32, ctypes.c_longlong, vp, vp], i32),
    "gcdt_canonicalize_host": ([vp, ctypes.c_int64], i32),
    "gcdt_write_mesh": ([ctypes.c_char_p, vp, i32, vp, ctypes.c_int64], i32),
    "gcdt_prof_enable": ([i32], i32),
    "gcdt_prof_reset": ([], i32),
    "gcdt_prof_only": ([ctypes.c_char_p], i32),
    "gcdt_prof_read": ([ctypes.c_char_p, pdbl, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)], i32),
    "gcdt_prof_names": ([ctypes.c_char_p, i32], i32),
    "gcdt_get_sample_points": ([vp, sz, vp, i32, pvp, pvp, pint, pdbl], i32),
    "gcdt_raw_mesh_before": ([vp, sz, vp, vp, vp, pint, i32, pdbl], i32),
    "gcdt_raw_mesh_after": ([vp, sz, vp, vp, vp, vp, vp, vp, pint, i32, pdbl], i32),
    "gcdt_raw_mesh": ([vp, sz, vp, vp, vp, vp, pint, i32, pdbl], i32),
    "gcdt_get_bitmap": ([vp, sz, vp, vp, pvp, pvp, i32, pdbl], i32),
    "gcdt_filter_triangles": ([vp, sz, vp, vp, vp, vp, vp, pint, pdbl], i32),
    "gcdt_convert_sample": ([vp, vp, i32, i32, i32], i32),
    "gcdt_delaunay_sample": ([vp, sz, vp, vp, vp, i32, i32, pdbl], i32),
    "gcdt_try_fast_delaunay": ([vp, sz, vp, vp, vp, i32, i32, pdbl], i32),
    "gcdt_delaunay_new": ([vp, sz, pvp, pvp, vp, i32, i32, i32, pdbl], i32),
    "gcdt_add_constraints": ([vp, sz, vp, vp, vp, vp, vp, vp, vp, pvp, i32, i32, i32, pint, pdbl], i32),
    "gcdt_solve": ([vp, i32, vp, i32, vp, vp, pint, ctypes.POINTER(Stats), vp], i32),
    "gcdt_solve_flags": ([vp, i32, vp, i32, vp, vp, vp, pint, ctypes.POINTER(Stats), vp], i32),
    "gcdt_solve_noalloc": ([vp, i32, vp, i32, vp, vp, pint, ctypes.POINTER(Stats), vp], i32),
    "gcdt_carve": ([vp, vp, vp, i32, i32, vp, pint, vp], i32),
    "gcdt_check_mesh": ([vp, ctypes.c_int64, vp, vp, vp, i32, i32, pint, pint, vp], i32),
    "gcdt_solve_steiner": ([vp, i32, vp, i32, i32, vp, vp, vp, vp, pint, pint, ctypes.POINTER(Stats), vp], i32),
    "gcdt_solve_batch": ([vp, ctypes.POINTER(ctypes.c_int64), i32, vp, ctypes.POINTER(ctypes.c_int64), vp,
                          ctypes.POINTER(BatchStats), vp, vp], i32),
    "gcdt_set_tunable": ([ctypes.c_char_p, i32], i32),
    "gcdt_get_tunable": ([ctypes.c_char_p, pint], i32),
}

_lib = None
_err = None


def load_error():
    return _err


def lib():
    """The loaded libgcdt.so; raises if it is missing (no fallback path)."""
    global _lib, _err
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        _err = f"{LIB_PATH} not built (run python -m gcdt_amd.build)"
        raise RuntimeError(_err)
    L = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(L, name, None)
        if f is None:
            continue
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(status: int, where: str):
    if status != 0:
        msg = lib().gcdt_last_error()
        raise GcdtError(status, where, msg.decode() if msg else "")
