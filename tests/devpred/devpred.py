"""Builds and binds tests/devpred/devpred.hip (test infrastructure): the
library's exact predicates run on the GPU over given point tuples."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "devpred.hip")
LIB = os.path.join(HERE, "libdevpred.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# the device branch of the same predicates compiled for the host (CPU tests)
HOST_SRC = os.path.join(HERE, "host_devbranch.cpp")
HOST_LIB = os.path.join(HERE, "libhost_devbranch.so")
EMU_INC = os.path.join(os.path.dirname(HERE), "emu")


def build() -> str:
    cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-shared", "--offload-arch=gfx950", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-o", LIB, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {SRC}:\n{r.stdout}\n{r.stderr}")
    return LIB


def build_host() -> str:
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", EMU_INC, "-o", HOST_LIB, HOST_SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed for {HOST_SRC}:\n{r.stdout}\n{r.stderr}")
    return HOST_LIB


_lib = None
_host = None


def host():
    """The device branch run on the host: same functions as the GPU ones."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB) or os.path.getmtime(HOST_LIB) < os.path.getmtime(HOST_SRC):
            build_host()
        L = ctypes.CDLL(HOST_LIB)
        vp = ctypes.c_void_p
        for f in (L.hostdev_orient2d, L.hostdev_incircle):
            f.argtypes = [vp, ctypes.c_int, vp]
            f.restype = ctypes.c_int
        L.hostdev_mul_s64.argtypes = [vp, ctypes.c_int, vp, vp]
        L.hostdev_mul_s64.restype = ctypes.c_int
        for f in (L.hostdev_orient_h, L.hostdev_incircle_h, L.hostdev_cross):
            f.argtypes = [vp, ctypes.c_int, vp]
            f.restype = ctypes.c_int
        _host = L
    return _host


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        for f in (L.devpred_orient2d, L.devpred_incircle):
            f.argtypes = [vp, ctypes.c_int, vp]
            f.restype = ctypes.c_int
        L.devpred_mul_s64.argtypes = [vp, ctypes.c_int, vp, vp]
        L.devpred_mul_s64.restype = ctypes.c_int
        for f in (L.devpred_orient_h, L.devpred_incircle_h, L.devpred_cross):
            f.argtypes = [vp, ctypes.c_int, vp]
            f.restype = ctypes.c_int
        _lib = L
    return _lib


def orient2d(tuples, on_host=False) -> np.ndarray:
    a = np.ascontiguousarray(tuples, np.int32).reshape(-1, 6)
    out = np.zeros(len(a), np.int32)
    f = host().hostdev_orient2d if on_host else lib().devpred_orient2d
    assert f(a.ctypes.data, len(a), out.ctypes.data) == 0
    return out


def incircle(tuples, on_host=False) -> np.ndarray:
    a = np.ascontiguousarray(tuples, np.int32).reshape(-1, 8)
    out = np.zeros(len(a), np.int32)
    f = host().hostdev_incircle if on_host else lib().devpred_incircle
    assert f(a.ctypes.data, len(a), out.ctypes.data) == 0
    return out


def mul_s64(pairs, on_host=False):
    a = np.ascontiguousarray(pairs, np.int64).reshape(-1, 2)
    lo = np.zeros(len(a), np.uint64)
    hi = np.zeros(len(a), np.int64)
    f = host().hostdev_mul_s64 if on_host else lib().devpred_mul_s64
    assert f(a.ctypes.data, len(a), lo.ctypes.data, hi.ctypes.data) == 0
    return lo, hi


def _words(h):
    """(X, Y, W) Python ints -> the 5 int64 words of a homogeneous point."""
    x, y, w = (int(v) for v in h)
    m = (1 << 64) - 1

    def s64(v):
        v &= m
        return v - (1 << 64) if v >> 63 else v
    return [s64(x), s64(x >> 64), s64(y), s64(y >> 64), int(w)]


def orient_h(triples, on_host=False) -> np.ndarray:
    a = np.array([[w for h in t for w in _words(h)] for t in triples], np.int64).reshape(-1, 15)
    out = np.zeros(len(a), np.int32)
    f = host().hostdev_orient_h if on_host else lib().devpred_orient_h
    assert f(a.ctypes.data, len(a), out.ctypes.data) == 0
    return out


def incircle_h(quads, on_host=False) -> np.ndarray:
    a = np.array([[w for h in t for w in _words(h)] for t in quads], np.int64).reshape(-1, 20)
    out = np.zeros(len(a), np.int32)
    f = host().hostdev_incircle_h if on_host else lib().devpred_incircle_h
    assert f(a.ctypes.data, len(a), out.ctypes.data) == 0
    return out


def cross(segs, on_host=False) -> list:
    """segs: (n, 8) grid coordinates p1 p2 q1 q2 -> [(X, Y, W)] (W 0: parallel)."""
    a = np.ascontiguousarray(segs, np.int32).reshape(-1, 8)
    out = np.zeros((len(a), 5), np.int64)
    f = host().hostdev_cross if on_host else lib().devpred_cross
    assert f(a.ctypes.data, len(a), out.ctypes.data) == 0
    res = []
    for o in out.tolist():
        x = (o[1] << 64) | (o[0] & ((1 << 64) - 1))
        y = (o[3] << 64) | (o[2] & ((1 << 64) - 1))
        res.append((x, y, o[4]))
    return res
