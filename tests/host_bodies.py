"""Builds and binds tests/host_bodies.cpp: the solve's per-item kernel bodies
executed on the host in seeded shuffled orders (test infrastructure)."""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "host_bodies.cpp")
LIB = os.path.join(HERE, "libhost_bodies.so")
# the same bodies unoptimised: every load the source names is made, in source
# order, so a read through an index that an optimised build would have
# dropped as unused still reaches memory (and faults on a guard page)
LIB_O0 = os.path.join(HERE, "libhost_bodies_O0.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _build_one(opt, out):
    cmd = [HIPCC, opt, "-fPIC", "-shared", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-o", out, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {SRC}:\n{r.stdout}\n{r.stderr}")
    return out


def build() -> str:
    _build_one("-O0", LIB_O0)
    return _build_one("-O2", LIB)


_lib = None


def lib():
    """The optimised build, or the unoptimised one when HOST_BODIES_O0=1 is
    set in the environment (child-process tests of memory accesses)."""
    global _lib
    if _lib is None:
        path = LIB_O0 if os.environ.get("HOST_BODIES_O0") == "1" else LIB
        if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(SRC):
            build()
        _lib = ctypes.CDLL(path)
        _lib.host_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_int]
        _lib.host_solve.restype = ctypes.c_int
        _lib.host_set_torn.argtypes = [ctypes.c_int]
        _lib.host_set_torn.restype = None
    return _lib


def set_torn(on: bool):
    """Torn-record mode of the packed flip rounds: a proposer whose record
    another proposal's flip rewrote earlier in the same apply pass reads the
    first 48 bytes new and the last 16 (meta, stamp, claim) old."""
    lib().host_set_torn(1 if on else 0)


def solve(g: np.ndarray, cons: np.ndarray, seed: int = 0, fused: bool = True, two_testers: bool = False):
    """(status, triangles (2n+2, 3), stats[8]).  two_testers: both triangles
    of an edge test it in every round (otherwise only in lists of at most 64
    triangles, as on the GPU with single_tester set)."""
    g = np.ascontiguousarray(g, np.int32)
    c = np.ascontiguousarray(cons, np.int32).reshape(-1, 2)
    n = len(g)
    out = np.zeros((2 * n + 2, 3), np.int32)
    st = np.zeros(8, np.int64)
    rc = lib().host_solve(g.ctypes.data, n, c.ctypes.data, len(c), out.ctypes.data, st.ctypes.data, seed,
                          1 if fused else 0, 1 if two_testers else 0)
    return rc, out, st
