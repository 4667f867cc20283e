"""Builds tests/emu/_build/libgcdt_emu.so: the library's HIP sources compiled
as host C++ against the CPU SIMT emulator (tests/emu/emu_rt.cpp and the
stand-in tests/emu/hip/hip_runtime.h).  Test infrastructure: it exports the
same C ABI as libgcdt.so plus emu_set_tunable / emu_launches, and runs only
in CPU tests.

Each kernel launch `k<<<grid, block, shmem, stream>>>(args)` is rewritten to
`::emu::launch("k", ::emu::LaunchCfg(grid, block, shmem, stream), &k, args)`,
and each `__shared__ T name[N];` to a reference to per-workgroup storage;
nothing else in the sources changes.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "gcdt_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libgcdt_emu.so")
SOURCES = ["mesh.hip", "cons.hip", "carve.hip", "io.hip", "noalloc.hip", "steiner.hip", "capi.cpp"]
HEADERS = ["gcdt_common.hpp", "predicates.hpp", "mesh_bodies.hpp", "cons_bodies.hpp", "mesh.hpp", "exact.hpp", "steiner_bodies.hpp"]
CXX = os.environ.get("CXX_EMU", "g++")
FLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-Wno-unknown-pragmas", "-Wno-attributes", "-fno-strict-aliasing",
         "-I", HERE]


def _kernel_name_start(s: str, i: int) -> int:
    """Start of the kernel expression that ends right before s[i] ('<<<')."""
    j = i - 1
    while j >= 0 and s[j].isspace():
        j -= 1
    if s[j] == ">":  # template arguments
        depth = 0
        while j >= 0:
            if s[j] == ">":
                depth += 1
            elif s[j] == "<":
                depth -= 1
                if depth == 0:
                    break
            j -= 1
        j -= 1
    while j >= 0 and (s[j].isalnum() or s[j] in "_:"):
        j -= 1
    return j + 1


def rewrite_launches(s: str) -> str:
    out = []
    pos = 0
    while True:
        i = s.find("<<<", pos)
        if i < 0:
            out.append(s[pos:])
            break
        a = _kernel_name_start(s, i)
        name = s[a:i].strip()
        # the configuration: up to the '>>>' at parenthesis depth 0
        depth = 0
        j = i + 3
        while j < len(s):
            c = s[j]
            if c in "([{":
                depth += 1
            elif c in ")]}":
                depth -= 1
            elif depth == 0 and s.startswith(">>>", j):
                break
            j += 1
        cfg = s[i + 3:j]
        k = j + 3
        while s[k].isspace():
            k += 1
        if s[k] != "(":
            raise ValueError(f"launch of {name} without an argument list")
        m = k + 1
        while s[m].isspace():
            m += 1
        sep = "" if s[m] == ")" else ", "
        out.append(s[pos:a])
        out.append(f'::emu::launch("{name}", ::emu::LaunchCfg({cfg}), &{name}{sep}')
        pos = k + 1
    return "".join(out)


_SHARED = re.compile(r"__shared__\s+((?:unsigned\s+)?(?:long\s+long|int|char|short)|int[234]|[A-Za-z_][\w:]*)\s+([^;]+);")
_lds_id = [0]


def rewrite_shared(s: str) -> str:
    """`__shared__ T a[N], b;` -> `auto& a = ::emu::lds<T[N], id>(); auto& b = ...;`"""
    def one(m):
        ty, decls = m.group(1), m.group(2)
        out = []
        for d in decls.split(","):
            d = d.strip()
            name = re.match(r"\w+", d).group(0)
            dims = d[len(name):].strip()
            _lds_id[0] += 1
            out.append(f"auto& {name} = ::emu::lds<{ty}{dims}, {_lds_id[0]}>();")
        return " ".join(out)
    return _SHARED.sub(one, s)


def _compile(src_path: str, obj: str, extra=()):
    cmd = [CXX, *FLAGS, *extra, "-x", "c++", "-c", src_path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emulator build failed for {src_path}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False, defines=(), lib: str = LIB) -> str:
    """defines: extra -D macros (a variant library, built in its own folder)."""
    out = OUT if not defines else os.path.join(OUT, "v_" + "_".join(d.replace("=", "") for d in defines))
    extra = [f"-D{d}" for d in defines]
    gen = os.path.join(out, "src")
    os.makedirs(gen, exist_ok=True)
    os.makedirs(os.path.join(OUT, "include"), exist_ok=True)
    # headers are copied unchanged beside the rewritten sources (relative
    # includes of ../../include resolve to the repository's include/)
    for h in HEADERS:
        with open(os.path.join(CSRC, h)) as f:
            txt = f.read()
        txt = txt.replace('"../../include/', f'"{ROOT}/include/')
        with open(os.path.join(gen, h), "w") as f:
            f.write(rewrite_shared(rewrite_launches(txt)))
    jobs = []
    for src in SOURCES:
        with open(os.path.join(CSRC, src)) as f:
            txt = f.read()
        txt = txt.replace('"../../include/', f'"{ROOT}/include/')
        dst = os.path.join(gen, os.path.splitext(src)[0] + ".cpp")
        with open(dst, "w") as f:
            f.write(rewrite_shared(rewrite_launches(txt)))
        jobs.append((dst, os.path.join(out, os.path.splitext(src)[0] + ".o"), extra))
    rt = os.path.join(HERE, "emu_rt.cpp")
    jobs.append((rt, os.path.join(out, "emu_rt.o"), ["-I", gen, *extra]))
    with ThreadPoolExecutor(max_workers=min(6, len(jobs))) as ex:
        objs = list(ex.map(lambda j: _compile(*j), jobs))
    tmp = lib + ".tmp"
    r = subprocess.run([CXX, "-shared", "-o", tmp, *objs, "-lpthread"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"emulator link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, lib)
    if verbose:
        print("built", lib)
    return lib


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
