"""Test infrastructure: CPU restatement of the reference's input path, used
only by tests/ as the checker of gcdt_read_file / gcdt_trans_points.

read_file(text, perm=None)  -- kernel.cu:5159-5291 readFile
trans_points(f, res)        -- kernel.cu:171-188 transPoints

Pinned by: a hand-worked known-answer file (tests/test_io.py), the closed
form of the normalisation (extremes land on 0.005 / 0.995 and the box centre
on 0.5), and invariance under the shuffle (the shuffled read is a
permutation of the plain read with constraints following their points).
The shuffle permutation itself is std::shuffle under std::mt19937(19999),
whose draws the C++ standard leaves to the library: it is taken from the
product as `perm` and only checked to be a permutation (parity unpinned for
the particular permutation, which differs between C++ libraries).
"""
from __future__ import annotations

import numpy as np


def _normalise(v: float, c: float, l: float) -> float:
    # kernel.cu:5227-5234: three rounded double operations in this order
    t = v - c
    t = t / l * 0.99
    return t + 0.5


def read_file(text: str, perm=None):
    tok = text.split()
    pos = 0

    def nxt():
        nonlocal pos
        pos += 1
        return tok[pos - 1]

    n_file = int(nxt())
    pts = []
    idxmap = [-1] * n_file
    seen = {}
    for _ in range(n_file):
        idx = int(nxt())
        x = float(nxt())
        y = float(nxt())
        if (x, y) not in seen:  # kernel.cu:5185-5197: first occurrence keeps its slot
            seen[(x, y)] = len(pts)
            idxmap[idx] = len(pts)
            pts.append((x, y))
        else:
            idxmap[idx] = seen[(x, y)]
    n = len(pts)
    if n:
        xs = [p[0] for p in pts]
        ys = [p[1] for p in pts]
        minx, maxx, miny, maxy = min(xs), max(xs), min(ys), max(ys)
    else:
        minx = maxx = miny = maxy = 0.0
    cx, cy = (minx + maxx) / 2, (miny + maxy) / 2
    lx = max(maxx - minx, maxy - miny)
    if lx == 0:
        lx = 1.0
    if perm is not None:  # point i moves to slot perm[i] (kernel.cu:5212-5227)
        moved = [None] * n
        for i in range(n):
            moved[perm[i]] = pts[i]
        pts = moved
    out = np.array([[_normalise(x, cx, lx), _normalise(y, cy, lx)] for x, y in pts], np.float64).reshape(-1, 2)
    cons = []
    have = set()
    if pos < len(tok):
        c_file = int(nxt())
        for _ in range(c_file):
            nxt()
            a = idxmap[int(nxt())]
            b = idxmap[int(nxt())]
            if perm is not None:
                a, b = perm[a], perm[b]
            if a == b or (a, b) in have or (b, a) in have:  # kernel.cu:5255-5269
                continue
            have.add((a, b))
            cons.append((a, b))
    return out, np.array(cons, np.int32).reshape(-1, 2)


def trans_points(f: np.ndarray, res: int = 1 << 30) -> np.ndarray:
    """kernel.cu:178-186: trunc(v*res) (toward zero), clamped to [0, res-1]."""
    v = np.trunc(np.asarray(f, np.float64) * float(res)).astype(np.int64)
    return np.clip(v, 0, res - 1).astype(np.int32)
