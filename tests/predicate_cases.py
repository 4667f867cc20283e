"""Shared inputs of the predicate parity tests (host and GPU): the golden
vectors (Python big-int determinants, tests/golden/predicates.json) and
64x64-bit products with exact values from Python integers."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def golden():
    with open(os.path.join(HERE, "golden", "predicates.json")) as f:
        g = json.load(f)
    o = np.array([sum(v["p"], []) for v in g["orient2d"]], np.int64)
    os_ = np.array([v["sign"] for v in g["orient2d"]], np.int32)
    i = np.array([sum(v["p"], []) for v in g["incircle"]], np.int64)
    is_ = np.array([v["sign"] for v in g["incircle"]], np.int32)
    return o, os_, i, is_


def product_pairs(count=4000, seed=3):
    """Signed factor pairs: the extremes of int64 and of the predicate's
    operand ranges (|x| < 2^62), zeros, ones, and random values of every
    magnitude, with every sign combination."""
    rng = np.random.default_rng(seed)
    ext = [0, 1, -1, 2, -2, (1 << 31) - 1, -(1 << 31), (1 << 32) - 1, 1 << 32, (1 << 61) - 1, -(1 << 61),
           (1 << 62) - 1, -(1 << 62), (1 << 63) - 1, -(1 << 63), -(1 << 63) + 1]
    pairs = [(a, b) for a in ext for b in ext]
    mags = rng.integers(0, 64, size=(count, 2))
    vals = rng.integers(0, 1 << 62, size=(count, 2), dtype=np.int64) >> (62 - np.minimum(mags, 62))
    signs = rng.choice([-1, 1], size=(count, 2))
    for (a, b), (sa, sb) in zip(vals.tolist(), signs.tolist()):
        pairs.append((sa * a, sb * b))
    return pairs


def exact_limbs(pairs):
    lo, hi = [], []
    for a, b in pairs:
        p = a * b
        lo.append(p & ((1 << 64) - 1))
        hi.append(p >> 64)
    return np.array(lo, np.uint64), np.array(hi, np.int64)
