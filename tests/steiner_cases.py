"""Inputs with crossing constraints and their checks, shared by the CPU
(emulator, oracle) and GPU tests of gcdt_solve_steiner."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from oracle import steiner as OS

R = (1 << 30) - 1


def random_crossing(seed: int, n: int, k: int):
    """n random grid points and k segments between random pairs of them."""
    rng = np.random.default_rng(seed)
    g = rng.integers(1, R, size=(n, 2)).astype(np.int32)
    cons = []
    while len(cons) < k:
        a, b = (int(v) for v in rng.integers(0, n, 2))
        if a != b:
            cons.append((a, b))
    return g, np.array(cons, np.int32)


def star(seed: int, arms: int = 4, extra: int = 5):
    """`arms` segments through one grid point that is not an input point."""
    rng = np.random.default_rng(seed)
    P = np.array([1 << 29, (1 << 29) + 12345], np.int64)
    pts, cons = [], []
    for _ in range(arms):
        d = rng.integers(-(1 << 24), 1 << 24, size=2)
        pts += [P - d, P + d]
        cons.append((len(pts) - 2, len(pts) - 1))
    g = np.concatenate([np.array(pts, np.int64), rng.integers(1, R, size=(extra, 2))]).astype(np.int32)
    return g, np.array(cons, np.int32)


def hd_family(seed: int, nh: int, nd: int, extra: int):
    """nh horizontal segments y = c_i and nd diagonal ones x - y = d_j, each
    crossing every one of the other family at the integer point (c_i + d_j,
    c_i), plus `extra` random points (none on a segment).  The crossings are
    grid points, so the result must equal the ordinary CDT of the points plus
    the crossings with every segment cut at them."""
    rng = np.random.default_rng(seed)
    span = 1 << 27
    ys = np.sort(rng.choice(np.arange(3 * span, 5 * span), size=nh, replace=False))
    ds = np.sort(rng.choice(np.arange(-span // 2, span // 2), size=nd, replace=False))
    x0, x1 = 2 * span - (3 << 20), 6 * span + (3 << 20)       # horizontals: x in [x0, x1]
    y0, y1 = 3 * span - (1 << 22), 5 * span + (1 << 22)       # diagonals: y in [y0, y1]
    pts, cons = [], []
    for y in ys:
        pts += [(x0 + int(rng.integers(0, 1 << 20)), int(y)), (x1 - int(rng.integers(0, 1 << 20)), int(y))]
        cons.append((len(pts) - 2, len(pts) - 1))
    for d in ds:
        a, b = y0 - int(rng.integers(0, 1 << 20)), y1 + int(rng.integers(0, 1 << 20))
        pts += [(a + int(d), a), (b + int(d), b)]
        cons.append((len(pts) - 2, len(pts) - 1))
    P = [(x, y, 1) for x, y in pts]
    while len(pts) < 2 * (nh + nd) + extra:
        q = (int(rng.integers(1, R)), int(rng.integers(1, R)))
        if all(OS.orient(P[a], P[b], (q[0], q[1], 1)) != 0 for a, b in cons):
            pts.append(q)
    return np.array(pts, np.int32), np.array(cons, np.int32)


def augmented_cdt(g, cons):
    """For inputs whose crossings are all grid points: the ordinary CDT of the
    points plus the crossings (oracle.cdt), in the numbering of a crossing
    result sorted by oracle.steiner.relabel (points, corners, crossings)."""
    g = np.asarray(g, np.int64).reshape(-1, 2)
    n = len(g)
    pcs = OS.pieces(g, cons)
    cr = OS.crossings(g, pcs)
    assert all(h[2] == 1 for h in cr), "crossings must be grid points"
    cr = sorted(cr)  # W = 1: plain (x, y) order, as relabel's
    S = len(cr)
    ga = np.concatenate([g, np.array([(x, y) for x, y, _ in cr], np.int64).reshape(-1, 2)]).astype(np.int32)
    V = [OS.grid_h(p) for p in O.with_corners(ga)]
    # every piece cut at the crossings on it (ids in ga's numbering)
    sub = []
    for a, b in pcs:
        on = [n + k for k in range(S) if OS.orient(V[a], V[b], V[n + k]) == 0 and OS._between(V[n + k], V[a], V[b])]
        gi = [[int(ga[i][0]), int(ga[i][1])] for i in (a, b)]
        key = lambda v: (int(ga[v][0]) - gi[0][0]) * (gi[1][0] - gi[0][0]) + (int(ga[v][1]) - gi[0][1]) * (gi[1][1] - gi[0][1])
        chain = [a] + sorted(on, key=key) + [b]
        sub.extend(zip(chain, chain[1:]))
    sub = np.array(sub, np.int32).reshape(-1, 2)
    tris, _ = O.cdt(ga, sub)
    v = O.validate(ga, tris, sub)
    # ga ids: points 0..n-1, crossings n..n+S-1, corners n+S..n+S+3
    remap = np.arange(n + S + 4)
    remap[n:n + S] = np.arange(n + 4, n + 4 + S)
    remap[n + S:] = np.arange(n, n + 4)
    return O.canonical(remap[np.asarray(tris, np.int64)]), v


def check(g, cons, tris, hpts, brute=False, augmented=False, segs=None):
    """Exact checks of a crossing result; returns the validator's dict."""
    v = OS.validate(g, cons, tris, hpts)
    assert v["valid"], {k: v[k] for k in v if k != "valid"}
    n = len(np.asarray(g).reshape(-1, 2))
    if segs is not None:
        # the two pieces named for each crossing pass through it
        P = [OS.grid_h(p) for p in np.asarray(g).reshape(-1, 2)]
        for h, s in zip(hpts, np.asarray(segs).reshape(-1, 4)):
            for a, b in ((s[0], s[1]), (s[2], s[3])):
                assert 0 <= a < n and 0 <= b < n and OS.orient(P[a], P[b], tuple(h)) == 0
    if brute and v["unique"]:
        bf, _ = OS.brute_force(g, cons)
        t2, _ = OS.relabel(n, tris, hpts)
        assert np.array_equal(O.canonical(t2), bf)
    if augmented:
        ref, va = augmented_cdt(g, cons)
        if va["unique"]:
            t2, _ = OS.relabel(n, tris, hpts)
            assert np.array_equal(O.canonical(t2), ref)
    return v


def rand_crossings(rng, k):
    """k crossing points of random long segments (W up to about 2^60)."""
    out = []
    while len(out) < k:
        s = rng.integers(1, R, size=(4, 2))
        P = [OS.grid_h(p) for p in s]
        if OS.orient(P[0], P[1], P[2]) * OS.orient(P[0], P[1], P[3]) < 0 and \
                OS.orient(P[2], P[3], P[0]) * OS.orient(P[2], P[3], P[1]) < 0:
            out.append(OS.cross_point(*[tuple(int(v) for v in p) for p in s]))
    return out


def homogeneous_cases(seed: int = 3):
    """Point triples and quadruples for orient_h / incircle_h with their exact
    signs: crossing points, mixtures with grid points, a crossing on its own
    segments (orient 0) and scaled rectangles (cocircular, incircle 0)."""
    rng = np.random.default_rng(seed)
    pts = rand_crossings(rng, 400)
    triples = [pts[i:i + 3] for i in range(0, 399, 3)]
    quads = [pts[i:i + 4] for i in range(0, 396, 4)]
    g = [OS.grid_h(p) for p in rng.integers(0, R + 1, size=(200, 2))]
    for i in range(100):
        triples.append((pts[i], g[i], g[i + 100]))
        quads.append((g[i], pts[i + 1], g[i + 100], pts[i + 2]))
    for _ in range(60):
        s = rng.integers(1, R, size=(4, 2))
        P = [OS.grid_h(p) for p in s]
        if OS.orient(P[0], P[1], P[2]) * OS.orient(P[0], P[1], P[3]) < 0 and \
                OS.orient(P[2], P[3], P[0]) * OS.orient(P[2], P[3], P[1]) < 0:
            x = OS.cross_point(*[tuple(int(v) for v in p) for p in s])
            triples.append((P[0], P[1], x))
            triples.append((x, P[2], P[3]))
    for _ in range(40):
        x0, y0 = (int(v) for v in rng.integers(1 << 20, R - (1 << 22), size=2))
        w, h = (int(v) for v in rng.integers(1, 1 << 21, size=2))
        k = int(rng.integers(2, 1 << 30))
        quads.append(((x0 * k, y0 * k, k), ((x0 + w) * k, y0 * k, k), ((x0 + w) * k, (y0 + h) * k, k),
                      (x0 * k, (y0 + h) * k, k)))
    return triples, quads, [OS.orient(*t) for t in triples], [OS.incircle(*q) for q in quads]


def crossing_segments(seed: int = 4):
    """Random segment pairs (p1 p2 q1 q2, grid), some parallel."""
    rng = np.random.default_rng(seed)
    segs = rng.integers(0, R + 1, size=(300, 8))
    segs[:20, 6:8] = segs[:20, 4:6] + (segs[:20, 2:4] - segs[:20, 0:2]) // 3
    return np.clip(segs, 0, R)


def check_crossing_points(segs, got):
    """got[i] = (X, Y, W) of devpred.cross: the reduced crossing, W 0 when parallel."""
    for s, h in zip(np.asarray(segs).tolist(), got):
        p1, p2, q1, q2 = (s[0], s[1]), (s[2], s[3]), (s[4], s[5]), (s[6], s[7])
        w = (p2[0] - p1[0]) * (q2[1] - q1[1]) - (p2[1] - p1[1]) * (q2[0] - q1[0])
        if w == 0:
            assert h[2] == 0
            continue
        assert h[2] > 0 and OS.reduce_h(h) == OS.cross_point(p1, p2, q1, q2)
        assert abs(h[0]) < 1 << 92 and abs(h[1]) < 1 << 92 and h[2] < 1 << 61
