"""Crossing constraints, restated exactly in Python integers (TEST
INFRASTRUCTURE ONLY: tests/ and bench.py's checker import it; the product
package gcdt_amd never does).

The reference has no result to restate here: its walk drops a segment that
meets a constrained edge (addConstraints.cu:768-774,942-946).  What
gcdt_solve_steiner must return is fixed by the mathematics instead: the
constrained Delaunay triangulation of the input points, the four corners
(oracle.with_corners) and every proper crossing of two constraint pieces, in
which each piece is a chain of edges through the crossings on it.

  pieces(pts, cons)                  constraints cut at the input points that
                                     lie exactly on them (the split pass)
  crossings(pts, pcs)                every proper crossing of two pieces, as
                                     reduced homogeneous (X, Y, W), W > 0
  validate(pts, cons, tris, hpts)    exact check of a result (T count, CCW,
                                     manifold edges, area, the crossing set,
                                     every sub-piece an edge, every other
                                     interior edge locally Delaunay)
  brute_force(pts, cons)             the CDT by its definition on tiny inputs
                                     (O(V^4)), crossing vertices numbered by
                                     sorted coordinates, as relabel() numbers
                                     a result's

Homogeneous points are (X, Y, W) with the point at (X/W, Y/W); a grid point is
(x, y, 1).  All predicates are signs of integer determinants.
"""
from __future__ import annotations

import itertools
from fractions import Fraction
from math import gcd

import numpy as np

from oracle.oracle import canonical, with_corners

R = (1 << 30) - 1


def _sgn(v: int) -> int:
    return (v > 0) - (v < 0)


def orient(p, q, r) -> int:
    """sign of | Xp Yp Wp ; Xq Yq Wq ; Xr Yr Wr | (+1 counter-clockwise)."""
    d = (p[0] * (q[1] * r[2] - r[1] * q[2]) - p[1] * (q[0] * r[2] - r[0] * q[2])
         + p[2] * (q[0] * r[1] - r[0] * q[1]))
    return _sgn(d)


def incircle(a, b, c, d) -> int:
    """+1 when d lies strictly inside the circle through counter-clockwise
    a, b, c: the lifted 3x3 determinant over points translated by d, in
    rationals scaled by positive W products."""
    rows = []
    for p in (a, b, c):
        # (x - xd, y - yd) * W_p W_d
        x = p[0] * d[2] - d[0] * p[2]
        y = p[1] * d[2] - d[1] * p[2]
        w = p[2] * d[2]
        rows.append((x * w, y * w, x * x + y * y))  # each row scaled by w^2 > 0
    (a0, a1, a2), (b0, b1, b2), (c0, c1, c2) = rows
    det = a0 * (b1 * c2 - b2 * c1) - a1 * (b0 * c2 - b2 * c0) + a2 * (b0 * c1 - b1 * c0)
    return _sgn(det)


def reduce_h(h) -> tuple[int, int, int]:
    x, y, w = (int(v) for v in h)
    if w < 0:
        x, y, w = -x, -y, -w
    g = gcd(gcd(abs(x), abs(y)), w)
    return x // g, y // g, w // g


def grid_h(p) -> tuple[int, int, int]:
    return int(p[0]), int(p[1]), 1


def _between(p, a, b) -> bool:
    """p collinear with a-b: strictly between them (exact, homogeneous)."""
    for k in (0, 1):
        pa = p[k] * a[2] - a[k] * p[2]  # sign(x_p - x_a)
        pb = p[k] * b[2] - b[k] * p[2]
        if _sgn(pa) != 0 or _sgn(pb) != 0:
            return _sgn(pa) * _sgn(pb) < 0
    return False


def pieces(pts_grid, cons) -> list[tuple[int, int]]:
    """Each constraint (a, b) cut at the input points strictly inside it, in
    order from a; duplicates kept once (as an unordered pair)."""
    P = [grid_h(p) for p in np.asarray(pts_grid).reshape(-1, 2)]
    out, seen = [], set()
    for a, b in np.asarray(cons).reshape(-1, 2):
        a, b = int(a), int(b)
        on = [v for v in range(len(P)) if v not in (a, b) and orient(P[a], P[b], P[v]) == 0
              and _between(P[v], P[a], P[b])]
        key = lambda v: (P[v][0] - P[a][0]) * (P[b][0] - P[a][0]) + (P[v][1] - P[a][1]) * (P[b][1] - P[a][1])
        chain = [a] + sorted(on, key=key) + [b]
        for u, v in zip(chain, chain[1:]):
            k = (min(u, v), max(u, v))
            if k not in seen:
                seen.add(k)
                out.append((u, v))
    return out


def cross_point(p1, p2, q1, q2):
    """Crossing of the lines p1p2 and q1q2 (grid points), reduced."""
    d1 = (p2[0] - p1[0], p2[1] - p1[1])
    d2 = (q2[0] - q1[0], q2[1] - q1[1])
    w = d1[0] * d2[1] - d1[1] * d2[0]
    t = (q1[0] - p1[0]) * d2[1] - (q1[1] - p1[1]) * d2[0]
    return reduce_h((p1[0] * w + d1[0] * t, p1[1] * w + d1[1] * t, w))


def crossings(pts_grid, pcs) -> set[tuple[int, int, int]]:
    """Every point where two pieces cross properly (interiors, not parallel)."""
    P = [grid_h(p) for p in np.asarray(pts_grid).reshape(-1, 2)]
    out = set()
    for (a, b), (c, d) in itertools.combinations(pcs, 2):
        if len({a, b, c, d}) < 4:
            continue
        o1, o2 = orient(P[a], P[b], P[c]), orient(P[a], P[b], P[d])
        o3, o4 = orient(P[c], P[d], P[a]), orient(P[c], P[d], P[b])
        if o1 * o2 < 0 and o3 * o4 < 0:
            out.add(cross_point(P[a], P[b], P[c], P[d]))
    return out


def relabel(n: int, tris, hpts):
    """Renumber crossing vertices n+4+k by their sorted (x, y): returns the
    triangles in that numbering and the sorted reduced points."""
    red = [reduce_h(h) for h in hpts]
    order = sorted(range(len(red)), key=lambda k: (Fraction(red[k][0], red[k][2]), Fraction(red[k][1], red[k][2])))
    new_id = {n + 4 + k: n + 4 + r for r, k in enumerate(order)}
    t = np.asarray(tris, dtype=np.int64).reshape(-1, 3).copy()
    for i in range(t.shape[0]):
        for j in range(3):
            t[i, j] = new_id.get(int(t[i, j]), int(t[i, j]))
    return t, [red[k] for k in order]


def _sub_edges(V, pcs, first_steiner):
    """The pieces cut at the vertices (crossings) lying on them."""
    out = []
    for a, b in pcs:
        on = [v for v in range(first_steiner, len(V)) if orient(V[a], V[b], V[v]) == 0 and _between(V[v], V[a], V[b])]
        key = lambda v: (Fraction(V[v][0], V[v][2]) - V[a][0]) * (V[b][0] - V[a][0]) + \
            (Fraction(V[v][1], V[v][2]) - V[a][1]) * (V[b][1] - V[a][1])
        chain = [a] + sorted(on, key=key) + [b]
        out.extend(zip(chain, chain[1:]))
    return out


def validate(pts_grid, cons, tris, hpts) -> dict:
    """Exact check of a crossing-resolved CDT.  tris use ids 0..n-1 for the
    points, n..n+3 for the corners and n+4+k for hpts[k] = (X, Y, W)."""
    pg = np.asarray(pts_grid).reshape(-1, 2)
    n = len(pg)
    pcs = pieces(pg, cons)
    want = crossings(pg, pcs)
    got = [reduce_h(h) for h in hpts]
    V = [grid_h(p) for p in with_corners(pg)] + [tuple(int(v) for v in h) for h in hpts]
    T = np.asarray(tris, dtype=np.int64).reshape(-1, 3)
    d = {"crossings_expected": len(want), "crossings_got": len(got),
         "crossings_ok": len(set(got)) == len(got) and set(got) == want,
         "count_ok": len(T) == 2 * (n + len(got)) + 2,
         "not_ccw": 0, "bad_edges": 0, "missing_constraints": 0, "not_delaunay": 0, "cocircular": 0}
    if T.size and (T.min() < 0 or T.max() >= len(V)):
        d["ids_ok"] = False
        d["valid"] = d["unique"] = False
        return d
    d["ids_ok"] = True
    area = Fraction(0)
    directed = {}
    for i, (a, b, c) in enumerate(T.tolist()):
        if orient(V[a], V[b], V[c]) <= 0:
            d["not_ccw"] += 1
        pa, pb, pc = V[a], V[b], V[c]
        det = (pa[0] * (pb[1] * pc[2] - pc[1] * pb[2]) - pa[1] * (pb[0] * pc[2] - pc[0] * pb[2])
               + pa[2] * (pb[0] * pc[1] - pc[0] * pb[1]))
        area += Fraction(det, pa[2] * pb[2] * pc[2])
        for k, (u, v) in enumerate(((a, b), (b, c), (c, a))):
            if (u, v) in directed:
                d["bad_edges"] += 1
            directed[(u, v)] = (i, (c, a, b)[k])  # triangle and the vertex opposite the edge
    d["area_ok"] = area == 2 * R * R
    hull = {(n, n + 1), (n + 1, n + 2), (n + 2, n + 3), (n + 3, n)}
    for (u, v) in directed:
        if (v, u) not in directed and (u, v) not in hull:
            d["bad_edges"] += 1
    cset = set()
    for u, v in _sub_edges(V, pcs, n + 4):
        if (u, v) not in directed and (v, u) not in directed:
            d["missing_constraints"] += 1
        cset.add((min(u, v), max(u, v)))
    for (u, v), (i, o) in directed.items():
        if u > v or (v, u) not in directed or (u, v) in cset:
            continue
        a, b, c = T[i].tolist()
        q = directed[(v, u)][1]
        s = incircle(V[a], V[b], V[c], V[q])
        if s > 0:
            d["not_delaunay"] += 1
        elif s == 0:
            d["cocircular"] += 1
    d["valid"] = (d["crossings_ok"] and d["count_ok"] and d["area_ok"] and d["not_ccw"] == 0
                  and d["bad_edges"] == 0 and d["missing_constraints"] == 0 and d["not_delaunay"] == 0)
    d["unique"] = d["valid"] and d["cocircular"] == 0
    return d


def _cross_open(p, q, r, s) -> bool:
    o1, o2 = orient(p, q, r), orient(p, q, s)
    o3, o4 = orient(r, s, p), orient(r, s, q)
    return o1 * o2 < 0 and o3 * o4 < 0


def brute_force(pts_grid, cons):
    """CDT of the points, the corners and the crossings, by definition: a
    counter-clockwise triangle is kept iff no sub-piece crosses one of its
    edges, no vertex lies inside or on it, and no vertex strictly inside its
    circumcircle is visible from its centroid.  Crossing vertices are
    numbered n+4+k in sorted (x, y) order.  O(V^4): tiny inputs only."""
    pg = np.asarray(pts_grid).reshape(-1, 2)
    n = len(pg)
    pcs = pieces(pg, cons)
    cr = sorted(crossings(pg, pcs), key=lambda h: (Fraction(h[0], h[2]), Fraction(h[1], h[2])))
    V = [grid_h(p) for p in with_corners(pg)] + cr
    sub = _sub_edges(V, pcs, n + 4)
    out = []
    for a, b, c in itertools.combinations(range(len(V)), 3):
        o = orient(V[a], V[b], V[c])
        if o == 0:
            continue
        tri = (a, b, c) if o > 0 else (a, c, b)
        edges = [(tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])]
        if any(_cross_open(V[u], V[v], V[s], V[t]) for (u, v) in edges for (s, t) in sub):
            continue
        # centroid, homogeneous: (sum X_i W_j W_k, ..., 3 W_a W_b W_c)
        pa, pb, pc = (V[i] for i in tri)
        g = (pa[0] * pb[2] * pc[2] + pb[0] * pa[2] * pc[2] + pc[0] * pa[2] * pb[2],
             pa[1] * pb[2] * pc[2] + pb[1] * pa[2] * pc[2] + pc[1] * pa[2] * pb[2],
             3 * pa[2] * pb[2] * pc[2])
        bad = False
        for dv in range(len(V)):
            if dv in tri:
                continue
            if all(orient(V[e[0]], V[e[1]], V[dv]) >= 0 for e in edges):
                bad = True
                break
            if incircle(pa, pb, pc, V[dv]) > 0:
                if not any(_cross_open(V[dv], g, V[s], V[t]) for (s, t) in sub):
                    bad = True
                    break
        if not bad:
            out.append(tri)
    return canonical(np.array(out, dtype=np.int64).reshape(-1, 3)), cr
