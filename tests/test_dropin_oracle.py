"""Pins of the drop-in intermediate checks (oracle/dropin.py): hand-worked
cells, an independent exact clipping of the cell square by the triangle's
half-planes, brute force over every cell, and planted mistakes in the
sample, child and adjacency layouts."""
from fractions import Fraction

import numpy as np
import pytest

from gcdt_amd import gen
from oracle import dropin as D
from oracle import oracle as O

R = D.GRID_R


def test_sample_ids_closed_form():
    s = 1 << 22  # one 256x256 cell
    pts = np.array([[5 * s + 9, 3 * s + 1],   # cell (5,3)
                    [1, 2],                    # cell (0,0)
                    [5 * s, 3 * s],            # cell (5,3) again
                    [0, s - 1],                # cell (0,0) again
                    [s, 0]])                   # cell (1,0)
    assert D.sample_ids(pts).tolist() == [1, 4, 0]  # cells (0,0), (1,0), (5,3): smallest id of each
    assert len(D.sample_ids(np.zeros((0, 2)))) == 0


def _tri(a, b, c):
    return (np.array([a], np.int64), np.array([b], np.int64), np.array([c], np.int64))


def _meets(a, b, c, box):
    A, B, C = _tri(a, b, c)
    x0, y0, x1, y1 = (np.array([v], np.int64) for v in box)
    return bool(D.tri_meets_cell(A, B, C, x0, y0, x1, y1)[0])


def test_tri_meets_cell_hand_worked():
    box = (8, 8, 15, 15)
    assert _meets((9, 9), (14, 9), (9, 14), box)           # inside
    assert _meets((0, 0), (100, 0), (0, 100), box)         # box inside the triangle
    assert _meets((15, 15), (30, 15), (15, 30), box)       # a vertex on the box corner (closed)
    assert _meets((0, 30), (30, 0), (30, 30), box)         # hypotenuse x + y = 30 through (15,15)
    assert not _meets((0, 31), (31, 0), (31, 31), box)     # x + y = 31 misses (15,15) by one unit
    assert _meets((0, 10), (20, 10), (10, 20), box)        # an edge crosses, no vertex inside
    assert not _meets((0, 0), (7, 0), (0, 20), box)        # left of the box
    assert _meets((0, 0), (8, 0), (8, 20), box)            # an edge along x = 8 touches


def _clip_nonempty(tri, box):
    """Independent: clip the closed box by the triangle's three closed
    half-planes with exact rationals (Sutherland-Hodgman); a point or a
    segment left over counts as meeting."""
    x0, y0, x1, y1 = box
    poly = [(Fraction(x0), Fraction(y0)), (Fraction(x1), Fraction(y0)), (Fraction(x1), Fraction(y1)),
            (Fraction(x0), Fraction(y1))]
    P = [tuple(map(Fraction, v)) for v in tri]
    for e in range(3):
        (ax, ay), (bx, by) = P[e], P[(e + 1) % 3]

        def side(q):
            return (bx - ax) * (q[1] - ay) - (by - ay) * (q[0] - ax)

        out = []
        for i in range(len(poly)):
            cur, nxt = poly[i], poly[(i + 1) % len(poly)]
            sc, sn = side(cur), side(nxt)
            if sc >= 0:
                out.append(cur)
            if (sc > 0 > sn) or (sc < 0 < sn):
                t = sc / (sc - sn)
                out.append((cur[0] + t * (nxt[0] - cur[0]), cur[1] + t * (nxt[1] - cur[1])))
        poly = out
        if not poly:
            return False
    return True


def test_tri_meets_cell_equals_exact_clipping():
    rng = np.random.default_rng(5)
    n = 0
    for _ in range(3000):
        p = rng.integers(0, 17, size=(3, 2))
        o = (p[1, 0] - p[0, 0]) * (p[2, 1] - p[0, 1]) - (p[1, 1] - p[0, 1]) * (p[2, 0] - p[0, 0])
        if o == 0:
            continue
        if o < 0:
            p = p[[0, 2, 1]]
        x0, y0 = rng.integers(0, 15, size=2)
        box = (int(x0), int(y0), int(x0 + rng.integers(0, 4)), int(y0 + rng.integers(0, 4)))
        tri = [tuple(map(int, v)) for v in p]
        assert _meets(*tri, box) == _clip_nonempty(tri, box), (tri, box)
        n += 1
    assert n > 2000


def test_bitmap_cells_equal_brute_force_over_every_cell():
    """At a 4x4 cell grid (cells of 2^28), the bounding-box enumeration must
    find exactly the cells that a test of every triangle against every cell
    finds."""
    rng = np.random.default_rng(2)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(60, 2)))
    g = gen.to_grid(p)
    tris, _ = O.cdt(g)
    pts = O.with_corners(g)
    off, ids = D.bitmap_cells(pts, tris, log_bitmap=2)
    sh = 28
    for cell in range(16):
        cx, cy = cell >> 2, cell & 3
        T = len(tris)
        A, B, C = pts[tris[:, 0]], pts[tris[:, 1]], pts[tris[:, 2]]
        hit = D.tri_meets_cell(A, B, C, np.full(T, cx << sh), np.full(T, cy << sh), np.full(T, ((cx + 1) << sh) - 1),
                               np.full(T, ((cy + 1) << sh) - 1))
        assert ids[off[cell]:off[cell + 1]].tolist() == np.nonzero(hit)[0].tolist(), cell
    # every cell of the square is met by some triangle (the mesh tiles it)
    assert all(off[c + 1] > off[c] for c in range(16))


def _split_dag():
    """Square corners 0..3 (ids n..n+3 with n = 2 points), one interior split
    and one edge split, laid out as rawMesh.cu:209-276."""
    s = 1 << 28
    # point 1 lies on the diagonal x + y = R between corners (R,0) and (0,R)
    pts = np.array([[s, s], [2 * s, R - 2 * s], [0, 0], [R, 0], [R, R], [0, R]], np.int64)
    # initial triangles (n, n+1, n+3) and (n+1, n+2, n+3) with n = 2
    tris = [(2, 3, 5), (3, 4, 5)]
    sons = [-1, -1]
    # point 0 strictly inside triangle 0 -> children 2,3,4 = (q,y,z),(q,z,x),(q,x,y)
    x, y, z = tris[0]
    tris += [(0, y, z), (0, z, x), (0, x, y)]
    sons[0] = 2
    sons += [-1, -1, -1]
    # point 1 on the diagonal (3,5): slot 2 = (0,3,5) holds the edge opposite vertex 0
    # t = (o,a,b) = (0,3,5) -> (0,3,1),(0,1,5); n = slot 1 = (3,4,5) rotated (4,5,3) -> (4,5,1),(4,1,3)
    tris += [(0, 3, 1), (0, 1, 5), (4, 5, 1), (4, 1, 3)]
    sons[2] = -5
    sons[1] = -7
    sons += [-1, -1, -1, -1]
    return pts, np.array(tris, np.int64), np.array(sons, np.int64)


def test_sons_layout_checker():
    pts, tris, sons = _split_dag()
    assert D.sons_violations(pts, tris, sons, len(tris)) == []
    bad = tris.copy()
    bad[[2, 3]] = bad[[3, 2]]  # interior children in the wrong order
    assert D.sons_violations(pts, bad, sons, len(tris))
    bad = tris.copy()
    bad[5] = (0, 3, 0)  # wrong inserted point in an edge child
    assert D.sons_violations(pts, bad, sons, len(tris))
    bad_s = sons.copy()
    bad_s[1] = -5  # the neighbour's children point at the other side's pair
    assert D.sons_violations(pts, tris, bad_s, len(tris))


def _brute_adj(tris):
    where = {}
    for i, t in enumerate(tris.tolist()):
        for k in range(3):
            where[(t[(k + 1) % 3], t[(k + 2) % 3])] = i
    adj = np.full(tris.shape, -1, np.int64)
    for i, t in enumerate(tris.tolist()):
        for k in range(3):
            adj[i, k] = where.get((t[(k + 2) % 3], t[(k + 1) % 3]), -1)
    return adj


def test_adjacency_checker():
    rng = np.random.default_rng(8)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(300, 2)))
    g = gen.to_grid(p)
    tris, _ = O.cdt(g)
    pts = O.with_corners(g)
    adj = _brute_adj(tris)
    assert D.adj_violations(pts, tris, adj) == []
    interior = np.argwhere(adj >= 0)[7]
    bad = adj.copy()
    bad[tuple(interior)] = -1  # an interior edge with no neighbour
    assert D.adj_violations(pts, tris, bad)
    bad = adj.copy()
    i, k = interior
    n = adj[i, k]
    bad[i, k] = adj[n, 0] if adj[n, 0] not in (i, -1) else adj[n, 1]  # a wrong neighbour
    assert D.adj_violations(pts, tris, bad)
    live = np.ones(len(tris), bool)
    live[n] = False  # pointing at a dead slot
    assert D.adj_violations(pts, tris, adj, live)
