"""Checks of the drop-in pipeline's intermediate outputs (TEST INFRASTRUCTURE
ONLY: imported by tests/ as the checker, never by the product package).

Restated from the reference:

  sample_ids(pts)              getSamplePoints, rawMesh.cu:678-746: one point
                               per occupied 256x256 cell of the 2^30 grid
                               (cell = (x >> 22, y >> 22)), compacted in cell
                               order (ad = cx * 256 + cy) by a scan.  The
                               reference keeps an arbitrary point of the cell
                               (last writer, rawMesh.cu:687); the build keeps
                               the smallest id, one of the reference's
                               possible answers, and this is its closed form.
  tri_meets_cell(...)          intersect(), boostmap.cu:30-66: closed triangle
                               against the closed cell square
                               [a*cx, a*(cx+1)-1] x [a*cy, a*(cy+1)-1]
                               (boostmap.cu:100-101): a square corner in the
                               triangle, a triangle edge across the square, or a
                               triangle vertex in the square.
  bitmap_cells(pts, tris)      getBitmap, boostmap.cu:146-181: for every
                               triangle, the cells of its bounding box
                               (>> 20) that it meets; returned as the sorted
                               triangle ids of every cell (CSR: offsets over
                               2^20 cells, ad = cx * 1024 + cy).
  sons_violations(...)         the insertion DAG's child layout,
                               rawMesh.cu:209-276: sons[t] > 0 is the first of
                               3 children (q,y,z),(q,z,x),(q,x,y) of t = (x,y,z)
                               split by q strictly inside; sons[t] < 0 is minus
                               the first of 2 children (o,a,q),(o,q,b) of
                               t = (o,a,b) split by q strictly inside edge (a,b).
  adj_violations(...)          testAdjTris, rawMesh.cu:931-992: adjTris[t][k]
                               is the triangle across the edge opposite vertex
                               k, holding that edge and pointing back at t;
                               -1 only on the square's boundary.

Pinned by tests/test_dropin_oracle.py: hand-worked cells, an independent
rational clipping of the square by the triangle's half-planes on small
coordinates, and planted mistakes in each layout that the checks must
report.
"""
from __future__ import annotations

import numpy as np

GRID_R = (1 << 30) - 1
LOG_SAMPLE = 8   # math_base.cuh:24 logSample
LOG_BITMAP = 10  # math_base.cuh:25 logBitmap


def sample_ids(pts) -> np.ndarray:
    """Smallest point id of every occupied 256x256 cell, in cell order."""
    p = np.asarray(pts, np.int64).reshape(-1, 2)
    if len(p) == 0:
        return np.zeros(0, np.int64)
    sh = 30 - LOG_SAMPLE
    cell = ((p[:, 0] >> sh) << LOG_SAMPLE) + (p[:, 1] >> sh)
    order = np.lexsort((np.arange(len(p)), cell))
    c = cell[order]
    first = np.ones(len(c), bool)
    first[1:] = c[1:] != c[:-1]
    return order[first]


def _orient(ax, ay, bx, by, cx, cy):
    """sign((a-c) x (b-c)) elementwise on int64 arrays (|coord| < 2^31)."""
    o1 = (ax - cx) * (by - cy)
    o2 = (ay - cy) * (bx - cx)
    return np.sign(o1 - o2)


def _edge_meets_box(ax, ay, bx, by, x0, y0, x1, y1):
    """boostmap.cu:30-39: segment a-b against the closed box."""
    out = ((ax < x0) & (bx < x0)) | ((ax > x1) & (bx > x1)) | ((ay < y0) & (by < y0)) | ((ay > y1) & (by > y1))
    f = ((bx >= ax) & (by >= ay)) | ((bx <= ax) & (by <= ay))
    # corners p[0]=(x0,y0) p[1]=(x1,y0) p[2]=(x1,y1) p[3]=(x0,y1); f picks p[f], p[f+2]
    px1 = np.where(f, x1, x0)
    py1 = y0
    px2 = np.where(f, x0, x1)
    py2 = y1
    s = _orient(ax, ay, bx, by, px1, py1) * _orient(ax, ay, bx, by, px2, py2)
    return ~out & (s <= 0)


def tri_meets_cell(A, B, C, x0, y0, x1, y1):
    """boostmap.cu:44-66 over arrays: closed CCW triangle ABC (each (n,2))
    against closed boxes [x0,x1] x [y0,y1]."""
    ax, ay, bx, by, cx, cy = A[:, 0], A[:, 1], B[:, 0], B[:, 1], C[:, 0], C[:, 1]
    hit = np.zeros(len(ax), bool)
    for qx, qy in ((x0, y0), (x1, y0), (x1, y1), (x0, y1)):
        hit |= (_orient(ax, ay, bx, by, qx, qy) >= 0) & (_orient(bx, by, cx, cy, qx, qy) >= 0) & \
               (_orient(cx, cy, ax, ay, qx, qy) >= 0)
    hit |= _edge_meets_box(ax, ay, bx, by, x0, y0, x1, y1)
    hit |= _edge_meets_box(bx, by, cx, cy, x0, y0, x1, y1)
    hit |= _edge_meets_box(cx, cy, ax, ay, x0, y0, x1, y1)
    for vx, vy in ((ax, ay), (bx, by), (cx, cy)):
        hit |= (x0 <= vx) & (vx <= x1) & (y0 <= vy) & (vy <= y1)
    return hit


def bitmap_cells(pts, tris, log_bitmap: int = LOG_BITMAP):
    """(offsets int64[4^L + 1], ids int64[...]): for each cell in order, the
    sorted ids of the triangles meeting it."""
    p = np.asarray(pts, np.int64).reshape(-1, 2)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    sh = 30 - log_bitmap
    A, B, C = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    lo = np.minimum(np.minimum(A, B), C) >> sh
    hi = np.maximum(np.maximum(A, B), C) >> sh
    w = hi[:, 0] - lo[:, 0] + 1
    h = hi[:, 1] - lo[:, 1] + 1
    cnt = w * h
    tid = np.repeat(np.arange(len(t)), cnt)
    start = np.repeat(np.cumsum(cnt) - cnt, cnt)
    local = np.arange(int(cnt.sum())) - start
    cx = lo[tid, 0] + local // h[tid]
    cy = lo[tid, 1] + local % h[tid]
    a = 1 << sh
    ok = tri_meets_cell(A[tid], B[tid], C[tid], a * cx, a * cy, a * (cx + 1) - 1, a * (cy + 1) - 1)
    cell = (cx << log_bitmap) + cy
    cell, tid = cell[ok], tid[ok]
    order = np.lexsort((tid, cell))
    cell, tid = cell[order], tid[order]
    offsets = np.zeros((1 << (2 * log_bitmap)) + 1, np.int64)
    np.add.at(offsets, cell + 1, 1)
    return np.cumsum(offsets), tid


def sons_violations(pts, tris, sons, nslots) -> list:
    """Every parent slot whose children break the layout of rawMesh.cu:209-276
    (as (slot, reason)).  Leaves have sons == -1; the initial two triangles
    and every child slot lie below nslots."""
    p = np.asarray(pts, np.int64).reshape(-1, 2)
    t = np.asarray(tris, np.int64).reshape(-1, 3)[:nslots]
    s = np.asarray(sons, np.int64)[:nslots]
    bad = []

    def orient(a, b, c):
        return int(np.sign((p[a, 0] - p[c, 0]) * (p[b, 1] - p[c, 1]) - (p[a, 1] - p[c, 1]) * (p[b, 0] - p[c, 0])))

    for i in np.nonzero(s != -1)[0].tolist():
        x, y, z = t[i].tolist()
        if s[i] > 0:
            c = int(s[i])
            if c + 2 >= nslots:
                bad.append((i, "children out of range"))
                continue
            q = int(t[c, 0])
            want = [(q, y, z), (q, z, x), (q, x, y)]
            if [tuple(t[c + j].tolist()) for j in range(3)] != want:
                bad.append((i, "interior children"))
            elif not (orient(y, z, q) > 0 and orient(z, x, q) > 0 and orient(x, y, q) > 0):
                bad.append((i, "inserted point not strictly inside"))
        elif s[i] < -1 or s[i] == 0:
            c = int(-s[i])
            if s[i] == 0 or c + 1 >= nslots:
                bad.append((i, "children out of range"))
                continue
            o, a, q = t[c].tolist()
            o2, q2, b = t[c + 1].tolist()
            rot = [(x, y, z), (y, z, x), (z, x, y)]
            if o2 != o or q2 != q or (o, a, b) not in rot:
                bad.append((i, "edge children"))
            elif orient(a, b, q) != 0 or not (
                    (p[q] - p[a]) @ (p[b] - p[a]) > 0 and (p[q] - p[b]) @ (p[a] - p[b]) > 0):
                bad.append((i, "inserted point not strictly inside the edge"))
        else:
            bad.append((i, "bad sons value"))
    return bad


def adj_violations(pts, tris, adj, live=None) -> list:
    """testAdjTris (rawMesh.cu:931-992) over the triangles with live[t] (all
    when None): (triangle, slot, reason) for every entry that is not the
    triangle across that edge pointing back, or -1 off the square's boundary."""
    p = np.asarray(pts, np.int64).reshape(-1, 2)
    t = np.asarray(tris, np.int64).reshape(-1, 3)
    a = np.asarray(adj, np.int64).reshape(-1, 3)
    idx = np.arange(len(t)) if live is None else np.nonzero(np.asarray(live))[0]
    bad = []
    for i in idx.tolist():
        for k in range(3):
            u, v = int(t[i, (k + 1) % 3]), int(t[i, (k + 2) % 3])
            n = int(a[i, k])
            if n == -1:
                pu, pv = p[u], p[v]
                on_side = (pu[0] == pv[0] and pu[0] in (0, GRID_R)) or (pu[1] == pv[1] and pu[1] in (0, GRID_R))
                if not on_side:
                    bad.append((i, k, "no neighbour off the boundary"))
                continue
            if n < 0 or n >= len(t) or (live is not None and not live[n]):
                bad.append((i, k, "neighbour id"))
                continue
            tn = t[n].tolist()
            if u not in tn or v not in tn:
                bad.append((i, k, "neighbour lacks the edge"))
                continue
            j = [m for m in range(3) if tn[m] != u and tn[m] != v][0]
            if int(a[n, j]) != i:
                bad.append((i, k, "not symmetric"))
    return bad
