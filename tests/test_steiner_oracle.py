"""Pins of the crossing-constraint oracle (oracle/steiner.py) and of the
homogeneous predicates and crossing points of gcdt_amd/csrc/exact.hpp (its
device branch compiled for the host, tests/devpred/host_devbranch.cpp):
against the grid predicates of the pinned oracle, against an independent
rational circumcircle computation, against closed-form crossings, and the
brute-force CDT against the oracle's CDT where every crossing is a grid
point; the validator rejects planted mistakes."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from oracle import steiner as OS
from tests import steiner_cases as SC
from tests.devpred import devpred as D

R = (1 << 30) - 1


def test_homogeneous_predicates_equal_the_grid_ones():
    rng = np.random.default_rng(1)
    for _ in range(300):
        q = [tuple(int(v) for v in p) for p in rng.integers(0, R + 1, size=(4, 2))]
        h = [OS.grid_h(p) for p in q]
        k = int(rng.integers(2, 1 << 20))
        hs = [(x * k, y * k, k) for x, y, _ in h]  # the same points, scaled
        assert OS.orient(*h[:3]) == O.orient2d(*q[:3]) == OS.orient(*hs[:3])
        assert OS.incircle(*h) == O.incircle(*q) == OS.incircle(*hs)


def _incircle_fraction(a, b, c, d):
    """Circumcentre of a, b, c by solving the perpendicular-bisector system in
    rationals; the sign of r^2 - |d - centre|^2 (independent of the
    determinant form)."""
    A, B, C, Dp = ([Fraction(p[0], p[2]), Fraction(p[1], p[2])] for p in (a, b, c, d))
    # 2 (B - A) . X = |B|^2 - |A|^2 ; 2 (C - A) . X = |C|^2 - |A|^2
    a11, a12 = 2 * (B[0] - A[0]), 2 * (B[1] - A[1])
    a21, a22 = 2 * (C[0] - A[0]), 2 * (C[1] - A[1])
    r1 = B[0] ** 2 + B[1] ** 2 - A[0] ** 2 - A[1] ** 2
    r2 = C[0] ** 2 + C[1] ** 2 - A[0] ** 2 - A[1] ** 2
    det = a11 * a22 - a12 * a21
    cx, cy = (r1 * a22 - a12 * r2) / det, (a11 * r2 - a21 * r1) / det
    r = (A[0] - cx) ** 2 + (A[1] - cy) ** 2
    e = (Dp[0] - cx) ** 2 + (Dp[1] - cy) ** 2
    return (r > e) - (r < e)


def test_homogeneous_incircle_against_a_rational_circumcircle():
    rng = np.random.default_rng(2)
    pts = SC.rand_crossings(rng, 160)
    for i in range(0, 160, 4):
        a, b, c, d = pts[i:i + 4]
        if OS.orient(a, b, c) == 0:
            continue
        if OS.orient(a, b, c) < 0:
            b, c = c, b
        assert OS.incircle(a, b, c, d) == _incircle_fraction(a, b, c, d)


def test_device_branch_homogeneous_predicates():
    """exact.hpp's orient_h / incircle_h, device branch (the 64x64 product
    from __umul64hi) on the host: equal to the Python determinants on
    crossing points, grid points and exact degeneracies."""
    triples, quads, want_o, want_i = SC.homogeneous_cases(3)
    assert 0 in want_o and 0 in want_i and 1 in want_i and -1 in want_i
    assert list(D.orient_h(triples, on_host=True)) == want_o
    assert list(D.incircle_h(quads, on_host=True)) == want_i


def test_device_branch_crossing_points():
    segs = SC.crossing_segments(4)
    SC.check_crossing_points(segs, D.cross(segs, on_host=True))


def test_crossings_closed_forms():
    # the diagonals of an axis-parallel rectangle meet at its centre
    g = np.array([[100, 200], [900, 600], [100, 600], [900, 200], [50, 50]], np.int32)
    assert OS.crossings(g, OS.pieces(g, [(0, 1), (2, 3)])) == {(500, 400, 1)}
    # a non-grid crossing: y = 1 (0..3) and x = y (0..3) * 1/2 slope -> (2, 1)
    g = np.array([[0, 1], [3, 1], [0, 0], [4, 2]], np.int32)
    assert OS.crossings(g, OS.pieces(g, [(0, 1), (2, 3)])) == {(2, 1, 1)}
    g = np.array([[0, 0], [3, 2], [0, 2], [3, 0]], np.int32)
    assert OS.crossings(g, OS.pieces(g, [(0, 1), (2, 3)])) == {(3, 2, 2)}   # (1.5, 1)
    # every horizontal crosses every diagonal at (c_i + d_j, c_i)
    g, c = SC.hd_family(5, 3, 4, 0)
    cr = OS.crossings(g, OS.pieces(g, c))
    ys = sorted({int(g[a][1]) for a, _ in c[:3]})
    ds = sorted({int(g[a][0]) - int(g[a][1]) for a, _ in c[3:]})
    assert cr == {(y + d, y, 1) for y in ys for d in ds}
    # segments through one point meet once there
    g, c = SC.star(6, arms=4, extra=0)
    assert OS.crossings(g, OS.pieces(g, c)) == {(1 << 29, (1 << 29) + 12345, 1)}
    # shared endpoints and touching are not crossings
    g = np.array([[0, 0], [10, 10], [10, 0], [5, 5 + 3]], np.int32)
    assert OS.crossings(g, OS.pieces(g, [(0, 1), (0, 2), (1, 2)])) == set()


def test_brute_force_without_crossings_equals_the_oracle_cdt():
    rng = np.random.default_rng(7)
    for _ in range(6):
        g = rng.integers(1, R, size=(9, 2)).astype(np.int32)
        cons = np.array([[0, 1], [2, 3]], np.int32)
        P = [OS.grid_h(p) for p in g]
        if OS.crossings(g, OS.pieces(g, cons)):
            continue
        bf, cr = OS.brute_force(g, cons)
        assert cr == []
        ref, _ = O.cdt(g, cons)
        assert np.array_equal(bf, O.canonical(ref))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_brute_force_equals_the_cdt_of_the_augmented_input(seed):
    """Grid crossings: the brute-force CDT with crossing vertices equals
    oracle.cdt of the points plus the crossings, segments cut at them."""
    g, c = SC.hd_family(seed, 2, 2, 2)
    bf, cr = OS.brute_force(g, c)
    assert len(cr) == 4
    ref, v = SC.augmented_cdt(g, c)
    assert v["valid"] and v["unique"]
    assert np.array_equal(bf, ref)


def _brute_result(seed):
    g, c = SC.random_crossing(seed, 7, 4)
    bf, cr = OS.brute_force(g, c)
    return g, c, bf, cr


def test_validator_accepts_brute_force_and_rejects_planted_mistakes():
    for seed in range(20):
        g, c, bf, cr = _brute_result(seed)
        if len(cr) >= 2:
            break
    assert len(cr) >= 2
    v = OS.validate(g, c, bf, cr)
    assert v["valid"] and v["unique"]
    # a crossing moved by one unit of X
    bad = list(cr)
    bad[0] = (bad[0][0] + 1, bad[0][1], bad[0][2])
    assert not OS.validate(g, c, bf, bad)["crossings_ok"]
    # a crossing dropped (ids out of range)
    assert not OS.validate(g, c, bf, cr[:-1])["valid"]
    # one triangle clockwise
    t = bf.copy()
    t[0] = t[0][::-1]
    assert not OS.validate(g, c, t, cr)["valid"]
    # one triangle missing
    assert not OS.validate(g, c, bf[1:], cr)["valid"]
    # an interior edge flipped: not Delaunay, or a sub-piece missing
    n = len(g)
    T = [tuple(int(x) for x in r) for r in bf]
    directed = {}
    for i, (a, b, cc) in enumerate(T):
        for u, w, o in ((a, b, cc), (b, cc, a), (cc, a, b)):
            directed[(u, w)] = (i, o)
    flipped = 0
    V = [OS.grid_h(p) for p in O.with_corners(g)] + list(cr)
    for (u, w), (i, o) in directed.items():
        if (w, u) not in directed:
            continue
        j, q = directed[(w, u)]
        # convex quad: o and q on opposite sides of u-w and u, w on opposite sides of o-q
        if OS.orient(V[o], V[q], V[u]) * OS.orient(V[o], V[q], V[w]) < 0:
            t = [list(r) for r in T]
            t[i] = [o, u, q]
            t[j] = [q, w, o]
            r = OS.validate(g, c, np.array(t), cr)
            assert not r["valid"] and (r["not_delaunay"] > 0 or r["missing_constraints"] > 0)
            flipped += 1
    assert flipped > 0
