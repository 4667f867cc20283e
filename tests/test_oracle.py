"""Pins the oracle (CPU restatement) before anything is compared with it:
golden predicate vectors, brute-force DT/CDT on tiny inputs, closed-form
counts, and the exact validator's ability to catch planted mistakes."""
import json
import os

import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden", "predicates.json")


@pytest.fixture(scope="module")
def vectors():
    with open(GOLD) as f:
        return json.load(f)


def test_orient2d_golden(vectors):
    for v in vectors["orient2d"]:
        a, b, c = v["p"]
        assert O.orient2d(a, b, c) == v["sign"], v


def test_incircle_golden(vectors):
    bad = [v for v in vectors["incircle"] if O.incircle(*v["p"]) != v["sign"]]
    assert not bad, bad[:3]


def test_golden_has_degenerate_cases(vectors):
    assert sum(v["sign"] == 0 for v in vectors["incircle"]) > 300
    assert sum(v["sign"] == 0 for v in vectors["orient2d"]) > 300


@pytest.mark.parametrize("seed", range(12))
def test_dt_matches_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 26))
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(n, 2)))
    g = gen.to_grid(p)
    t, _ = O.cdt(g)
    assert np.array_equal(O.canonical(t), O.brute_force_dt(g))


def _random_constraints(rng, g, k):
    P = [tuple(map(int, q)) for q in g]
    cons = []
    for _ in range(k):
        a, b = (int(x) for x in rng.integers(0, len(P), 2))
        if a == b:
            continue
        ok = True
        for s, t in cons:
            if {a, b} == {s, t} or (len({a, b, s, t}) == 4 and O._segments_cross(P[a], P[b], P[s], P[t])):
                ok = False
        for d in range(len(P)):
            if d not in (a, b) and O._orient_py(P[a], P[b], P[d]) == 0 and \
                    min(P[a][0], P[b][0]) <= P[d][0] <= max(P[a][0], P[b][0]) and \
                    min(P[a][1], P[b][1]) <= P[d][1] <= max(P[a][1], P[b][1]):
                ok = False
        if ok:
            cons.append((a, b))
    return np.array(cons, dtype=np.int32).reshape(-1, 2)


@pytest.mark.parametrize("seed", range(12))
def test_cdt_matches_brute_force(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(4, 18))
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(n, 2)))
    g = gen.to_grid(p)
    c = _random_constraints(rng, g, 6)
    t, _ = O.cdt(g, c)
    assert np.array_equal(O.canonical(t), O.brute_force_cdt(g, c))


def test_counts_closed_form():
    p, c = gen.uniform_cdt(n=5000, c=50)
    g = gen.to_grid(p)
    t, st = O.cdt(g, c)
    n = len(p)
    assert len(t) == 2 * (n + 4) - 6
    v = O.validate(g, t, c)
    assert v["valid"] and v["edges_checked"] == 3 * (n + 4) - 7 - 4 - v["constraint_pieces"]
    assert st["cavity_flips"] == 0


def test_known_answer_four_points():
    # kernel.cu:49-52 fixed points (0.75,.5),(.75,.2),(.25,.2),(.5,.75) mapped by
    # kernel.cu:178-186 (x * 2^30, clamped) and chain (0,1),(1,2),(2,3) (kernel.cu:5348-5351)
    s = 1 << 30
    pts = np.array([[int(0.75 * s), int(0.5 * s)], [int(0.75 * s), int(0.2 * s)],
                    [int(0.25 * s), int(0.2 * s)], [int(0.5 * s), int(0.75 * s)]], dtype=np.int32)
    cons = np.array([[0, 1], [1, 2], [2, 3]], dtype=np.int32)
    t, _ = O.cdt(pts, cons)
    assert np.array_equal(O.canonical(t), O.brute_force_cdt(pts, cons))
    assert O.validate(pts, t, cons)["valid"]


def test_validator_catches_planted_mistakes():
    p, c = gen.uniform_cdt(n=2000, c=20)
    g = gen.to_grid(p)
    t, _ = O.cdt(g, c)
    assert O.validate(g, t, c)["valid"]
    # 1) swap two vertices of a triangle -> clockwise
    bad = t.copy(); bad[5, [1, 2]] = bad[5, [2, 1]]
    assert not O.validate(g, bad, c)["valid"]
    # 2) drop a triangle -> count/area/edges wrong
    assert not O.validate(g, t[1:], c)["valid"]
    # 3) flip one unconstrained interior edge -> not locally Delaunay
    from itertools import combinations
    key = {}
    for i, tri in enumerate(t):
        for k in range(3):
            key.setdefault(tuple(sorted((tri[(k + 1) % 3], tri[(k + 2) % 3]))), []).append((i, k))
    n = len(p)
    flipped = None
    for e, lst in key.items():
        if len(lst) != 2 or max(e) >= n:
            continue
        (i, k), (j, l) = lst
        pv, u, v = t[i][k], t[i][(k + 1) % 3], t[i][(k + 2) % 3]
        o = t[j][l]
        if O.orient2d(g[pv], g[u], g[o]) > 0 and O.orient2d(g[pv], g[o], g[v]) > 0:
            flipped = t.copy()
            flipped[i] = [pv, u, o]
            flipped[j] = [pv, o, v]
            break
    assert flipped is not None
    r = O.validate(g, flipped, c)
    assert not r["valid"] and (r["not_delaunay"] >= 1 or r["missing_constraints"] >= 1)
    # 4) the flipped-away edge declared as a constraint is reported missing
    e = (int(u), int(v))
    r = O.validate(g, flipped, np.concatenate([c, np.array([e], dtype=np.int32)]))
    assert r["missing_constraints"] >= 1


def test_lattice_is_valid_not_unique():
    p, c = gen.lattice_cdt(k=48, c=30, max_len=16)
    g = gen.to_grid(p)
    t, st = O.cdt(g, c)
    v = O.validate(g, t, c)
    assert v["valid"] and not v["unique"] and st["splits"] > 0


def test_generators_deterministic():
    a = gen.uniform_cdt(n=1000, c=10)
    b = gen.uniform_cdt(n=1000, c=10)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    p, c = gen.clustered_cdt(n=20000, circles=10)
    assert len(c) == 10 * 100
    s = gen.Stream(1)
    assert s.u64(1)[0] == np.uint64(0x910A2DEC89025CC1)  # SplitMix64(seed=1) first output
