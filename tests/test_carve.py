"""Even-odd carving restated in the oracle, pinned by exact areas: the
triangles it keeps cover exactly the polygons' area minus their holes."""
import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O


def _expected_doubled_area(g, cons, nested):
    areas = sorted((O.loop_doubled_area(g, cyc) for cyc in O.loops(cons)), reverse=True)
    if nested:  # one outer polygon, every other loop a hole inside it
        return areas[0] - sum(areas[1:])
    return sum(areas)


@pytest.mark.parametrize("kind,seed,n", [("star", 3, 600), ("star", 8, 1500), ("poly", 5, 800)])
def test_oracle_carving_matches_polygon_area(kind, seed, n):
    s = gen.Stream(seed)
    p, c = gen.star_instance(s, n) if kind == "star" else gen.polygon_instance(s, n)
    g = gen.to_grid(p)
    tris, _ = O.cdt(g, c)
    bits = O.constraint_bits(g, tris, c)
    inside = O.carve_even_odd(tris, bits, len(g))
    assert inside.any() and not inside.all()
    assert O.doubled_area(g, tris[inside]) == _expected_doubled_area(g, c, kind == "star")
    # every constraint piece is an edge: the bits mark at least one side of each
    assert int(sum(bin(int(b)).count("1") for b in bits)) >= len(c)


def test_planted_mistake_without_constraint_bits_keeps_nothing():
    s = gen.Stream(3)
    p, c = gen.star_instance(s, 600)
    g = gen.to_grid(p)
    tris, _ = O.cdt(g, c)
    inside = O.carve_even_odd(tris, np.zeros(len(tris), np.int32), len(g))
    assert not inside.any()


def test_loops_recover_closed_polygons():
    s = gen.Stream(4)
    p, c = gen.star_instance(s, 900)
    cyc = O.loops(c)
    assert sum(len(x) for x in cyc) == len(c)
