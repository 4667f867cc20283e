"""GPU even-odd carving (gcdt_solve_flags + gcdt_carve) against the oracle's
restatement on the GPU's own triangulation, and the exact polygon area."""
import numpy as np
import pytest

from gcdt_amd import gen, solve_carved
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _expected_doubled_area(g, cons, nested):
    areas = sorted((O.loop_doubled_area(g, cyc) for cyc in O.loops(cons)), reverse=True)
    return areas[0] - sum(areas[1:]) if nested else sum(areas)


@pytest.mark.parametrize("kind,seed,n", [("star", 3, 600), ("star", 11, 20_000), ("poly", 5, 5_000)])
def test_carving_matches_oracle_and_area(kind, seed, n):
    s = gen.Stream(seed)
    p, c = gen.star_instance(s, n) if kind == "star" else gen.polygon_instance(s, n)
    g = gen.to_grid(p)
    tris, cflag, inside, st = solve_carved(g, c)
    v = O.validate(g, tris, c)
    assert v["valid"], v
    assert np.array_equal(cflag, O.constraint_bits(g, tris, c))
    assert np.array_equal(inside, O.carve_even_odd(tris, cflag, len(g)))
    assert O.doubled_area(g, tris[inside]) == _expected_doubled_area(g, c, kind == "star")


def test_carving_without_constraints_keeps_nothing():
    s = gen.Stream(2)
    p, _ = gen.dedup(np.stack([s.bits(500, 20), s.bits(500, 20)], axis=1))
    tris, cflag, inside, _ = solve_carved(gen.to_grid(p), np.zeros((0, 2), np.int32))
    assert not cflag.any() and not inside.any()
