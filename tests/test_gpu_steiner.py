"""gcdt_solve_steiner on the GPU: crossing constraints resolved with exact
crossing vertices.  Each result passes the exact validator of
oracle/steiner.py (the crossing set, CCW, manifold, area, every sub-piece an
edge, every other edge locally Delaunay); tiny inputs equal the brute-force
CDT, inputs whose crossings are grid points equal the oracle's CDT of the
points plus the crossings."""
import numpy as np
import pytest

from gcdt_amd import GcdtError, solve
from gcdt_amd.api import solve_steiner
from oracle import oracle as O
from tests import steiner_cases as SC

pytestmark = pytest.mark.gpu


def _run(g, c, max_steiner=None, **kw):
    tris, hp, segs, cflag, st = solve_steiner(g, c, max_steiner)
    v = SC.check(g, c, tris, hp, segs=segs, **kw)
    assert st["steiner"] == len(hp) and len(tris) == 2 * (len(g) + len(hp)) + 2
    return tris, hp, cflag, v


@pytest.mark.parametrize("seed", range(1, 9))
def test_tiny_inputs_equal_the_brute_force_cdt(seed):
    g, c = SC.random_crossing(seed, 12, 6)
    _run(g, c, brute=True)


def test_segments_through_one_point():
    g, c = SC.star(6, arms=4, extra=5)
    _, hp, _, _ = _run(g, c, brute=True)
    assert len(hp) == 1


@pytest.mark.parametrize("seed,nh,nd,extra", [(8, 5, 6, 40), (9, 20, 20, 5000)])
def test_grid_crossings_equal_the_cdt_of_the_augmented_input(seed, nh, nd, extra):
    g, c = SC.hd_family(seed, nh, nd, extra)
    _, hp, _, v = _run(g, c, max_steiner=nh * nd + 16, augmented=True)
    assert len(hp) == nh * nd


def test_thousands_of_crossings():
    g, c = SC.random_crossing(10, 20_000, 150)
    _, hp, cflag, v = _run(g, c, max_steiner=20_000)
    assert len(hp) > 2000
    # constrained bits: two per crossing vertex at least (its two pieces)
    assert int(np.sum(cflag != 0)) > 2 * len(hp)


def test_without_crossings_it_equals_the_plain_solve():
    from gcdt_amd import gen
    p, c = gen.uniform_cdt(n=5000, c=60, seed_pts=3, seed_cons=4)
    g = gen.to_grid(p)
    tris, hp, _, _, _ = solve_steiner(g, c)
    assert hp == []
    ref, _ = solve(g, c)
    assert np.array_equal(O.canonical(tris), O.canonical(ref))


def test_crossings_beyond_max_steiner_fail_with_capacity():
    g, c = SC.random_crossing(9, 400, 40)
    with pytest.raises(GcdtError) as e:
        solve_steiner(g, c, max_steiner=3)
    assert e.value.status == 8


def test_plain_solve_still_refuses_crossings():
    g, c = SC.random_crossing(9, 400, 40)
    with pytest.raises(GcdtError) as e:
        solve(g, c)
    assert e.value.status == 5
