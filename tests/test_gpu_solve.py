"""GPU parity of the full CDT solve (gcdt_solve through the C ABI) against
the oracle: canonical triangle sets must be bit-identical when the CDT is
unique (no cocircular unconstrained edge), and every output must pass the
exact validator (tiling, constraint chains, local Delaunay)."""
import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _solve(g, c):
    from gcdt_amd import solve
    return solve(g, c)


def _check(g, c, expect_unique=None):
    tris, st = _solve(g, c)
    assert st["num_triangles"] == 2 * len(g) + 2
    v = O.validate(g, tris, c)
    assert v["valid"], v
    ref, _ = O.cdt(g, c)
    vr = O.validate(g, ref, c)
    if vr["unique"]:
        assert np.array_equal(O.canonical(tris), O.canonical(ref))
    if expect_unique is not None:
        assert vr["unique"] == expect_unique
    return st


@pytest.mark.parametrize("seed", range(8))
def test_random_small(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 2000))
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(n, 2)))
    _check(gen.to_grid(p), np.zeros((0, 2), np.int32))


def test_tiny_inputs():
    for n in (1, 2, 3, 4, 5):
        p = gen.dedup(np.array([[10 + 7 * i, 20 + 3 * i * i] for i in range(n)]))[0]
        _check(gen.to_grid(p), np.zeros((0, 2), np.int32))


def test_uniform_small_constraints():
    p, c = gen.uniform_cdt(n=20000, c=200)
    st = _check(gen.to_grid(p), c, expect_unique=True)
    assert st["segments"] > len(c)  # some segments pass through points and are split


def test_config_small():
    """100k uniform points + 1000 horizontal segments (the CPU-runnable config)."""
    p, c = gen.uniform_cdt()
    st = _check(gen.to_grid(p), c, expect_unique=True)
    assert st["splits"] > 0


def test_lattice_degenerate():
    p, c = gen.lattice_cdt(k=128, c=150, max_len=60)
    _check(gen.to_grid(p), c, expect_unique=False)


def test_clustered_circles():
    p, c = gen.clustered_cdt(n=200_000, circles=100)
    _check(gen.to_grid(p), c)


def test_star_instance():
    s = gen.Stream(11)
    p, c = gen.star_instance(s, 5000)
    _check(gen.to_grid(p), c)


def test_rejects_duplicate_point():
    from gcdt_amd import GcdtError
    g = gen.to_grid(np.array([[5, 5], [9, 9], [5, 5], [100, 7]]))
    with pytest.raises(GcdtError):
        _solve(g, np.zeros((0, 2), np.int32))


def test_rejects_out_of_range():
    from gcdt_amd import GcdtError
    g = np.array([[0, 5], [9, 9]], dtype=np.int32)
    with pytest.raises(GcdtError):
        _solve(g, np.zeros((0, 2), np.int32))


def test_rejects_crossing_constraints():
    """Two segments crossing at a non-vertex: the intersection is not a grid
    point, so the solve reports GCDT_ERR_CROSSING instead of inventing one."""
    from gcdt_amd import GcdtError
    rng = np.random.default_rng(1)
    p = np.concatenate([rng.integers(0, 1 << 20, size=(300, 2)),
                        [[1000, 1000], [900000, 900000], [1000, 900000], [900000, 1000]]])
    p, _ = gen.dedup(p)
    idx = {tuple(x): i for i, x in enumerate(p.tolist())}
    a, b, c, d = idx[(1000, 1000)], idx[(900000, 900000)], idx[(1000, 900000)], idx[(900000, 1000)]
    g = gen.to_grid(p)
    with pytest.raises(GcdtError) as e:
        _solve(g, np.array([[a, b], [c, d]], np.int32))
    assert e.value.status == 5
    tris, _ = _solve(g, np.array([[a, b]], np.int32))   # one of them alone is fine
    assert O.validate(g, tris, np.array([[a, b]]))["valid"]


def test_empty_input():
    tris, st = _solve(np.zeros((0, 2), np.int32), np.zeros((0, 2), np.int32))
    assert len(tris) == 2 and O.validate(np.zeros((0, 2), np.int32), tris)["valid"]


def test_collinear_points():
    x = np.arange(500, dtype=np.int64) * 1000 + 7
    p = np.stack([x, x // 2 + 11], 1)
    _check(gen.to_grid(p), np.zeros((0, 2), np.int32))


def test_extreme_grid_coordinates():
    """Points one unit inside the square (the largest inCircle terms) used
    directly on the 2^30 grid, with and without constraints."""
    R = (1 << 30) - 1
    ex = np.array([[1, 1], [R - 1, R - 1], [1, R - 1], [R - 1, 1], [1 << 29, 1 << 29], [1, 1 << 29],
                   [R - 1, 1 << 29], [1 << 29, 1], [1 << 29, R - 1], [2, 3], [R - 2, R - 3]], np.int32)
    rng = np.random.default_rng(9)
    ex = np.concatenate([ex, rng.integers(1, R - 1, size=(2000, 2)).astype(np.int32)])
    ex = np.unique(ex, axis=0)
    _check(ex, np.zeros((0, 2), np.int32))
    idx = {tuple(v): i for i, v in enumerate(ex.tolist())}
    cons = np.array([[idx[(1, 1)], idx[(1, R - 1)]], [idx[(1, R - 1)], idx[(R - 1, R - 1)]]], np.int32)
    _check(ex, cons)


def test_duplicate_and_reversed_constraints():
    rng = np.random.default_rng(3)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(3000, 2)))
    g = gen.to_grid(p)
    cons = np.array([[10, 20], [20, 10], [10, 20], [20, 30], [30, 10]], np.int32)
    tris, st = _solve(g, cons)
    assert O.validate(g, tris, cons)["valid"]


@pytest.mark.parametrize("bad", [[[5, 5]], [[0, 10 ** 6]], [[-1, 3]]])
def test_rejects_bad_constraint_ids(bad):
    from gcdt_amd import GcdtError
    rng = np.random.default_rng(4)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(200, 2)))
    with pytest.raises(GcdtError) as e:
        _solve(gen.to_grid(p), np.array(bad, np.int32))
    assert e.value.status == 2
