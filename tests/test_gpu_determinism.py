"""Run-to-run determinism on the GPU: the same input gives the same triangle
and adjacency arrays, bit for bit -- also on degenerate inputs, whose CDT is
not unique, and for batches.  Each Morton cell's points are reordered by
input index, samples are each cell's extreme point, each cell's leaves are
reordered by vertex ids before every flip phase, on-edge points rank by their
edge, and batched outputs are placed by slot, so no atomic's arrival order
reaches a decision."""
import numpy as np
import pytest

from gcdt_amd import gen, solve, solve_batch
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _twice(g, c):
    t1, _, a1 = solve(g, c, with_adj=True)
    t2, _, a2 = solve(g, c, with_adj=True)
    return np.array_equal(t1, t2) and np.array_equal(a1, a2)


def test_lattice_full_size_is_deterministic():
    p, c = gen.lattice_cdt()
    assert _twice(gen.to_grid(p), c)


def test_clustered_full_size_is_deterministic():
    """The 10M clustered workload with its 100,000 segments: the same arrays
    and the same round counts (the flip rounds no longer vary run to run)."""
    p, c = gen.clustered_cdt()
    g = gen.to_grid(p)
    t1, s1, a1 = solve(g, c, with_adj=True)
    t2, s2, a2 = solve(g, c, with_adj=True)
    assert np.array_equal(t1, t2) and np.array_equal(a1, a2)
    for k in ("insert_rounds", "flip_rounds", "constraint_rounds", "flips", "segments"):
        assert s1[k] == s2[k], k


def test_batch_is_deterministic():
    inst = [(gen.to_grid(p), c) for p, c in gen.batched_instances(count=256)]
    a, _ = solve_batch(inst)
    b, _ = solve_batch(inst)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def _dense_cell(seed, n_uniform, n_dense):
    """n_uniform grid points over the square plus n_dense distinct points in
    one 2^20 x 2^20 cell of the 2^30 grid (the cell of the 20-bit Morton
    sort), so that cell and the leaf cells inside it hold far more elements
    than one workgroup sorts."""
    rng = np.random.default_rng(seed)
    u = rng.integers(1, (1 << 30) - 1, size=(n_uniform, 2))
    lo = 3 << 28  # a multiple of 2^20: the cell [lo, lo + 2^20)^2
    r = rng.choice(1 << 40, size=n_dense, replace=False)
    dense = lo + np.stack([r >> 20, r & ((1 << 20) - 1)], axis=1)
    g = np.unique(np.concatenate([u, dense]), axis=0)
    return rng.permutation(g).astype(np.int32)


def test_cells_beyond_one_workgroup_are_deterministic():
    """40,000 points in one 20-bit cell: its points and its leaves take the
    large-cell merge sort; two solves give the same arrays, and the result
    is the oracle's CDT."""
    g = _dense_cell(11, 200_000, 40_000)
    cell = (g[:, 0].astype(np.int64) >> 20) * 1024 + (g[:, 1].astype(np.int64) >> 20)
    assert np.bincount(cell).max() > 2048 * 10
    c = np.zeros((0, 2), np.int32)
    t1, _, a1 = solve(g, c, with_adj=True)
    t2, _, a2 = solve(g, c, with_adj=True)
    assert np.array_equal(t1, t2) and np.array_equal(a1, a2)
    ref, _ = O.cdt(g, c)
    v = O.validate(g, ref, c)
    if v["unique"]:
        assert np.array_equal(O.canonical(t1), O.canonical(ref))
    else:
        assert O.validate(g, t1, c)["valid"]


def test_crossing_constraints_are_deterministic():
    """gcdt_solve_steiner: the same triangles, crossing vertices (numbered in
    piece order) and named pieces on every run."""
    from gcdt_amd.api import solve_steiner
    from tests import steiner_cases as SC
    g, c = SC.random_crossing(10, 20_000, 150)
    t1, h1, s1, f1, _ = solve_steiner(g, c, 20_000)
    t2, h2, s2, f2, _ = solve_steiner(g, c, 20_000)
    assert len(h1) > 2000
    assert np.array_equal(t1, t2) and h1 == h2 and np.array_equal(s1, s2) and np.array_equal(f1, f2)
