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

pytestmark = pytest.mark.gpu


def _twice(g, c):
    t1, _, a1 = solve(g, c, with_adj=True)
    t2, _, a2 = solve(g, c, with_adj=True)
    return np.array_equal(t1, t2) and np.array_equal(a1, a2)


def test_lattice_full_size_is_deterministic():
    p, c = gen.lattice_cdt()
    assert _twice(gen.to_grid(p), c)


def test_clustered_is_deterministic():
    p, c = gen.clustered_cdt(n=1_000_000, circles=100)
    assert _twice(gen.to_grid(p), c)


def test_batch_is_deterministic():
    inst = [(gen.to_grid(p), c) for p, c in gen.batched_instances(count=256)]
    a, _ = solve_batch(inst)
    b, _ = solve_batch(inst)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
