"""The device mesh check (gcdt_check_mesh) on the GPU, and the solves with
the per-phase check switched on (the "validate" tunable)."""
import numpy as np
import pytest

from gcdt_amd import gen, set_tunable, solve, solve_batch
from gcdt_amd.api import check_mesh, solve_carved

pytestmark = pytest.mark.gpu


def test_solved_meshes_pass():
    p, c = gen.clustered_cdt(n=200_000, circles=20)
    g = gen.to_grid(p)
    tris, cflag, _, _ = solve_carved(g, c)
    _, _, adj = solve(g, c, with_adj=True)
    assert check_mesh(g, tris, adj, cflag) == (0, -1)
    # without the constraint bits the constrained edges may be non-Delaunay
    f, _ = check_mesh(g, tris, adj, None)
    assert f in (0, 16)


def test_planted_mistakes_are_caught():
    p, _ = gen.uniform_cdt(n=5000, c=0, seed_pts=3)
    g = gen.to_grid(p)
    tris, _, adj = solve(g, None, with_adj=True)
    assert check_mesh(g, tris, adj) == (0, -1)
    t2 = tris.copy()
    t2[5, [1, 2]] = t2[5, [2, 1]]
    f, first = check_mesh(g, t2, adj)
    assert f & 2 and f & 4 and first <= 5
    a2 = adj.copy()
    a2[9, 0] = len(tris) + 5
    assert check_mesh(g, tris, a2)[0] & 4
    t3 = tris.copy()
    t3[11, 0] = len(g) + 4 + 100
    assert check_mesh(g, t3, adj)[0] & 1


@pytest.fixture
def validating():
    set_tunable("validate", 1)
    try:
        yield
    finally:
        set_tunable("validate", 0)


def test_solves_pass_their_phase_checks(validating):
    p, c = gen.clustered_cdt(n=1_000_000, circles=100)
    solve(gen.to_grid(p), c)
    p, c = gen.lattice_cdt(k=512, c=200)
    solve(gen.to_grid(p), c)
    inst = [(gen.to_grid(p), c) for p, c in gen.batched_instances(count=64)]
    solve_batch(inst)
