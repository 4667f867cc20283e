"""The same inputs through each round path of the solve: every round as
separate launches (tail threshold 0), every round inside the persistent
workgroup (threshold 2^20: short rounds keep their flip claims in LDS, long
ones on the records), and the tuned defaults.  Each result must pass the exact
validator and match the oracle's canonical CDT when it is unique."""
import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = {
    "uniform": lambda: gen.uniform_cdt(n=30_000, c=300),
    "clustered": lambda: gen.clustered_cdt(n=60_000, circles=40),
    "lattice": lambda: gen.lattice_cdt(k=96, c=80, max_len=40),
}


@pytest.fixture(scope="module")
def references():
    out = {}
    for name, make in CASES.items():
        p, c = make()
        g = gen.to_grid(p)
        ref, _ = O.cdt(g, c)
        out[name] = (g, c, ref, O.validate(g, ref, c)["unique"])
    return out


@pytest.mark.parametrize("threshold", [0, 1 << 20])
@pytest.mark.parametrize("case", sorted(CASES))
def test_round_paths_agree(references, case, threshold):
    from gcdt_amd import get_tunable, set_tunable, solve
    g, c, ref, unique = references[case]
    old = get_tunable("tail_threshold")
    set_tunable("tail_threshold", threshold)
    try:
        tris, st = solve(g, c)
    finally:
        set_tunable("tail_threshold", old)
    v = O.validate(g, tris, c)
    assert v["valid"], (case, threshold, v)
    if unique:
        assert np.array_equal(O.canonical(tris), O.canonical(ref))
    assert st["num_triangles"] == 2 * len(g) + 2


def test_batch_all_in_tail():
    """Batched instances with every round in the persistent workgroup."""
    from gcdt_amd import get_tunable, set_tunable, solve_batch
    s = gen.Stream(7)
    inst = [gen.polygon_instance(s, int(n)) for n in (64, 300, 2000, 9000)]
    inst = [(gen.to_grid(p), c) for p, c in inst]
    old = get_tunable("tail_threshold")
    set_tunable("tail_threshold", 1 << 20)
    try:
        res, _ = solve_batch(inst)
    finally:
        set_tunable("tail_threshold", old)
    for (g, c), tris in zip(inst, res):
        assert O.validate(g, tris, c)["valid"]
