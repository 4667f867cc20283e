"""Every value of every round tunable that the autotuner tries, on the GPU:
a mixed batch (star polygons with holes, polygon boundaries, plain point
sets, sizes from 64 to 12,000 points) and one single solve with constraints,
each result checked against the oracle (canonical equality where the CDT is
unique, the exact validator always).  The tunables change launch shapes and
list paths only (persistent tails, grid-stride caps, items per thread,
insertion pivots, one or two testers per edge), never the result's
correctness."""
import numpy as np
import pytest

from gcdt_amd import gen, tune
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(k, v) for k, vs in tune.CANDIDATES.items() for v in vs]
# the coordinate descent can end on several non-default values at once
COMBOS = [{"tail_threshold": 16384, "grid_cap": 16384, "flip_items": 4, "ins_items": 4, "pivot": 0, "single_tester": 0},
          {"tail_threshold": 1024, "grid_cap": 2048, "flip_items": 2, "ins_items": 2, "pivot": 0, "single_tester": 0}]


def _instances(seed):
    s = gen.Stream(seed)
    out = []
    for i, n in enumerate([64, 300, 2000, 900, 150, 5000, 70, 12000, 64, 800, 3000, 64]):
        if i % 3 == 0:
            p, c = gen.star_instance(s, n)
        elif i % 3 == 1:
            p, c = gen.polygon_instance(s, n)
        else:
            p, _ = gen.dedup(np.stack([s.bits(n, 20), s.bits(n, 20)], axis=1))
            c = np.zeros((0, 2), np.int32)
        out.append((gen.to_grid(p), c))
    return out


@pytest.fixture(scope="module")
def workload():
    inst = _instances(41)
    refs = []
    for g, c in inst:
        ref, _ = O.cdt(g, c)
        refs.append(O.canonical(ref) if O.validate(g, ref, c)["unique"] else None)
    p, c = gen.uniform_cdt(n=20_000, c=200, seed_pts=43, seed_cons=44)
    single = (gen.to_grid(p), c)
    ref, _ = O.cdt(*single)
    sref = O.canonical(ref) if O.validate(single[0], ref, single[1])["unique"] else None
    return inst, refs, single, sref


@pytest.fixture(autouse=True)
def _defaults():
    yield
    tune.apply(tune.DEFAULTS)


def _check(workload, choice):
    from gcdt_amd import solve, solve_batch
    inst, refs, single, sref = workload
    print(f"tunables {choice}", flush=True)
    tune.apply(choice)
    res, _ = solve_batch(inst)
    for i, ((g, c), tris, ref) in enumerate(zip(inst, res, refs)):
        assert tris.shape == (2 * len(g) + 2, 3), i
        assert O.validate(g, tris, c)["valid"], i
        if ref is not None:
            assert np.array_equal(O.canonical(tris), ref), i
    tris, _ = solve(*single)
    assert O.validate(single[0], tris, single[1])["valid"]
    if sref is not None:
        assert np.array_equal(O.canonical(tris), sref)


@pytest.mark.parametrize("knob,value", CASES)
def test_candidate_value(workload, knob, value):
    _check(workload, dict(tune.DEFAULTS, **{knob: value}))


@pytest.mark.parametrize("combo", range(len(COMBOS)))
def test_candidate_combination(workload, combo):
    _check(workload, COMBOS[combo])
