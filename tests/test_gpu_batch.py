"""GPU parity of the batched solve (gcdt_solve_batch): every instance of a
mixed batch must be the oracle's CDT, and for instances whose CDT is unique
the batched result must equal the instance solved alone."""
import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _instances(k, seed=1, nmax=4000):
    s = gen.Stream(seed)
    out = []
    for i in range(k):
        if i % 3 == 0:
            n = 64 + int(s.below(1, nmax)[0])
            p, c = gen.star_instance(s, n)
        elif i % 3 == 1:
            p, c = gen.polygon_instance(s, 500 + int(s.below(1, nmax)[0]))
        else:
            p, _ = gen.dedup(np.stack([s.bits(300, 20), s.bits(300, 20)], axis=1))
            c = np.zeros((0, 2), np.int32)
        out.append((gen.to_grid(p), c))
    return out


def test_batch_instances_are_exact():
    from gcdt_amd import solve_batch
    inst = _instances(30)
    res, st = solve_batch(inst)
    assert st["num_instances"] == len(inst)
    for (g, c), tris in zip(inst, res):
        assert len(tris) == 2 * len(g) + 2
        v = O.validate(g, tris, c)
        assert v["valid"], v
        ref, _ = O.cdt(g, c)
        if O.validate(g, ref, c)["unique"]:
            assert np.array_equal(O.canonical(tris), O.canonical(ref))


def test_batch_equals_alone():
    """Each result is the same alone and in the mix (unique CDTs compare bit-exactly)."""
    from gcdt_amd import solve, solve_batch
    inst = _instances(12, seed=7)
    res, _ = solve_batch(inst)
    for (g, c), tris in zip(inst, res):
        alone, _ = solve(g, c)
        ref, _ = O.cdt(g, c)
        if O.validate(g, ref, c)["unique"]:
            assert np.array_equal(O.canonical(tris), O.canonical(alone))
        else:
            assert O.validate(g, tris, c)["valid"] and O.validate(g, alone, c)["valid"]


def test_batch_tiny_and_empty_instances():
    from gcdt_amd import solve_batch
    z = np.zeros((0, 2), np.int32)
    inst = [(gen.to_grid(np.zeros((0, 2), np.int64)), z),
            (gen.to_grid(np.array([[5, 7]])), z),
            (gen.to_grid(np.array([[5, 7], [100, 3], [50, 900]])), np.array([[0, 2]], np.int32)),
            (gen.to_grid(gen.lattice_cdt(k=16, c=4, max_len=6)[0]), gen.lattice_cdt(k=16, c=4, max_len=6)[1])]
    res, _ = solve_batch(inst)
    for (g, c), tris in zip(inst, res):
        assert len(tris) == 2 * len(g) + 2
        assert O.validate(g, tris, c)["valid"]


def test_batch_rejects_cross_instance_constraint():
    from gcdt_amd import GcdtError, solve_batch
    a = gen.to_grid(np.array([[5, 7], [100, 3], [50, 900]]))
    bad = np.array([[0, 5]], np.int32)  # id 5 does not exist in a 3-point instance
    with pytest.raises(GcdtError):
        solve_batch([(a, bad), (a, np.zeros((0, 2), np.int32))])


def _crossing_instance():
    """Two constraints crossing at a point that is not a vertex."""
    rng = np.random.default_rng(1)
    p = np.concatenate([rng.integers(0, 1 << 20, size=(300, 2)),
                        [[1000, 1000], [900000, 900000], [1000, 900000], [900000, 1000]]])
    p, _ = gen.dedup(p)
    idx = {tuple(x): i for i, x in enumerate(p.tolist())}
    a, b, c, d = idx[(1000, 1000)], idx[(900000, 900000)], idx[(1000, 900000)], idx[(900000, 1000)]
    return gen.to_grid(p), np.array([[a, b], [c, d]], np.int32)


def test_batch_isolates_a_failing_instance():
    """With per-instance status, an instance whose constraints cross fails
    alone (GCDT_ERR_CROSSING) and every other instance is still solved."""
    from gcdt_amd import solve_batch
    inst = _instances(9, seed=5)
    inst.insert(4, _crossing_instance())
    res, st = solve_batch(inst, isolate=True)
    assert st["status"][4] == 5
    assert [s for i, s in enumerate(st["status"]) if i != 4] == [0] * (len(inst) - 1)
    for i, ((g, c), tris) in enumerate(zip(inst, res)):
        if i != 4:
            assert O.validate(g, tris, c)["valid"], i
    # without per-instance status the whole call fails
    from gcdt_amd import GcdtError
    with pytest.raises(GcdtError) as e:
        solve_batch(inst)
    assert e.value.status == 5
