"""GPU parity at the workloads' full sizes, through the C ABI:

* the 10M-point clustered CDT with 1,000 circles (100,000 segments);
* the 4096 x 4096 lattice with 2,000 diagonal constraints (16.7M points);
* all 4,096 instances of the batched workload (38.7M points, one launch
  sequence);
* serving requests of up to 10^6 points, merged and alone.

Every result must pass the oracle's exact validator (tiling, constraint
chains, local Delaunay by the exact inCircle, triangle count), and must equal
the oracle's canonical triangle set wherever the CDT is unique.  The 10M
workload has cocircular unconstrained edges (the Gaussian clusters are dense
on the integer grid) and the lattice is degenerate throughout, so for those two
the exact validator is the bar."""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# the oracle releases the GIL: check instances on the host cores this run may
# use (16 per GPU), or fewer where the machine has fewer
WORKERS = max(1, min(16, os.cpu_count() or 1))


def _exact(g, tris, c):
    assert tris.shape == (2 * len(g) + 2, 3)
    v = O.validate(g, tris, c)
    assert v["valid"], v
    return v


def test_config_large_10M_clustered():
    from gcdt_amd import solve
    p, c = gen.clustered_cdt()
    g = gen.to_grid(p)
    assert len(g) > 10_000_000 and len(c) == 100_000
    tris, st = solve(g, c)
    v = _exact(g, tris, c)
    assert v["missing_constraints"] == 0 and v["not_delaunay"] == 0
    assert st["segments"] >= len(c)


def test_config_degenerate_lattice_4096():
    from gcdt_amd import solve
    p, c = gen.lattice_cdt()
    g = gen.to_grid(p)
    assert len(g) == 4096 * 4096 and len(c) == 2000
    tris, st = solve(g, c)
    v = _exact(g, tris, c)
    assert v["cocircular"] > 0  # degenerate: many Delaunay triangulations exist
    # each diagonal runs through lattice points, so it is split into unit pieces
    assert st["segments"] > 10 * len(c)


def _oracle_canonical(gc):
    g, c = gc
    ref, _ = O.cdt(g, c)
    v = O.validate(g, ref, c)
    return O.canonical(ref) if v["unique"] else None


def test_config_batched_all_4096_instances():
    from gcdt_amd import solve_batch
    inst = [(gen.to_grid(p), c) for p, c in gen.batched_instances()]
    assert len(inst) == 4096
    res, st = solve_batch(inst)
    assert st["num_instances"] == 4096

    def check(i):
        (g, c), tris = inst[i], res[i]
        if tris.shape != (2 * len(g) + 2, 3):
            return i, "count"
        v = O.validate(g, tris, c)
        if not v["valid"]:
            return i, "invalid"
        ref = _oracle_canonical((g, c))
        if ref is not None and not np.array_equal(O.canonical(tris), ref):
            return i, "differs from the oracle"
        return i, None

    with ThreadPoolExecutor(WORKERS) as ex:
        bad = [(i, why) for i, why in ex.map(check, range(len(inst))) if why]
    assert not bad, bad[:20]


def test_serving_requests_up_to_1e6_points():
    """Requests log-uniform in [10^3, 10^6] points (plus one of 10^6 exactly),
    served with merging, all waiting at once; each result must be the
    oracle's CDT, and the largest are also solved alone."""
    from gcdt_amd import serve as S
    from gcdt_amd import solve
    raw = gen.serving_requests(count=20, qps=100.0, nmin=1e3, nmax=1e6, seed=21)
    reqs = [S.Request(i, 0.0, gen.to_grid(p), c) for i, (_, p, c) in enumerate(raw)]
    big_p, big_c = gen.polygon_instance(gen.Stream(22), 1_000_000)
    reqs.append(S.Request(len(reqs), 0.0, gen.to_grid(big_p), big_c))
    assert max(r.n for r in reqs) >= 999_000
    st = S.Server(workers=3).run(reqs, realtime=False)
    assert st.completed == len(reqs) and st.failed == 0, [r.error for r in reqs if r.error]

    def check(r):
        tris = r.fetch()
        if tris.shape != (2 * r.n + 2, 3):
            return r.rid, "count"
        if not O.validate(r.pts, tris, r.cons)["valid"]:
            return r.rid, "invalid"
        ref = _oracle_canonical((r.pts, r.cons))
        if ref is not None and not np.array_equal(O.canonical(tris), ref):
            return r.rid, "differs from the oracle"
        return r.rid, None

    for r in reqs:
        r.fetch()  # device -> host on this thread
    with ThreadPoolExecutor(WORKERS) as ex:
        bad = [(rid, why) for rid, why in ex.map(check, reqs) if why]
    assert not bad, bad
    largest = sorted(reqs, key=lambda r: r.n)[-2:]
    for r in largest:
        alone, _ = solve(r.pts, r.cons)
        if O.validate(r.pts, alone, r.cons)["unique"]:
            assert np.array_equal(O.canonical(alone), O.canonical(r.fetch())), r.rid
