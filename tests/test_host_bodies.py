"""CPU parity of the solve's kernel bodies: the insertion, packed flip
(fused fix + test, stamp-pending slots), cavity and constrained flip bodies,
executed on the host one item at a time in seeded shuffled orders, must give
a valid CDT equal to the oracle's whenever the CDT is unique.  Several seeds
per input vary which of two neighbouring items is handled first."""
import numpy as np
import pytest

from gcdt_amd import gen
from oracle import oracle as O
from tests import host_bodies as HB


def _check(g, c, seed, fused=True, two_testers=False):
    rc, tris, st = HB.solve(g, c, seed=seed, fused=fused, two_testers=two_testers)
    assert rc == 0, (rc, seed)
    v = O.validate(g, tris, c)
    assert v["valid"], (seed, v)
    ref, _ = O.cdt(g, c)
    if O.validate(g, ref, c)["unique"]:
        assert np.array_equal(O.canonical(tris), O.canonical(ref)), seed
    return st


@pytest.mark.parametrize("seed", range(12))
def test_random_small(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 400))
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(n, 2)))
    _check(gen.to_grid(p), np.zeros((0, 2), np.int32), seed)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_uniform_with_constraints(seed):
    p, c = gen.uniform_cdt(n=12_000, c=120)
    _check(gen.to_grid(p), c, seed)


@pytest.mark.parametrize("fused", [True, False])
def test_lattice_degenerate(fused):
    p, c = gen.lattice_cdt(k=40, c=30, max_len=20)
    _check(gen.to_grid(p), c, 7, fused=fused)


def test_clustered_circles():
    p, c = gen.clustered_cdt(n=20_000, circles=20)
    _check(gen.to_grid(p), c, 3)


@pytest.mark.parametrize("seed", [1, 2])
def test_clamped_cluster_fans(seed):
    """A Gaussian cluster clamped at x = 0: a dense collinear column beside
    sparse points, whose vertex fans take hundreds of flip rounds."""
    rng = np.random.default_rng(seed)
    u = rng.integers(0, 1 << 20, size=(20_000, 2))
    cl = np.clip(np.rint(rng.normal(0, 4096, size=(6_000, 2)) + [800.0, 500_000.0]), 0, (1 << 20) - 1)
    p, _ = gen.dedup(np.concatenate([u, cl.astype(np.int64)]))
    st = _check(gen.to_grid(p), np.zeros((0, 2), np.int32), seed)
    assert st[2] > 50  # flip rounds: the fans do form


@pytest.mark.parametrize("seed", [4, 5])
def test_two_testers_per_edge(seed):
    """Both triangles of an edge test it in every round (the single_tester 0
    setting): more proposals lose, and every round must still list each
    triangle at most once (the emulator fails a case that lists one twice)."""
    p, c = gen.uniform_cdt(n=12_000, c=120, seed_pts=500 + seed, seed_cons=600 + seed)
    _check(gen.to_grid(p), c, seed, two_testers=True)
    _check(gen.to_grid(p), c, seed, fused=False, two_testers=True)


_TORN_CHILD = r"""
import sys
import numpy as np
from gcdt_amd import gen
from oracle import oracle as O
from tests import host_bodies as HB
HB.set_torn(True)
bad = []
for seed in range(int(sys.argv[1]), int(sys.argv[2])):
    rng = np.random.default_rng(seed)
    p, _ = gen.dedup(rng.integers(0, 1 << 20, size=(int(rng.integers(3, 12)), 2)))
    g = gen.to_grid(p)
    for s in range(2):
        rc, tris, st = HB.solve(g, np.zeros((0, 2), np.int32), seed=s, two_testers=True)
        if rc != 0 or not O.validate(g, tris)["valid"]:
            bad.append((seed, s, rc))
print("BAD", bad)
"""


def test_torn_record_reads_by_losing_proposals():
    """Both triangles of an edge test it, and a proposal whose triangle another
    proposal flips earlier in the same apply pass reads its own record torn:
    vertices, neighbours and coordinates already rewritten, meta and claim
    not yet -- as a GPU load of the record can while that flip's 16-byte
    stores land.  The new neighbour in the proposal's slot can be -1 (a
    square edge), and the records sit after an inaccessible page, so an apply
    that follows that neighbour before checking its own claim reads record
    -1 and the child faults.  Tiny inputs (3-11 points) make the square's
    edges common; seeds 57 and 69 hit the -1 case.  The unoptimised build
    makes every load the source names, as the GPU code does."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, HOST_BODIES_O0="1")
    r = subprocess.run([sys.executable, "-c", _TORN_CHILD, "0", "400"], cwd=root, env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "BAD []" in r.stdout, r.stdout
