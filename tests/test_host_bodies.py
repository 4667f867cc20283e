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
