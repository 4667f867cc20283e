"""The exact predicates evaluated by GPU code (tests/devpred/devpred.hip
includes the library's predicates.hpp) over the golden vectors, and the
signed 64x64 -> 128-bit product against exact Python integers: the device
branch of mul_s64 (__umul64hi plus the signed correction) is the one run."""
import numpy as np
import pytest

from tests import predicate_cases as PC
from tests import steiner_cases as SC
from tests.devpred import devpred as D

pytestmark = pytest.mark.gpu


def test_orient2d_golden_on_gpu():
    o, osign, _, _ = PC.golden()
    assert np.array_equal(D.orient2d(o), osign)


def test_incircle_golden_on_gpu():
    _, _, i, isign = PC.golden()
    assert np.array_equal(D.incircle(i), isign)


def test_mul_s64_exact_on_gpu():
    pairs = PC.product_pairs()
    lo, hi = D.mul_s64(pairs)
    elo, ehi = PC.exact_limbs(pairs)
    assert np.array_equal(lo, elo) and np.array_equal(hi, ehi)


def test_homogeneous_predicates_on_gpu():
    """exact.hpp's orient_h / incircle_h (crossing vertices of
    gcdt_solve_steiner) on the device: crossing points, grid points, exact
    collinear and cocircular cases."""
    triples, quads, want_o, want_i = SC.homogeneous_cases(3)
    assert list(D.orient_h(triples)) == want_o
    assert list(D.incircle_h(quads)) == want_i


def test_crossing_points_on_gpu():
    segs = SC.crossing_segments(4)
    SC.check_crossing_points(segs, D.cross(segs))
