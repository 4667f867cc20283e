"""The GPU code path of the exact predicates -- mul_s64 built from
__umul64hi with its signed correction -- compiled for the host (the
emulator's stand-in header makes __HIP_DEVICE_COMPILE__ true) and checked
against the golden vectors and exact Python products.  The same checks run
on the GPU itself in tests/test_gpu_predicates.py."""
import numpy as np
import pytest

from tests import predicate_cases as PC
from tests.devpred import devpred as D


@pytest.fixture(scope="module", autouse=True)
def _built():
    D.build_host()


def test_orient2d_golden():
    o, osign, _, _ = PC.golden()
    assert np.array_equal(D.orient2d(o, on_host=True), osign)


def test_incircle_golden():
    _, _, i, isign = PC.golden()
    assert np.array_equal(D.incircle(i, on_host=True), isign)


def test_mul_s64_exact():
    pairs = PC.product_pairs()
    lo, hi = D.mul_s64(pairs, on_host=True)
    elo, ehi = PC.exact_limbs(pairs)
    assert np.array_equal(lo, elo) and np.array_equal(hi, ehi)


def test_a_dropped_correction_is_caught():
    """The product check is sharp: without one signed correction term the
    high limb of a negative factor's product is off (restated here in
    Python on the same pairs)."""
    pairs = PC.product_pairs(count=200)
    elo, ehi = PC.exact_limbs(pairs)
    bad = 0
    for (a, b), h in zip(pairs, ehi.tolist()):
        ua, ub = a & ((1 << 64) - 1), b & ((1 << 64) - 1)
        hi = (ua * ub) >> 64
        if a < 0:
            hi -= ub  # the b < 0 correction is left out
        hi &= (1 << 64) - 1
        hi = hi - (1 << 64) if hi >= 1 << 63 else hi
        bad += hi != h
    assert bad > 0
