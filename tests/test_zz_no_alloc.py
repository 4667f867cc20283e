"""gcdt_solve_noalloc (the solve whose buffers all come from the library's
static scratch array) against gcdt_solve, through the C ABI, on two calls in
a row -- the second with a smaller input, so it runs over the first call's
leftovers in the array -- and against the oracle: canonical triangle sets
equal where the CDT is unique, exact-valid always, adjacency symmetric."""
import ctypes

import numpy as np
import pytest

from gcdt_amd import DeviceBuffer, gen
from gcdt_amd.lib import Stats, check, lib
from oracle import dropin as D
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GCDT_ERR_SCRATCH = 6
MAX_POINTS = 750000  # GCDT_NOALLOC_MAX_POINTS


def _run(fn, g, c):
    n = len(g)
    dp = DeviceBuffer.from_array(g)
    dc = DeviceBuffer.from_array(np.ascontiguousarray(c, np.int32)) if len(c) else None
    dt, da = DeviceBuffer(12 * (2 * n + 2)), DeviceBuffer(12 * (2 * n + 2))
    try:
        nt = ctypes.c_int(0)
        st = Stats()
        check(fn(dp.ptr, n, dc.ptr if dc else None, len(c), dt.ptr, da.ptr, ctypes.byref(nt), ctypes.byref(st), None),
              fn.__name__)
        return dt.to_array(np.int32, (nt.value, 3)), da.to_array(np.int32, (nt.value, 3)), st.as_dict()
    finally:
        for b in (dp, dc, dt, da):
            if b is not None:
                b.free()


def _check_pair(g, c):
    L = lib()
    t_na, a_na, st = _run(L.gcdt_solve_noalloc, g, c)
    t_ref, _, _ = _run(L.gcdt_solve, g, c)
    assert t_na.shape == (2 * len(g) + 2, 3)
    v = O.validate(g, t_na, c)
    assert v["valid"], v
    assert D.adj_violations(O.with_corners(g), t_na, a_na) == []
    if v["unique"]:
        assert np.array_equal(O.canonical(t_na), O.canonical(t_ref))
        ref, _ = O.cdt(g, c)
        assert np.array_equal(O.canonical(t_na), O.canonical(ref))
    else:
        assert O.validate(g, t_ref, c)["valid"]
    return st


def test_noalloc_equals_solve_two_calls_second_smaller():
    check(lib().gcdt_pool_init(), "gcdt_pool_init")
    p1, c1 = gen.uniform_cdt(n=300_000, c=3_000, seed_pts=31, seed_cons=32)
    p2, c2 = gen.star_instance(gen.Stream(33), 20_000)
    _check_pair(gen.to_grid(p1), c1)
    _check_pair(gen.to_grid(p2), c2)
    # a degenerate input after them: the lattice with diagonals
    p3, c3 = gen.lattice_cdt(k=100, c=60, max_len=40)
    _check_pair(gen.to_grid(p3), c3)


def test_noalloc_rejects_inputs_above_its_maximum():
    check(lib().gcdt_pool_init(), "gcdt_pool_init")
    g = gen.to_grid(gen.uniform_cdt(n=1000, c=0)[0])
    d = DeviceBuffer.from_array(g)
    t = DeviceBuffer(12 * 2002)
    try:
        # the size check comes before any access to the buffers
        rc = lib().gcdt_solve_noalloc(d.ptr, MAX_POINTS + 1, None, 0, t.ptr, None, None, None, None)
        assert rc == GCDT_ERR_SCRATCH
        assert b"at most" in lib().gcdt_last_error()
        # the same buffers at their real size still solve
        rc = lib().gcdt_solve_noalloc(d.ptr, len(g), None, 0, t.ptr, None, None, None, None)
        assert rc == 0
    finally:
        d.free()
        t.free()


def test_noalloc_reports_a_solve_that_outgrows_the_array():
    """Long constraint segments crossing many triangles need more than the
    array holds (about 274 MB here): the call fails with GCDT_ERR_SCRATCH and
    the bytes it needed, and the next call on the same array succeeds."""
    L = lib()
    check(L.gcdt_pool_init(), "gcdt_pool_init")
    p, c = gen.uniform_cdt(n=100_000, c=20_000, seed_pts=35, seed_cons=36)
    with pytest.raises(Exception) as e:
        _run(L.gcdt_solve_noalloc, gen.to_grid(p), c)
    assert e.value.status == GCDT_ERR_SCRATCH and "too small" in str(e.value)
    p2, c2 = gen.uniform_cdt(n=20_000, c=100, seed_pts=37, seed_cons=38)
    _check_pair(gen.to_grid(p2), c2)
