"""Solves on six streams at once, three in a row on each: every host thread
polls its own context's counter snapshots while the other contexts run their
rounds (and contexts reuse one another's snapshot rings), and every result
must equal the same input solved alone (or be a valid CDT where the CDT is
not unique)."""
import ctypes
import threading

import numpy as np
import pytest

from gcdt_amd import DeviceBuffer, gen, lib, solve, solve_device
from gcdt_amd.lib import check
from oracle import oracle as O

pytestmark = pytest.mark.gpu

REPEATS = 3


def test_concurrent_solves_match_alone():
    L = lib()
    inputs = []
    for i in range(6):
        p, c = gen.uniform_cdt(n=40_000 + 10_000 * i, c=200, seed_pts=300 + i, seed_cons=400 + i)
        inputs.append((gen.to_grid(p), c))
    alone = [solve(g, c)[0] for g, c in inputs]
    results = [[] for _ in inputs]
    errors = []

    def work(i):
        g, c = inputs[i]
        T = 2 * len(g) + 2
        s = ctypes.c_void_p()
        try:
            check(L.gcdt_stream_create(ctypes.byref(s)), "gcdt_stream_create")
            with DeviceBuffer.from_array(g) as dp, DeviceBuffer.from_array(c) as dc, DeviceBuffer(12 * T) as dt:
                for _ in range(REPEATS):
                    solve_device(dp, len(g), dc, len(c), dt, None, stream=s)
                    check(L.gcdt_stream_synchronize(s), "gcdt_stream_synchronize")
                    results[i].append(dt.to_array(np.int32, (T, 3)))
        except Exception as e:  # reported below, with the failing input
            errors.append((i, repr(e)))
        finally:
            if s.value:
                L.gcdt_stream_destroy(s)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(inputs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a solve thread did not finish"
    assert not errors, errors
    for i, (g, c) in enumerate(inputs):
        assert len(results[i]) == REPEATS, i
        v = O.validate(g, alone[i], c)
        assert v["valid"], (i, v)
        for tris in results[i]:
            if v["unique"]:
                assert np.array_equal(O.canonical(tris), O.canonical(alone[i])), i
            else:
                assert O.validate(g, tris, c)["valid"], i
