"""CPU tests of the serving loop's policy: continuous batching, HBM-footprint
admission, back-pressure and latency accounting.  The executor is a CPU
stand-in (the oracle, or a timed sleep), so no GPU is needed; the GPU run of
the same server is in test_gpu_serve.py."""
import threading
import time

import numpy as np
import pytest

from gcdt_amd import gen
from gcdt_amd import serve as S
from oracle import oracle as O


def _req(rid, n, t=0.0):
    return S.Request(rid, t, np.zeros((n, 2), np.int32), np.zeros((0, 2), np.int32))


def test_take_batch_merges_small_requests_in_arrival_order():
    srv = S.Server(executor_factory=None, small_n=100, batch_points=250)
    for i, n in enumerate([50, 60, 70, 80, 500, 10]):
        srv.waiting.append(_req(i, n, 0.001 * i))
    b, need = srv.take_batch()
    # 180 points; 80 would pass 250 and 500 is large, but 10 still fits
    assert [r.rid for r in b] == [0, 1, 2, 5]
    assert need == S.batch_footprint([50, 60, 70, 10])
    b, _ = srv.take_batch()
    assert [r.rid for r in b] == [3]                # the large request behind it waits its turn
    b, need = srv.take_batch()
    assert [r.rid for r in b] == [4] and need == S.footprint(500)
    assert srv.take_batch() == (None, 0)


def test_footprints_cover_the_measured_pool_peaks():
    """The per-path figures are at least the MemCache pool's measured peaks
    (bench.py's pool_peak_bytes, inputs and outputs included): the single
    solve of the 10M workload (10,099,763 points, 100,000 segments) peaked at
    9,674,490,880 bytes in use, the batch of 4,096 instances (38,759,094
    points) at 17,474,913,280."""
    assert S.footprint(10_099_763) >= 9_674_490_880
    from gcdt_amd import gen
    ns = [len(p) for p, _ in gen.batched_instances(count=4096)]
    assert sum(ns) == 38_759_094
    assert S.batch_footprint(ns) >= 17_474_913_280


def test_results_hold_their_bytes_until_fetched():
    reqs = [_req(i, 1000 + i, 0.0) for i in range(4)]
    srv = S.Server(executor_factory=SleepExecutor, workers=2, small_n=10)
    srv.run(reqs, realtime=False)
    held = sum(S.result_bytes(r.n) for r in reqs)
    assert srv.ledger.used == held
    for r in reqs:
        r.fetch()
    assert srv.ledger.used == 0
    reqs[0].fetch()                                 # a second fetch releases nothing more
    assert srv.ledger.used == 0


def test_waiting_past_the_deadline_fails_with_a_timeout():
    reqs = [_req(i, 50_000, 0.0) for i in range(3)]
    reqs[2].deadline = 0.0005                       # needs admission within 0.5 ms
    srv = S.Server(executor_factory=SleepExecutor, workers=1, small_n=10)
    st = srv.run(reqs, realtime=False)
    assert st.completed == 2 and st.failed == 1 and st.timed_out == 1
    assert reqs[2].error.startswith("timeout") and reqs[2].t_admit < 0


def test_admission_holds_back_when_budget_is_full():
    srv = S.Server(executor_factory=None, small_n=10, hbm_budget=S.footprint(1000) + 10)
    srv.waiting.append(_req(0, 1000))
    srv.waiting.append(_req(1, 1000))
    b, need = srv.take_batch()
    srv.ledger.reserve(need)
    assert srv.take_batch() == (None, 0)            # second one waits for the first
    assert len(srv.waiting) == 1
    srv.ledger.release(need)
    b2, _ = srv.take_batch()
    assert b2[0].rid == 1


def test_oversize_request_runs_alone_on_an_idle_device():
    srv = S.Server(executor_factory=None, small_n=10, hbm_budget=1 << 20)
    srv.waiting.append(_req(0, 10_000_000))
    b, need = srv.take_batch()
    assert b and need > srv.ledger.budget           # admitted only because nothing runs


class SleepExecutor:
    """Simulated device: time proportional to points; tracks reserved bytes."""
    live = 0
    peak_live = 0
    lock = threading.Lock()

    def __init__(self):
        pass

    def run(self, batch):
        with SleepExecutor.lock:
            SleepExecutor.live += 1
            SleepExecutor.peak_live = max(SleepExecutor.peak_live, SleepExecutor.live)
        time.sleep(0.002 + 1e-8 * sum(r.n for r in batch))
        for r in batch:
            r.result = np.full((2 * r.n + 2, 3), r.rid, np.int32)
        with SleepExecutor.lock:
            SleepExecutor.live -= 1


def test_server_completes_every_request_within_budget():
    rng = np.random.default_rng(3)
    reqs = [_req(i, int(n), t) for i, (n, t) in
            enumerate(zip(rng.integers(10, 200_000, 80), np.cumsum(rng.exponential(0.002, 80))))]
    budget = S.footprint(400_000) + sum(S.result_bytes(r.n) for r in reqs)
    srv = S.Server(executor_factory=SleepExecutor, workers=3, hbm_budget=budget, small_n=50_000,
                   batch_points=150_000)
    st = srv.run(reqs)
    assert st.completed == len(reqs) and st.failed == 0
    assert st.peak_reserved <= budget
    assert SleepExecutor.peak_live <= 3
    for r in reqs:
        assert r.t_done >= r.t_admit >= r.arrival - 1e-6
        assert np.all(r.fetch() == r.rid)           # each request got its own result
    s = st.summary(qps=500)
    assert s["completed"] == 80 and s["p99_ms"] >= s["p50_ms"] > 0
    assert sum(st.batch_sizes) == 80 and max(st.batch_sizes) > 1


class FailingExecutor:
    def run(self, batch):
        if any(r.rid == 2 for r in batch):
            raise RuntimeError("device error")
        for r in batch:
            r.result = np.zeros((2 * r.n + 2, 3), np.int32)


def test_a_failing_batch_does_not_stop_the_server():
    reqs = [_req(i, 100 + i, 0.0) for i in range(6)]
    srv = S.Server(executor_factory=FailingExecutor, workers=1, small_n=10)   # every request alone
    st = srv.run(reqs, realtime=False)
    assert st.failed == 1 and st.completed == 5
    assert reqs[2].error == "device error" and reqs[3].error is None


class PerRequestExecutor:
    """Like the GPU executor with per-instance status: a bad request inside a
    merged batch fails alone."""

    def run(self, batch):
        errors = []
        for r in batch:
            if r.rid == 3:
                errors.append("gcdt_solve_batch: crossing constraints (status 5)")
            else:
                r.result = np.full((2 * r.n + 2, 3), r.rid, np.int32)
                errors.append(None)
        return errors


def test_a_bad_request_in_a_merged_batch_fails_alone():
    reqs = [_req(i, 100 + i, 0.0) for i in range(8)]
    srv = S.Server(executor_factory=PerRequestExecutor, workers=1, small_n=1000)
    st = srv.run(reqs, realtime=False)
    assert max(st.batch_sizes) == 8                 # all of them in one batch
    assert st.failed == 1 and st.completed == 7
    assert "crossing" in reqs[3].error
    for r in reqs:
        if r.rid != 3:
            assert r.error is None and np.all(r.fetch() == r.rid)


class OracleExecutor:
    """Reference CDT per request (CPU): checks that batching routes each
    result to its own request."""

    def run(self, batch):
        for r in batch:
            r.result, _ = O.cdt(r.pts, r.cons)


def test_results_are_routed_to_their_requests():
    reqs = []
    for i, (t, p, c) in enumerate(gen.serving_requests(count=10, qps=1000.0, nmin=50, nmax=600, seed=11)):
        reqs.append(S.Request(i, t, gen.to_grid(p), c))
    srv = S.Server(executor_factory=OracleExecutor, workers=2, small_n=400)
    st = srv.run(reqs)
    assert st.completed == 10
    for r in reqs:
        v = O.validate(r.pts, r.fetch(), r.cons)
        assert v["valid"], v


def test_serving_requests_are_deterministic():
    a = gen.serving_requests(count=3, qps=10.0, nmin=100, nmax=300, seed=5)
    b = gen.serving_requests(count=3, qps=10.0, nmin=100, nmax=300, seed=5)
    for (ta, pa, ca), (tb, pb, cb) in zip(a, b):
        assert ta == tb and np.array_equal(pa, pb) and np.array_equal(ca, cb)
    assert all(x[0] < y[0] for x, y in zip(a, a[1:]))


@pytest.mark.parametrize("n", [1000, 1_000_000])
def test_footprint_is_monotone(n):
    assert S.footprint(n) < S.footprint(2 * n)
