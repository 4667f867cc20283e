"""Continuous-batching CDT server on one GPU.

Requests (points on the 2^30 grid + constraint vertex-id pairs) arrive over
time.  Worker threads, each owning one non-blocking HIP stream, repeatedly
take whatever requests are waiting when they become free:

* small requests (n <= small_n) are merged into one batched solve
  (gcdt_solve_batch) up to `batch_points` points;
* a large request is solved alone (gcdt_solve);

so requests that arrive while others are on the GPU join the next batch of
the first worker that frees up, and the workers' streams overlap on the
device.  Admission is by HBM footprint: a batch starts only when its estimated
device bytes fit in the budget next to the batches in flight (the ledger also
records the MemCache pool's own peak).  Results stay on the device until the
client reads them (Request.fetch()).

The executor is pluggable: GpuExecutor drives libgcdt.so; tests substitute a
CPU executor to exercise admission and back-pressure without a GPU.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

# bytes of device memory per point during a solve: the insertion DAG (4N
# slots x 28 B) alive while the leaves are written into the packed flip
# records (2N x 64 B), then records + final mesh (2N x 24 B) + flip lists, plus
# per-point state and the sorted copy; a 4,096-instance batched arena peaks
# at about 378 B per point, rounded up here; per request overhead for the
# bitmap and scratch
BYTES_PER_POINT = 384
BYTES_PER_REQUEST = 16 << 20


def footprint(n: int) -> int:
    return BYTES_PER_POINT * max(n, 1) + BYTES_PER_REQUEST


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the stream start
    pts: np.ndarray                # (n,2) int32 grid coordinates
    cons: np.ndarray               # (c,2) int32 local vertex ids
    t_admit: float = -1.0
    t_done: float = -1.0
    batch_size: int = 0
    error: str | None = None
    result: np.ndarray | None = None   # host copy once fetched
    _dev: object = None                # device triangle buffer (GPU executor)

    @property
    def n(self) -> int:
        return len(self.pts)

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    def fetch(self) -> np.ndarray:
        if self.result is None and self._dev is not None:
            self.result = self._dev.to_array(np.int32, (2 * self.n + 2, 3))
            self._dev.free()
            self._dev = None
        return self.result


class Ledger:
    """HBM-footprint admission: reserve before a batch starts, release after."""

    def __init__(self, budget: int):
        self.budget = int(budget)
        self.used = 0
        self.peak = 0
        self.cv = threading.Condition()

    def fits(self, nbytes: int) -> bool:
        return self.used + nbytes <= self.budget or self.used == 0

    def reserve(self, nbytes: int):
        self.used += nbytes
        self.peak = max(self.peak, self.used)

    def release(self, nbytes: int):
        with self.cv:
            self.used -= nbytes
            self.cv.notify_all()


class GpuExecutor:
    """Runs one batch on its own stream through the C ABI."""

    def __init__(self):
        import ctypes

        from .lib import check, lib
        self._lib = lib()
        self._check = check
        s = ctypes.c_void_p()
        check(self._lib.gcdt_stream_create(ctypes.byref(s)), "gcdt_stream_create")
        self.stream = s.value

    def close(self):
        if self.stream:
            self._lib.gcdt_stream_destroy(self.stream)
            self.stream = None

    def run(self, batch):
        from .api import DeviceBuffer, pack_instances, solve_batch_device, solve_device
        L = self._lib
        if len(batch) == 1 and batch[0].n > 0:
            r = batch[0]
            dp = DeviceBuffer(8 * r.n)
            self._check(L.gcdt_memcpy_h2d_async(dp.ptr, r.pts.ctypes.data, 8 * r.n, self.stream), "h2d")
            dc = None
            if len(r.cons):
                dc = DeviceBuffer(8 * len(r.cons))
                self._check(L.gcdt_memcpy_h2d_async(dc.ptr, r.cons.ctypes.data, 8 * len(r.cons), self.stream), "h2d")
            dt = DeviceBuffer(12 * (2 * r.n + 2))
            try:
                solve_device(dp, r.n, dc, len(r.cons), dt, None, stream=self.stream)
                self._check(L.gcdt_stream_synchronize(self.stream), "sync")
            finally:
                dp.free()
                if dc is not None:
                    dc.free()
            r._dev = dt
            return
        flat_p, poff, flat_c, coff, toff = pack_instances([(r.pts, r.cons) for r in batch])
        dp = DeviceBuffer(flat_p.nbytes)
        self._check(L.gcdt_memcpy_h2d_async(dp.ptr, flat_p.ctypes.data, flat_p.nbytes, self.stream), "h2d")
        dc = None
        if len(flat_c):
            dc = DeviceBuffer(flat_c.nbytes)
            self._check(L.gcdt_memcpy_h2d_async(dc.ptr, flat_c.ctypes.data, flat_c.nbytes, self.stream), "h2d")
        dt = DeviceBuffer(12 * int(toff[-1]))
        try:
            solve_batch_device(dp, poff, dc, coff, dt, stream=self.stream)
            allt = np.empty((int(toff[-1]), 3), np.int32)
            self._check(L.gcdt_memcpy_d2h_async(allt.ctypes.data, dt.ptr, allt.nbytes, self.stream), "d2h")
            self._check(L.gcdt_stream_synchronize(self.stream), "sync")
        finally:
            for b in (dp, dc, dt):
                if b is not None:
                    b.free()
        for i, r in enumerate(batch):
            r.result = allt[toff[i]:toff[i + 1]]


@dataclass
class ServeStats:
    completed: int = 0
    failed: int = 0
    makespan_s: float = 0.0
    latencies_s: list = field(default_factory=list)
    batches: int = 0
    batch_sizes: list = field(default_factory=list)
    peak_reserved: int = 0

    def summary(self, qps: float):
        lat = np.array(self.latencies_s) * 1000.0 if self.latencies_s else np.zeros(1)
        return {"qps_offered": qps, "requests": self.completed + self.failed, "completed": self.completed,
                "failed": self.failed, "cdts_per_s": round(self.completed / max(self.makespan_s, 1e-9), 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "batches": self.batches,
                "mean_batch": round(float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0, 2),
                "peak_reserved_bytes": int(self.peak_reserved)}


class Server:
    def __init__(self, executor_factory=GpuExecutor, workers: int = 3, hbm_budget: int = 64 << 30,
                 small_n: int = 65536, batch_points: int = 2_000_000, clock=time.perf_counter):
        self.executor_factory = executor_factory
        self.workers = workers
        self.ledger = Ledger(hbm_budget)
        self.small_n = small_n
        self.batch_points = batch_points
        self.clock = clock
        self.waiting = collections.deque()
        self.cv = self.ledger.cv
        self.closed = False
        self.stats = ServeStats()
        self.t0 = 0.0

    # ---- batching policy (pure; unit-tested on the CPU)
    def take_batch(self):
        """Pop the next batch from `waiting` (FIFO) if its footprint fits the
        ledger; returns (batch, bytes) or (None, 0).  Caller holds the lock."""
        if not self.waiting:
            return None, 0
        head = self.waiting[0]
        if head.n > self.small_n:
            need = footprint(head.n)
            if not self.ledger.fits(need):
                return None, 0
            self.waiting.popleft()
            return [head], need
        batch, pts, need = [], 0, BYTES_PER_REQUEST
        while self.waiting:
            r = self.waiting[0]
            if r.n > self.small_n or (batch and pts + r.n > self.batch_points):
                break
            extra = BYTES_PER_POINT * max(r.n, 1)
            if batch and not self.ledger.fits(need + extra):
                break
            batch.append(self.waiting.popleft())
            pts += r.n
            need += extra
        if not self.ledger.fits(need) and self.ledger.used > 0:
            for r in reversed(batch):
                self.waiting.appendleft(r)
            return None, 0
        return batch, need

    def now(self) -> float:
        return self.clock() - self.t0

    def _worker(self):
        ex = self.executor_factory()
        try:
            while True:
                with self.cv:
                    while True:
                        batch, need = self.take_batch()
                        if batch:
                            self.ledger.reserve(need)
                            break
                        if self.closed and not self.waiting:
                            return
                        self.cv.wait(timeout=0.05)
                t = self.now()
                for r in batch:
                    r.t_admit = t
                    r.batch_size = len(batch)
                try:
                    ex.run(batch)
                    err = None
                except Exception as e:  # one failing batch does not stop the server
                    err = str(e)
                t = self.now()
                with self.cv:
                    for r in batch:
                        r.t_done = t
                        r.error = err
                        if err is None:
                            self.stats.completed += 1
                            self.stats.latencies_s.append(r.latency)
                        else:
                            self.stats.failed += 1
                    self.stats.batches += 1
                    self.stats.batch_sizes.append(len(batch))
                self.ledger.release(need)
        finally:
            if hasattr(ex, "close"):
                ex.close()

    def run(self, requests, realtime: bool = True):
        """Feed `requests` at their arrival times (seconds, relative) and serve
        them to completion; with realtime=False all of them are waiting at
        the start.  Returns the stats object."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.t0 = self.clock()
        threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(self.workers)]
        if not realtime:  # everything is waiting before the first worker looks
            with self.cv:
                self.waiting.extend(reqs)
        for th in threads:
            th.start()
        for r in reqs if realtime else ():
            dt = r.arrival - self.now()
            if dt > 0:
                time.sleep(dt)
            with self.cv:
                self.waiting.append(r)
                self.cv.notify_all()
        with self.cv:
            self.closed = True
            self.cv.notify_all()
        for th in threads:
            th.join()
        self.stats.makespan_s = max((r.t_done for r in reqs), default=0.0) - min((r.arrival for r in reqs), default=0.0)
        self.stats.peak_reserved = self.ledger.peak
        return self.stats
