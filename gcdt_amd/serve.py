"""Continuous-batching CDT server on one GPU.

Requests (points on the 2^30 grid + constraint vertex-id pairs) arrive over
time.  Worker threads, each owning one non-blocking HIP stream, repeatedly
take whatever requests are waiting when they become free:

* small requests (n <= small_n) are merged into one batched solve
  (gcdt_solve_batch) of up to `batch_points` points, taken in arrival order
  from anywhere in the queue (a large request waiting ahead of them does not
  stop the merge);
* a large request is solved alone (gcdt_solve);

so requests that arrive while others are on the GPU join the next batch of
the first worker that frees up, and the workers' streams overlap on the
device.

Admission is by HBM footprint.  A batch starts only when its device bytes --
the solve's working memory plus the triangles it leaves on the device -- fit
next to what is in flight and what waits to be fetched, against the budget.
The per-point figures are the MemCache pool's measured peaks for each path
(single solve and batched solve), so the ledger never admits more than fits;
the same budget is given to the pool (gcdt_pool_set_budget), which refuses an
allocation past it.  Results stay on the device until the client reads them
(Request.fetch()), which also returns their bytes to the ledger.

A request that fails -- alone, or inside a merged batch, where
gcdt_solve_batch's per-instance status isolates it -- fails alone with its
own error; a request still waiting when its deadline passes fails with a
timeout instead of running.

The executor is pluggable: GpuExecutor drives libgcdt.so; tests substitute a
CPU executor to exercise admission and back-pressure without a GPU.
"""
from __future__ import annotations

import collections
import threading
import time
from dataclasses import dataclass, field

import numpy as np

# Device bytes a solve needs while it runs, from the MemCache pool's measured
# peaks: the single solve peaked at 9,674,490,880 B for 10,099,763 points
# with 100,000 segments (958 B per point, inputs and outputs included); the
# batched solve of the 4,096-instance workload at 17,474,913,280 B for
# 38,759,094 points (451 B per point: its sample meshes and insertion DAG are
# per instance, the flip records shared).  Per call, the fixed buffers (cell
# tables, bitmap, counters).
SINGLE_BYTES_PER_POINT = 1024
SINGLE_BYTES_PER_CALL = 64 << 20
BATCH_BYTES_PER_POINT = 480
BATCH_BYTES_PER_CALL = 16 << 20
# the triangles a finished request leaves on the device until it is fetched
RESULT_BYTES_PER_POINT = 24


def result_bytes(n: int) -> int:
    return 12 * (2 * max(n, 0) + 2)


def footprint(n: int) -> int:
    """Device bytes of one request solved alone: working memory + result."""
    return SINGLE_BYTES_PER_POINT * max(n, 1) + SINGLE_BYTES_PER_CALL + result_bytes(n)


def batch_footprint(ns) -> int:
    """Device bytes of one merged batch of requests with these point counts."""
    return BATCH_BYTES_PER_CALL + sum(BATCH_BYTES_PER_POINT * max(n, 1) + result_bytes(n) for n in ns)


class DeviceResult:
    """A request's triangles on the device: a range of a batch's output
    buffer, shared by the batch's requests; the buffer is freed when the
    last of them has been read or dropped."""

    def __init__(self, buf, offset_tris: int, count: int, owner):
        self.buf = buf
        self.offset = offset_tris
        self.count = count
        self.owner = owner  # _SharedBuffer

    def to_array(self) -> np.ndarray:
        from .lib import check, lib
        out = np.empty((self.count, 3), np.int32)
        if out.nbytes:
            check(lib().gcdt_memcpy_d2h(out.ctypes.data, self.buf.ptr + 12 * self.offset, out.nbytes), "d2h")
        return out

    def free(self):
        if self.owner is not None:
            self.owner.drop()
            self.owner = None


class _SharedBuffer:
    def __init__(self, buf, refs: int):
        self.buf = buf
        self.refs = refs
        self.lock = threading.Lock()

    def drop(self):
        with self.lock:
            self.refs -= 1
            last = self.refs == 0
        if last:
            self.buf.free()


@dataclass
class Request:
    rid: int
    arrival: float                 # seconds after the stream start
    pts: np.ndarray                # (n,2) int32 grid coordinates
    cons: np.ndarray               # (c,2) int32 local vertex ids
    deadline: float | None = None  # seconds after arrival to be admitted by, or fail
    t_admit: float = -1.0
    t_done: float = -1.0
    batch_size: int = 0
    error: str | None = None
    result: np.ndarray | None = None   # host copy once fetched
    _dev: object = None                # DeviceResult (GPU executor)
    _ledger: object = None             # returns the result's bytes on fetch
    _held: int = 0

    @property
    def n(self) -> int:
        return len(self.pts)

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival

    def fetch(self) -> np.ndarray | None:
        """The triangles (host copy), read from the device on first call; the
        device range and its ledger bytes are released then."""
        if self.result is None and self._dev is not None:
            try:
                self.result = self._dev.to_array()
            finally:
                self._dev.free()
                self._dev = None
        self._release()
        return self.result

    def _release(self):
        if self._ledger is not None and self._held:
            self._ledger.release(self._held)
        self._held = 0
        self._ledger = None


class Ledger:
    """HBM-footprint admission: reserve before a batch starts; the working
    part is released when the batch ends, the result part at fetch."""

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


STATUS_TEXT = {1: "out of device memory", 2: "invalid input", 3: "not converged", 5: "crossing constraints",
               7: "duplicate point or location failure", 8: "triangle capacity exceeded", 9: "internal check failed"}


class GpuExecutor:
    """Runs one batch on its own stream through the C ABI; returns one error
    (None for success) per request."""

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
            except Exception:
                dt.free()
                raise
            finally:
                dp.free()
                if dc is not None:
                    dc.free()
            r._dev = DeviceResult(dt, 0, 2 * r.n + 2, _SharedBuffer(dt, 1))
            return [None]
        flat_p, poff, flat_c, coff, toff = pack_instances([(r.pts, r.cons) for r in batch])
        dp = DeviceBuffer(flat_p.nbytes)
        self._check(L.gcdt_memcpy_h2d_async(dp.ptr, flat_p.ctypes.data, flat_p.nbytes, self.stream), "h2d")
        dc = None
        if len(flat_c):
            dc = DeviceBuffer(flat_c.nbytes)
            self._check(L.gcdt_memcpy_h2d_async(dc.ptr, flat_c.ctypes.data, flat_c.nbytes, self.stream), "h2d")
        dt = DeviceBuffer(12 * int(toff[-1]))
        status = np.zeros(len(batch), np.int32)
        try:
            solve_batch_device(dp, poff, dc, coff, dt, stream=self.stream, inst_status=status)
            self._check(L.gcdt_stream_synchronize(self.stream), "sync")
        except Exception:
            dt.free()
            raise
        finally:
            for b in (dp, dc):
                if b is not None:
                    b.free()
        ok = [i for i in range(len(batch)) if status[i] == 0]
        if not ok:
            dt.free()
        shared = _SharedBuffer(dt, len(ok)) if ok else None
        errors = []
        for i, r in enumerate(batch):
            if status[i] == 0:
                r._dev = DeviceResult(dt, int(toff[i]), int(toff[i + 1] - toff[i]), shared)
                errors.append(None)
            else:
                errors.append(f"gcdt_solve_batch: {STATUS_TEXT.get(int(status[i]), int(status[i]))} "
                              f"(status {int(status[i])})")
        return errors


@dataclass
class ServeStats:
    completed: int = 0
    failed: int = 0
    timed_out: int = 0
    makespan_s: float = 0.0
    latencies_s: list = field(default_factory=list)
    batches: int = 0
    batch_sizes: list = field(default_factory=list)
    peak_reserved: int = 0

    def summary(self, qps: float):
        lat = np.array(self.latencies_s) * 1000.0 if self.latencies_s else np.zeros(1)
        return {"qps_offered": qps, "requests": self.completed + self.failed, "completed": self.completed,
                "failed": self.failed, "timed_out": self.timed_out,
                "cdts_per_s": round(self.completed / max(self.makespan_s, 1e-9), 2),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "mean_ms": round(float(lat.mean()), 3), "batches": self.batches,
                "mean_batch": round(float(np.mean(self.batch_sizes)) if self.batch_sizes else 0.0, 2),
                "peak_reserved_bytes": int(self.peak_reserved)}


class Server:
    def __init__(self, executor_factory=GpuExecutor, workers: int = 3, hbm_budget: int = 64 << 30,
                 small_n: int = 65536, batch_points: int = 2_000_000, clock=time.perf_counter,
                 pool_budget: bool | None = None):
        """pool_budget: also give hbm_budget to the device pool, which then
        refuses allocations past it (default: with the GPU executor)."""
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
        if pool_budget is None:
            pool_budget = executor_factory is GpuExecutor
        self.pool_budget = pool_budget

    # ---- batching policy (pure; unit-tested on the CPU)
    def _expire(self):
        """Requests whose deadline passed while waiting fail with a timeout.
        Caller holds the lock."""
        now = self.now()
        keep = collections.deque()
        for r in self.waiting:
            if r.deadline is not None and now - r.arrival > r.deadline:
                r.error = f"timeout: not admitted within {r.deadline:.3f} s"
                r.t_done = now
                self.stats.failed += 1
                self.stats.timed_out += 1
            else:
                keep.append(r)
        self.waiting = keep

    def take_batch(self):
        """Pop the next batch from `waiting` if its footprint fits the ledger;
        returns (batch, bytes) or (None, 0).  Caller holds the lock.

        The oldest request decides: a large one runs alone; a small one is
        merged with the next small requests in arrival order, from anywhere in
        the queue, up to batch_points points."""
        self._expire()
        if not self.waiting:
            return None, 0
        head = self.waiting[0]
        if head.n > self.small_n:
            need = footprint(head.n)
            if not self.ledger.fits(need):
                return None, 0
            self.waiting.popleft()
            return [head], need
        batch, pts = [], 0
        rest = collections.deque()
        while self.waiting:
            r = self.waiting.popleft()
            if r.n > self.small_n or (batch and pts + r.n > self.batch_points) or (
                    batch and not self.ledger.fits(batch_footprint([q.n for q in batch] + [r.n]))):
                rest.append(r)
                continue
            batch.append(r)
            pts += r.n
        self.waiting = rest
        need = batch_footprint([r.n for r in batch])
        if not self.ledger.fits(need):
            # put them back in arrival order
            self.waiting = collections.deque(sorted(list(self.waiting) + batch, key=lambda q: (q.arrival, q.rid)))
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
                        self.cv.wait(timeout=0.005)
                t = self.now()
                for r in batch:
                    r.t_admit = t
                    r.batch_size = len(batch)
                try:
                    errors = ex.run(batch)
                    if errors is None:
                        errors = [None] * len(batch)
                except Exception as e:  # one failing batch does not stop the server
                    errors = [str(e)] * len(batch)
                t = self.now()
                held = 0
                with self.cv:
                    for r, err in zip(batch, errors):
                        r.t_done = t
                        r.error = err
                        if err is None:
                            self.stats.completed += 1
                            self.stats.latencies_s.append(r.latency)
                            # the result stays reserved until it is fetched
                            r._held = result_bytes(r.n)
                            r._ledger = self.ledger
                            held += r._held
                        else:
                            self.stats.failed += 1
                    self.stats.batches += 1
                    self.stats.batch_sizes.append(len(batch))
                self.ledger.release(need - held)
        finally:
            if hasattr(ex, "close"):
                ex.close()

    def run(self, requests, realtime: bool = True):
        """Feed `requests` at their arrival times (seconds, relative) and serve
        them to completion; with realtime=False all of them are waiting at
        the start.  Returns the stats object."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        if self.pool_budget:
            from .lib import check, lib
            check(lib().gcdt_pool_set_budget(self.ledger.budget), "gcdt_pool_set_budget")
        try:
            self._serve(reqs, realtime)
        finally:
            if self.pool_budget:
                from .lib import lib
                lib().gcdt_pool_set_budget(0)
        self.stats.makespan_s = max((r.t_done for r in reqs), default=0.0) - min((r.arrival for r in reqs), default=0.0)
        self.stats.peak_reserved = self.ledger.peak
        return self.stats

    def _serve(self, reqs, realtime):
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
