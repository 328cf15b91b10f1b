"""Continuous-batching spectrum server over device-resident line databases.

Requests (database name + conditions) arrive at any time.  The loop:

* plans each arrival (``LineDatabase.request_bytes``: workspace bytes and
  sizes, nothing allocated) and queues it;
* admits queued requests in arrival order into a new batch while the bytes
  of all in-flight batches stay under the HBM budget and a stream is free;
* launches the batch on its own stream (one launch set for all its
  requests) followed by an asynchronous copy of the outputs to pinned host
  memory and an event;
* polls in-flight events, completes finished batches, frees their budget.

Several batches are in flight on different streams, so requests that
arrive while others run are admitted as soon as memory and a stream free
up, and their kernels overlap on the device.  Each request's latency is
completion minus arrival.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from . import _native, engine


@dataclasses.dataclass
class ServedRequest:
    rid: int
    db: str
    cond: object
    t_arrival: float
    bytes: int = 0
    n_lines: int = 0
    n_v: int = 0
    t_start: float = 0.0
    t_done: float = 0.0
    result: np.ndarray | None = None  # abscoeff: float32, or float64 when its range needs it (_unscale)
    error: str | None = None

    @property
    def latency(self):
        return self.t_done - self.t_arrival


def _unscale(a, s):
    """The stored spectrum times 2**s, out of the pinned buffer: float32, as
    the reference's gpuApp.iterate returns it (a power-of-two scale is exact
    there), or float64 when the scaled values leave float32's normal range."""
    if s == 0:
        return a.copy()
    peak = float(np.max(np.abs(a))) if a.size else 0.0
    if abs(s) <= 126 and (peak == 0.0 or 2.0 ** -100 < peak * 2.0 ** s < 2.0 ** 100):
        return a * np.float32(2.0 ** s)  # exact: a power of two, results stay normal
    return np.ldexp(a.astype(np.float64), s)


class _Flight:
    def __init__(self, batch, reqs, outs, host, event, stream, nbytes):
        self.batch, self.reqs, self.outs, self.host = batch, reqs, outs, host
        self.event, self.stream, self.bytes = event, stream, nbytes


class SpectrumServer:
    """Serve spectra from resident databases with continuous batching.

    ``dbs``: mapping name -> LineDatabase (same device).  ``budget_bytes``:
    HBM allowed for in-flight workspaces.  ``max_batch``: requests per
    launch set.  ``n_streams``: batches in flight."""

    def __init__(self, device, dbs, budget_bytes=32 << 30, max_batch=32, n_streams=4):
        import torch

        self.torch = torch
        self.device = device
        self.dbs = dict(dbs)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.free_streams = [torch.cuda.Stream() for _ in range(n_streams)]
        self.queue = collections.deque()
        self.flights = []
        self.in_flight_bytes = 0
        self.done = []
        self.rejected = []
        self._next = 0
        self.max_in_flight = 0  # most batches in flight at once

    # -- requests ---------------------------------------------------------------
    def submit(self, db_name, cond, t_arrival=None):
        """Queue one request; returns its ServedRequest record."""
        r = ServedRequest(self._next, db_name, cond, time.perf_counter() if t_arrival is None else t_arrival)
        self._next += 1
        db = self.dbs[db_name]
        try:
            req = engine.request_from_conditions(db, cond)
            r.bytes, inf = db.request_bytes(req)
            r.n_v = int(inf.n_v)
            r.n_lines = int(inf.line_end - inf.line_begin)
            r._req = req
        except _native.LdmError as e:
            r.error = str(e)
            r.t_done = time.perf_counter()
            self.rejected.append(r)
            return r
        if r.bytes > self.budget:
            r.error = f"request needs {r.bytes} B, above the {self.budget} B budget"
            r.t_done = time.perf_counter()
            self.rejected.append(r)
            return r
        self.queue.append(r)
        return r

    # -- loop ---------------------------------------------------------------------
    def _launch(self):
        torch = self.torch
        if not self.queue or not self.free_streams:
            return False
        head_db = self.queue[0].db
        picked, nbytes = [], 0
        rest = collections.deque()
        while self.queue and len(picked) < self.max_batch:
            r = self.queue.popleft()
            if r.db != head_db or self.in_flight_bytes + nbytes + r.bytes > self.budget:
                rest.append(r)
                if r.db == head_db:
                    break  # keep arrival order for this database
                continue
            picked.append(r)
            nbytes += r.bytes
        while rest:
            self.queue.appendleft(rest.pop())
        if not picked:
            return False
        stream = self.free_streams.pop()
        db = self.dbs[head_db]
        t0 = time.perf_counter()
        try:
            batch = engine.SpectrumBatch(db, [r._req for r in picked])
        except _native.BudgetExceeded:
            for r in reversed(picked):
                self.queue.appendleft(r)
            self.free_streams.append(stream)
            return False
        nbytes = batch.workspace_bytes()
        with torch.cuda.stream(stream):
            outs = batch.alloc_outputs()
            batch.run(outs, stream)
            host = [torch.empty(o.numel(), dtype=torch.float32, pin_memory=True) for o in outs]
            for h, o in zip(host, outs):
                h.copy_(o, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        for r in picked:
            r.t_start = t0
        self.flights.append(_Flight(batch, picked, outs, host, ev, stream, nbytes))
        self.in_flight_bytes += nbytes
        self.max_in_flight = max(self.max_in_flight, len(self.flights))
        return True

    def _complete(self):
        still = []
        progressed = False
        for f in self.flights:
            if not f.event.query():
                still.append(f)
                continue
            t = time.perf_counter()
            for i, r in enumerate(f.reqs):
                inf = f.batch.info(i)
                if inf.status:
                    r.error = f"device status {inf.status:#x} ({engine.describe_status(inf.status)})"
                else:
                    r.result = _unscale(f.host[i].numpy(), int(inf.scale_log2))
                r.t_done = t
                self.done.append(r)
            f.batch.close()
            self.in_flight_bytes -= f.bytes
            self.free_streams.append(f.stream)
            progressed = True
        self.flights = still
        return progressed

    def step(self):
        """One scheduling round: complete finished batches, then launch."""
        p = self._complete()
        while self._launch():
            p = True
        return p

    def drain(self, timeout=120.0):
        t_end = time.perf_counter() + timeout
        while (self.queue or self.flights) and time.perf_counter() < t_end:
            if not self.step():
                time.sleep(50e-6)
        if self.queue or self.flights:
            raise TimeoutError("server did not drain")

    # -- driving a request stream ---------------------------------------------------
    def run_stream(self, arrivals, timeout=300.0):
        """Feed (t_offset, db_name, cond) arrivals by wall clock, serve them
        and return the completed records."""
        t0 = time.perf_counter()
        pending = sorted(arrivals, key=lambda a: a[0])
        k = 0
        t_end = t0 + timeout
        while (k < len(pending) or self.queue or self.flights) and time.perf_counter() < t_end:
            now = time.perf_counter()
            while k < len(pending) and t0 + pending[k][0] <= now:
                self.submit(pending[k][1], pending[k][2], t_arrival=t0 + pending[k][0])
                k += 1
            if not self.step():
                time.sleep(50e-6)
        if k < len(pending) or self.queue or self.flights:
            raise TimeoutError("request stream not served in time")
        return self.done


def poisson_arrivals(rate, duration, seed=1004):
    """Arrival offsets (s) of a Poisson process of ``rate`` per second."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        out.append(t)


def latency_summary(records):
    lat = np.array([r.latency for r in records if r.error is None])
    if lat.size == 0:
        return {}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_ms": float(lat.max() * 1e3), "served": int(lat.size)}
