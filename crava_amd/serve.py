"""Continuous batching of inversion requests on one GPU.

Requests arrive at any time.  The server keeps
  * a waiting queue, from which requests are admitted in arrival order while the device bytes
    of every admitted request (its grids, held from admission to release) stay within an HBM
    budget;
  * an ingest stage that materialises an admitted request's inputs on the device, after which the
    request is ready.  An executor with `ingest_async` runs ingests off the scheduling loop (each
    on its own stream), so that dispatch, completions and admission go on while inputs are built;
  * `lanes` execution lanes (one HIP stream each).  Whenever a lane is idle, every ready request
    (up to `max_batch`) is gathered into one batch (cg_batch: each FFT pass, the posterior and
    the colouring of all of them share launches) and enqueued on that lane.  Lanes run
    concurrently on the device, so requests that arrive while others are computing are batched
    and run as soon as a lane frees up, overlapping the work already on the device;
  * completion polling (non-blocking event queries); a finished request hands its
    device-resident results to the caller's `on_done`, then its grids are released and its bytes
    return to the budget.

Each result is the one the request gives alone: a batch computes every inversion with the same
kernels and descriptors whatever else is in it.

The device work goes through an executor (GpuExecutor below, over the C ABI); the scheduling
itself is device-independent so that it can be tested with a simulated executor.
"""
import collections
import time

import numpy as np


class Request:
    __slots__ = ("rid", "spec", "arrival", "admitted", "ready", "started", "finished", "lane", "bytes", "work",
                 "cells")

    def __init__(self, rid, spec, arrival, nbytes, cells):
        self.rid, self.spec, self.arrival = rid, spec, arrival
        self.admitted = self.ready = self.started = self.finished = None
        self.lane = None
        self.bytes = nbytes
        self.cells = cells
        self.work = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    """Scheduler: admission against `hbm_budget` bytes, `lanes` concurrent batch lanes,
    batches of up to `max_batch` ready requests.  `clock` returns seconds."""

    def __init__(self, executor, hbm_budget=200e9, lanes=3, max_batch=32, ingest_per_step=1,
                 clock=time.perf_counter, asynchronous_ingest=True):
        if lanes < 1 or max_batch < 1 or ingest_per_step < 1:
            raise ValueError("lanes, max_batch and ingest_per_step must be >= 1")
        self.ingest_per_step = ingest_per_step
        # ingest off the scheduling loop when the executor can (asynchronous_ingest=False: every
        # request admitted in a step is ready before that step dispatches)
        self.asynchronous = asynchronous_ingest and hasattr(executor, "ingest_async")
        self.ex = executor
        self.budget = float(hbm_budget)
        self.in_use = 0.0
        self.peak = 0.0
        self.max_batch = max_batch
        self.clock = clock
        self.waiting = collections.deque()
        self.ingesting = collections.deque()  # (request, handle) of asynchronous ingests, admission order
        self.ready = collections.deque()
        self.running = [None] * lanes  # per lane: (batch handle, [requests]) or None
        self.done = []
        self.batches = 0
        self.batch_sizes = []
        self._next = 0

    # ---------------------------------------------------------------- client side
    def submit(self, spec, arrival=None):
        nbytes, cells = self.ex.footprint(spec), self.ex.cells(spec)
        if nbytes > self.budget:
            raise ValueError("request needs %.1f GB, more than the HBM budget" % (nbytes / 1e9))
        r = Request(self._next, spec, self.clock() if arrival is None else arrival, nbytes, cells)
        self._next += 1
        self.waiting.append(r)
        return r

    def idle(self):
        return (not self.waiting and not self.ingesting and not self.ready
                and all(x is None for x in self.running))

    # ---------------------------------------------------------------- scheduling step
    def step(self, on_done=None):
        """One pass of the loop: completions, dispatch to idle lanes, admission and ingest of up
        to `ingest_per_step` waiting requests (in arrival order, while they fit the budget),
        dispatch again.  Returns True if anything moved."""
        moved = self._complete(on_done)
        moved |= self._ingested()
        moved |= self._dispatch()
        asynchronous = self.asynchronous
        n = 0
        while n < self.ingest_per_step and self.waiting and self.in_use + self.waiting[0].bytes <= self.budget:
            r = self.waiting.popleft()
            r.admitted = self.clock()
            self.in_use += r.bytes
            self.peak = max(self.peak, self.in_use)
            if asynchronous:
                self.ingesting.append((r, self.ex.ingest_async(r)))
            else:
                r.work = self.ex.ingest(r)
                r.ready = self.clock()
                self.ready.append(r)
            n += 1
        if n:
            moved = True
            moved |= self._ingested()
            moved |= self._dispatch()
        return moved

    def _ingested(self):
        """Ready requests from finished asynchronous ingests, in admission order."""
        moved = False
        while self.ingesting and self.ingesting[0][1].done():
            r, h = self.ingesting.popleft()
            r.work = h.result()
            r.ready = self.clock()
            self.ready.append(r)
            moved = True
        return moved

    def _dispatch(self):
        moved = False
        for lane, slot in enumerate(self.running):
            if slot is not None or not self.ready:
                continue
            group = []
            while self.ready and len(group) < self.max_batch:
                group.append(self.ready.popleft())
            now = self.clock()
            for r in group:
                r.started, r.lane = now, lane
            self.running[lane] = (self.ex.launch(lane, group), group)
            self.batches += 1
            self.batch_sizes.append(len(group))
            moved = True
        return moved

    def _complete(self, on_done):
        moved = False
        for lane, slot in enumerate(self.running):
            if slot is None or not self.ex.finished(slot[0]):
                continue
            now = self.clock()
            for r in slot[1]:
                r.finished = now
                if on_done is not None:
                    on_done(r)
                self.ex.release(r)
                self.in_use -= r.bytes
                self.done.append(r)
            self.ex.retire(slot[0])
            self.running[lane] = None
            moved = True
        return moved

    def drain(self, on_done=None, poll_s=50e-6):
        while not self.idle():
            if not self.step(on_done):
                time.sleep(poll_s)

    # ---------------------------------------------------------------- report
    def report(self, t0=None, t1=None):
        done = self.done
        if not done:
            return {}
        lat = np.array([r.latency for r in done]) * 1e3
        t0 = min(r.arrival for r in done) if t0 is None else t0
        t1 = max(r.finished for r in done) if t1 is None else t1
        cells = sum(r.cells for r in done)
        return {"requests": len(done), "cells": cells, "seconds": t1 - t0, "cells_per_s": cells / (t1 - t0),
                "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                               "mean": float(lat.mean()), "max": float(lat.max())},
                "batches": self.batches, "mean_batch": float(np.mean(self.batch_sizes)),
                "peak_admitted_gb": self.peak / 1e9, "budget_gb": self.budget / 1e9}


class _Worker:
    """A host thread that owns one device context: it runs the tasks given to it in order, so a
    context is only ever used by its own thread (contexts are not re-entrant)."""

    def __init__(self, device):
        import queue
        import threading
        from . import Context
        self.ctx = Context(device)
        self.q = queue.Queue()
        self.pending = 0  # tasks submitted and not finished (for choosing the least busy worker)
        self._lock = threading.Lock()
        self.t = threading.Thread(target=self._run, daemon=True)
        self.t.start()

    def _run(self):
        while True:
            task = self.q.get()
            if task is None:
                return
            fn, fut = task
            if fut.set_running_or_notify_cancel():
                try:
                    fut.set_result(fn(self.ctx))
                except BaseException as e:  # reported to the caller through the future
                    fut.set_exception(e)
            with self._lock:
                self.pending -= 1

    def submit(self, fn):
        from concurrent.futures import Future
        fut = Future()
        with self._lock:
            self.pending += 1
        self.q.put((fn, fut))
        return fut

    def close(self):
        self.q.put(None)
        self.t.join()
        self.ctx.close()


class GpuExecutor:
    """Device side of the server over the C ABI: inputs are materialised by `ingest_workers` host
    threads, each with its own context and stream (crava_amd.synthetic.Problem: seeded synthetic
    seismic and priors, built on the device), concurrently with the batches, which run on one
    context per lane.  A request's grids are released by the worker that built them."""

    def __init__(self, device=0, lanes=3, graph=False, ingest_workers=2):
        from . import Context
        self.workers = [_Worker(device) for _ in range(max(1, ingest_workers))]
        self.lane_ctx = [Context(device) for _ in range(lanes)]
        self.graph = graph
        self.owner = {}

    def footprint(self, spec):
        from . import synthetic as S
        return S.footprint(spec)

    def cells(self, spec):
        from . import synthetic as S
        return S.cells(spec)

    def ingest_async(self, r):
        from . import synthetic as S
        w = min(self.workers, key=lambda x: x.pending)
        self.owner[r.rid] = w
        spec = r.spec
        return w.submit(lambda ctx: S.Problem(ctx, spec))  # synchronises the worker's stream at the end

    def ingest(self, r):
        return self.ingest_async(r).result()

    def launch(self, lane, group):
        from . import Batch
        b = Batch(self.lane_ctx[lane], [r.work.item() for r in group])
        if self.graph:
            b.use_graph(True)
        b.run()
        return b

    def finished(self, b):
        return b.done()

    def retire(self, b):
        b.close()

    def release(self, r):
        work = r.work
        r.work = None
        self.owner.pop(r.rid).submit(lambda ctx: work.close())

    def close(self):
        for w in self.workers:
            w.close()
        for c in self.lane_ctx:
            c.close()


def poisson_arrivals(n, rate_hz, seed):
    """Arrival times (s, from 0) of n requests of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))


def run_open_loop(server, specs, arrivals, on_done=None, poll_s=50e-6):
    """Drives the server with requests arriving at the given times (s from the start): submits
    each when its time comes, steps the scheduler in between, and returns the server report over
    the run."""
    t0 = server.clock()
    i = 0
    while i < len(specs) or not server.idle():
        now = server.clock() - t0
        while i < len(specs) and arrivals[i] <= now:
            server.submit(specs[i], arrival=t0 + arrivals[i])
            i += 1
        if not server.step(on_done):
            time.sleep(poll_s)
    return server.report(t0=t0)
