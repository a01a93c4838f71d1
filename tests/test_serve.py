"""Scheduling of the continuous-batching server (crava_amd.serve) with a simulated executor:
admission against the HBM budget, batching of requests that wait while lanes are busy,
concurrent lanes, completion of every request."""
import time

import pytest

from crava_amd import serve


class FakeExecutor:
    """Batches take (sum of request work) seconds; memory is the spec's `bytes`."""

    def __init__(self, clock):
        self.clock = clock
        self.live = 0
        self.max_live_lanes = 0
        self.running = set()

    def footprint(self, spec):
        return spec["bytes"]

    def cells(self, spec):
        return spec["cells"]

    def ingest(self, r):
        return {"spec": r.spec}

    def launch(self, lane, group):
        h = (lane, self.clock() + sum(r.spec["work"] for r in group))
        self.running.add(h)
        self.max_live_lanes = max(self.max_live_lanes, len(self.running))
        return h

    def finished(self, h):
        return self.clock() >= h[1]

    def retire(self, h):
        self.running.discard(h)

    def release(self, r):
        r.work = None


def _specs(n, nbytes=10.0, work=0.004):
    return [dict(bytes=nbytes, cells=1000 + i, work=work) for i in range(n)]


def test_every_request_completes_within_budget():
    clock = time.perf_counter
    ex = FakeExecutor(clock)
    srv = serve.Server(ex, hbm_budget=35.0, lanes=2, max_batch=4, clock=clock)
    seen = []
    orig = srv.step

    def step(on_done=None):
        moved = orig(on_done)
        seen.append(srv.in_use)
        return moved

    srv.step = step
    specs = _specs(20)
    arrivals = serve.poisson_arrivals(len(specs), 2000.0, seed=1)
    rep = serve.run_open_loop(srv, specs, arrivals)
    assert rep["requests"] == 20
    assert max(seen) <= 35.0 and srv.peak <= 35.0  # at most 3 requests of 10 admitted at once
    assert srv.in_use == 0
    assert all(r.finished >= r.started >= r.ready >= r.admitted >= r.arrival for r in srv.done)
    assert rep["latency_ms"]["p99"] >= rep["latency_ms"]["p50"] > 0


def test_waiting_requests_are_batched_and_lanes_overlap():
    clock = time.perf_counter
    ex = FakeExecutor(clock)
    srv = serve.Server(ex, hbm_budget=1e9, lanes=2, max_batch=8, clock=clock)
    specs = _specs(24, work=0.003)
    arrivals = serve.poisson_arrivals(len(specs), 5000.0, seed=2)  # faster than the lanes drain
    rep = serve.run_open_loop(srv, specs, arrivals)
    assert rep["requests"] == 24
    assert max(srv.batch_sizes) > 1  # requests that arrived while lanes were busy share a batch
    assert ex.max_live_lanes == 2  # both lanes ran at once
    assert sum(srv.batch_sizes) == 24


def test_admission_is_first_come_first_served():
    clock = time.perf_counter
    ex = FakeExecutor(clock)
    srv = serve.Server(ex, hbm_budget=25.0, lanes=1, max_batch=1, clock=clock)
    specs = [dict(bytes=20.0, cells=1, work=0.002), dict(bytes=20.0, cells=2, work=0.002),
             dict(bytes=5.0, cells=3, work=0.002)]
    for s in specs:
        srv.submit(s)
    srv.drain()
    order = sorted(srv.done, key=lambda r: r.admitted)
    assert [r.rid for r in order] == [0, 1, 2]


def test_oversized_request_is_refused():
    srv = serve.Server(FakeExecutor(time.perf_counter), hbm_budget=10.0)
    with pytest.raises(ValueError):
        srv.submit(dict(bytes=11.0, cells=1, work=0.0))


class AsyncFakeExecutor(FakeExecutor):
    """Ingest runs off the scheduling loop and takes `ingest_s` seconds."""

    class _Handle:
        def __init__(self, clock, t, value):
            self.clock, self.t, self.value = clock, t, value

        def done(self):
            return self.clock() >= self.t

        def result(self):
            return self.value

    def __init__(self, clock, ingest_s):
        super().__init__(clock)
        self.ingest_s = ingest_s
        self.async_calls = 0

    def ingest_async(self, r):
        self.async_calls += 1
        return self._Handle(self.clock, self.clock() + self.ingest_s, self.ingest(r))


def test_asynchronous_ingest_and_the_synchronous_switch():
    clock = time.perf_counter
    ex = AsyncFakeExecutor(clock, ingest_s=0.002)
    srv = serve.Server(ex, hbm_budget=1e9, lanes=2, max_batch=4, ingest_per_step=12, clock=clock)
    for s in _specs(12, work=0.002):
        srv.submit(s)
    srv.drain()
    assert ex.async_calls == 12 and len(srv.done) == 12 and sum(srv.batch_sizes) == 12
    assert all(r.started >= r.ready > r.admitted for r in srv.done)

    ex = AsyncFakeExecutor(clock, ingest_s=0.002)
    srv = serve.Server(ex, hbm_budget=1e9, lanes=2, max_batch=4, ingest_per_step=12, clock=clock,
                       asynchronous_ingest=False)
    for s in _specs(12, work=0.002):
        srv.submit(s)
    srv.drain()
    # every request ingested in the first step: the two lanes start with full batches of 4
    assert ex.async_calls == 0 and len(srv.done) == 12
    assert srv.batch_sizes[:2] == [4, 4] and sum(srv.batch_sizes) == 12
