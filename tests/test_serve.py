"""Scheduling of the continuous-batching request loop, on the CPU.

A fake executor sleeps for a per-cell service time and returns marker
values; contexts are plain objects whose footprint is a per-cell number, so
admission by memory, eviction, bypass, cancellation, failures and draining
are all forced from the host."""
import threading
import time

import numpy as np
import pytest

from exciting_amd import serve, synth


class FakeCell:
    def __init__(self, name, nbytes, service, nk=4):
        self.name, self.nbytes, self.service = name, nbytes, service
        self.kpts = list(range(nk))


class Harness:
    def __init__(self, cells, budget, max_inflight=8, fail_on=(), fail_make=()):
        self.cells = cells
        self.live = {}            # id(ctx) -> cell name, contexts that exist now
        self.lock = threading.Lock()
        self.max_live_bytes = 0
        self.running = 0
        self.max_running = 0
        self.executed = []
        self.fail_on = set(fail_on)
        self.fail_make = set(fail_make)
        self.srv = serve.Server(cells, self.execute, self.make, self.close, budget,
                                max_inflight=max_inflight, footprint=lambda c: c.nbytes)

    def make(self, cell):
        if cell.name in self.fail_make:
            raise MemoryError("cannot build context")
        ctx = {"cell": cell.name}
        with self.lock:
            self.live[id(ctx)] = cell
            b = sum(c.nbytes for c in self.live.values())
            self.max_live_bytes = max(self.max_live_bytes, b)
        return ctx

    def close(self, ctx):
        with self.lock:
            del self.live[id(ctx)]

    def execute(self, req, ctx):
        assert ctx["cell"] == req.cell
        with self.lock:
            self.running += 1
            self.max_running = max(self.max_running, self.running)
            self.executed.append(req.rid)
        try:
            time.sleep(self.cells[req.cell].service)
            if req.rid in self.fail_on:
                raise RuntimeError("planted executor failure")
            return [np.full(3, 100.0 * req.rid + k) for k in req.kpoints]
        finally:
            with self.lock:
                self.running -= 1


def test_all_requests_complete_with_their_own_results():
    cells = {"a": FakeCell("a", 10, 0.02), "b": FakeCell("b", 20, 0.03)}
    h = Harness(cells, budget=1000, max_inflight=4)
    reqs = [serve.Request(rid=i, cell="ab"[i % 2], kpoints=[i % 3, 3], arrival=0.005 * i)
            for i in range(12)]
    h.srv.run(reqs)
    for r in reqs:
        assert r.status == "done"
        assert [w[0] for w in r.eigenvalues] == [100.0 * r.rid + k for k in r.kpoints]
        assert r.arrival <= r.admitted <= r.finished
    assert h.max_running <= 4
    assert h.srv.peak_inflight <= 4


def test_continuous_admission_overlaps_running_work():
    """A request that arrives while another runs starts before that one ends."""
    cells = {"a": FakeCell("a", 10, 0.2)}
    h = Harness(cells, budget=100, max_inflight=4)
    r0 = serve.Request(rid=0, cell="a", kpoints=[0], arrival=0.0)
    r1 = serve.Request(rid=1, cell="a", kpoints=[1], arrival=0.05)
    h.srv.run([r0, r1])
    assert r1.admitted < r0.finished
    assert h.max_running == 2
    assert h.srv.pool.created == 2      # second context built: first is busy


def test_idle_contexts_are_reused():
    cells = {"a": FakeCell("a", 10, 0.01)}
    h = Harness(cells, budget=100, max_inflight=1)
    reqs = [serve.Request(rid=i, cell="a", kpoints=[0]) for i in range(5)]
    h.srv.run(reqs)
    assert all(r.status == "done" for r in reqs)
    assert h.srv.pool.created == 1


def test_admission_respects_memory_budget():
    cells = {"a": FakeCell("a", 40, 0.05), "b": FakeCell("b", 70, 0.05)}
    h = Harness(cells, budget=100, max_inflight=8)
    reqs = [serve.Request(rid=i, cell="ab"[(i // 2) % 2], kpoints=[0], arrival=0.002 * i)
            for i in range(12)]
    h.srv.run(reqs)
    assert all(r.status == "done" for r in reqs)
    assert h.max_live_bytes <= 100
    assert h.srv.peak_bytes <= 100
    assert h.srv.pool.evicted >= 1


def test_bypass_lets_small_requests_pass_a_waiting_large_one():
    """small(long) runs; big does not fit beside it and waits; a later small
    request fits and is admitted before big."""
    cells = {"s": FakeCell("s", 30, 0.25), "big": FakeCell("big", 60, 0.02),
             "t": FakeCell("t", 30, 0.02)}
    h = Harness(cells, budget=89, max_inflight=8)
    r0 = serve.Request(rid=0, cell="s", kpoints=[0], arrival=0.0)
    r1 = serve.Request(rid=1, cell="big", kpoints=[0], arrival=0.02)
    r2 = serve.Request(rid=2, cell="t", kpoints=[0], arrival=0.04)
    h.srv.run([r0, r1, r2])
    assert all(r.status == "done" for r in (r0, r1, r2))
    assert r2.admitted < r1.admitted
    assert r1.admitted >= r0.finished - 1e-3     # big waited for memory
    assert h.max_live_bytes <= 89
    assert h.srv.bypassed == 1 and h.srv.hbm_waits == 1


def test_request_larger_than_budget_fails_alone():
    cells = {"a": FakeCell("a", 10, 0.01), "huge": FakeCell("huge", 1000, 0.01)}
    h = Harness(cells, budget=100)
    reqs = [serve.Request(rid=0, cell="a", kpoints=[0]),
            serve.Request(rid=1, cell="huge", kpoints=[0]),
            serve.Request(rid=2, cell="a", kpoints=[1])]
    h.srv.run(reqs)
    assert [r.status for r in reqs] == ["done", "failed", "done"]
    assert "budget" in reqs[1].error
    assert 1 not in h.executed


def test_cancel_queued_request():
    cells = {"a": FakeCell("a", 10, 0.1)}
    h = Harness(cells, budget=100, max_inflight=1)
    reqs = [serve.Request(rid=i, cell="a", kpoints=[0], arrival=0.0) for i in range(3)]
    threading.Timer(0.03, h.srv.cancel, args=(2,)).start()
    h.srv.run(reqs)
    assert [r.status for r in reqs] == ["done", "done", "cancelled"]
    assert 2 not in h.executed


def test_executor_failure_is_contained():
    cells = {"a": FakeCell("a", 10, 0.01)}
    h = Harness(cells, budget=100, max_inflight=2, fail_on={1})
    reqs = [serve.Request(rid=i, cell="a", kpoints=[0]) for i in range(4)]
    h.srv.run(reqs)
    assert [r.status for r in reqs] == ["done", "failed", "done", "done"]
    assert "planted" in reqs[1].error
    # the failed request's context went back to the pool and was reused
    assert h.srv.pool.created <= 2


def test_context_build_failure_releases_its_reservation():
    cells = {"a": FakeCell("a", 60, 0.01), "bad": FakeCell("bad", 60, 0.01)}
    h = Harness(cells, budget=100, max_inflight=1, fail_make={"bad"})
    reqs = [serve.Request(rid=0, cell="bad", kpoints=[0]),
            serve.Request(rid=1, cell="a", kpoints=[0])]
    h.srv.run(reqs)
    assert [r.status for r in reqs] == ["failed", "done"]
    assert h.srv.pool.bytes == 60


def test_drains_before_returning():
    cells = {"a": FakeCell("a", 10, 0.1)}
    h = Harness(cells, budget=100, max_inflight=4)
    reqs = [serve.Request(rid=i, cell="a", kpoints=[0], arrival=0.01 * i) for i in range(4)]
    h.srv.run(reqs)
    assert not h.srv.inflight
    assert all(r.finished > 0 for r in reqs)


def test_poisson_requests_and_summary():
    cells = {"a": FakeCell("a", 10, 0.0, nk=16), "b": FakeCell("b", 10, 0.0, nk=1)}
    reqs = serve.poisson_requests(400, rate=50.0, cells=cells, seed=3, kmax=16)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 50.0) < 0.3 / 50.0
    assert all(1 <= len(r.kpoints) <= len(cells[r.cell].kpts) for r in reqs)
    assert all(len(set(r.kpoints)) == len(r.kpoints) for r in reqs)
    assert {r.nmod for r in reqs} == {0, 9}
    for i, r in enumerate(reqs):
        r.status, r.admitted, r.finished = "done", r.arrival, r.arrival + 0.01 * (i % 100 + 1)
    s = serve.summarize(reqs, wall=10.0)
    assert s["done"] == 400 and s["setups"] == sum(len(r.kpoints) for r in reqs)
    assert s["latency_s"]["p50"] == pytest.approx(0.505, abs=0.011)
    assert s["latency_s"]["p99"] == pytest.approx(0.99, abs=0.011)


def test_context_footprint_covers_setup_and_solve():
    s = synth.build(synth.config(0))
    n = s.ngkmax + s.nlotot
    b = serve.context_bytes(s)
    assert b > 7 * 16 * n * n          # H, O and the eigensolve temporaries
    assert serve.context_bytes(s, nmod_max=0) < b
    # every slot holds its own apwalm, H, O and eigensolve temporaries
    assert serve.context_bytes(s, slots=2) - b >= (2 * 16 + 8 * 16) * n * n


def test_serving_cells_span_two_to_sixty_four_atoms():
    specs = synth.serving_cells(kmax=16)
    natoms = sorted(sum(len(p) for p in spec.positions) for spec, _ in specs.values())
    assert natoms == [2, 8, 16, 32, 64]
    assert all(1 <= len(ks) <= 16 for _, ks in specs.values())


def test_segment_records_pool_into_one_summary(tmp_path):
    """Record files of consecutive segments of one stream merge into the
    summary of the whole stream: percentiles over every request, setups/s
    over the summed wall time, counters summed; overlapping segments are
    refused."""
    cells = {"a": FakeCell("a", 10, 0.0, nk=16), "b": FakeCell("b", 10, 0.0, nk=1)}
    reqs = serve.poisson_requests(300, rate=50.0, cells=cells, seed=4, kmax=16)
    for i, r in enumerate(reqs):
        r.status, r.admitted, r.finished = "done", r.arrival, r.arrival + 0.01 * (i % 100 + 1)
        r.launches = 1 + i % 3
    reqs[7].status = "failed"
    whole = serve.summarize(reqs, wall=12.0)
    paths = []
    for j, (a, b, wall) in enumerate(((0, 120, 5.0), (120, 300, 7.0))):
        head = {"segment": [a, b], "wall_s": wall, "contexts_created": 2,
                "contexts_evicted": j, "bypassed": 3 * j, "hbm_waits": 1,
                "peak_inflight": 4 + j, "peak_context_gib": 1.5 - j,
                "batching": {"launches": 10 + j, "kpoints_per_launch": {"max": 5 + j},
                             "cells_per_launch": {"max": 2}}}
        paths.append(str(tmp_path / f"seg{j}.jsonl"))
        serve.write_records(paths[-1], reqs[a:b], head)
    m = serve.merge_records(paths)
    for key in ("requests", "done", "failed", "setups", "setups_per_s", "latency_s",
                "queue_wait_s", "launches_per_request"):
        assert m[key] == pytest.approx(whole[key]), key
    assert m["segments"] == [[0, 120], [120, 300]]
    assert (m["contexts_evicted"], m["bypassed"], m["hbm_waits"]) == (1, 3, 2)
    assert m["peak_inflight"] == 5 and m["max_kpoints_per_launch"] == 6
    assert m["setup_launches"] == 21
    with pytest.raises(ValueError, match="overlap"):
        serve.merge_records([paths[0], paths[0]])
