"""Admission policy of the served form, on the CPU with stand-in workers
(each 'runs' a query by sleeping): at most one heavy query at a time, light
queries overlapping it and passing a waiting heavy one, heavy queries in
arrival order, the HBM budget respected, every query completed once."""

import threading
import time

from gpudbmetalcodegen_amd import server as S


class FakeDB:
    rows = {"scan": 120, "join": 150, "anti": 30, "like": 7}

    def rows_for(self, kind):
        return self.rows[kind]


class FakeWorker:
    def __init__(self, log, lock, dur):
        self.log, self.lock, self.dur = log, lock, dur

    def run(self, q):
        with self.lock:
            self.log.append(("start", q.qid, q.kind, time.perf_counter()))
        time.sleep(self.dur[q.kind])
        with self.lock:
            self.log.append(("end", q.qid, q.kind, time.perf_counter()))
        q.columns = {"qid": [q.qid]}


def _stream(kinds, footprint=1):
    return [S.Query(i, k, {}, 0.0, rows=FakeDB.rows[k], footprint=footprint) for i, k in enumerate(kinds)]


def _run(kinds, workers=4, budget=1 << 40, footprint=1):
    log, lock = [], threading.Lock()
    dur = {"scan": 0.03, "join": 0.04, "anti": 0.005, "like": 0.005}
    pool = [FakeWorker(log, lock, dur) for _ in range(workers)]
    qs = _stream(kinds, footprint)
    rep = S.serve(FakeDB(), qs, workers=workers, budget_bytes=budget, pool=pool)
    return rep, qs, sorted(log, key=lambda e: e[3])


def test_one_heavy_query_at_a_time_light_ones_overlap():
    kinds = ["join", "scan", "anti", "like", "join", "anti", "scan", "like"]
    rep, qs, log = _run(kinds)
    assert rep["errors"] == 0 and rep["queries"] == len(kinds)
    assert all(q.columns == {"qid": [q.qid]} for q in qs)
    heavy, light, max_heavy, overlap = 0, 0, 0, False
    heavy_order = []
    for ev, qid, kind, _ in log:
        is_heavy = kind in ("join", "scan")
        d = 1 if ev == "start" else -1
        if is_heavy:
            heavy += d
            if ev == "start":
                heavy_order.append(qid)
        else:
            light += d
        max_heavy = max(max_heavy, heavy)
        overlap |= heavy > 0 and light > 0
    assert max_heavy == 1                       # heavy queries never share the device
    assert overlap                               # light queries run beside a heavy one
    assert heavy_order == sorted(heavy_order)    # heavy ones in arrival order
    # the light queries queued behind the second heavy one started before it ended
    start = {qid: t for ev, qid, _, t in log if ev == "start"}
    end = {qid: t for ev, qid, _, t in log if ev == "end"}
    assert start[2] < end[0] and start[3] < end[0]


def test_plain_form_is_one_at_a_time():
    rep, qs, log = _run(["join", "anti", "scan", "like"], workers=1)
    depth, worst = 0, 0
    for ev, *_ in log:
        depth += 1 if ev == "start" else -1
        worst = max(worst, depth)
    assert worst == 1 and rep["max_inflight"] == 1


def test_budget_keeps_memory_order():
    # every query needs the whole budget: one at a time, in arrival order
    rep, qs, log = _run(["anti", "like", "anti", "like"], budget=10, footprint=10)
    assert rep["max_inflight"] == 1
    assert [qid for ev, qid, *_ in log if ev == "start"] == [0, 1, 2, 3]


def _run_with(kinds, min_overlap_ms, service_ms):
    log, lock = [], threading.Lock()
    dur = {"scan": 0.03, "join": 0.04, "anti": 0.005, "like": 0.005}
    pool = [FakeWorker(log, lock, dur) for _ in range(4)]
    qs = _stream(kinds)
    rep = S.serve(FakeDB(), qs, workers=4, budget_bytes=1 << 40, pool=pool, min_overlap_ms=min_overlap_ms,
                  service_ms=service_ms)
    start = {qid: t for ev, qid, _, t in log if ev == "start"}
    end = {qid: t for ev, qid, _, t in log if ev == "end"}
    return rep, start, end


def test_no_overlap_with_a_query_about_to_finish():
    """With min_overlap_ms, a query that arrives while the running one is
    expected to end within that time waits for it instead of sharing the
    device; one that would wait long still starts beside it."""
    est = {"scan": 30.0, "join": 40.0, "anti": 5.0, "like": 5.0}
    rep, start, end = _run_with(["anti", "join"], 10.0, est)
    assert rep["errors"] == 0 and start[1] >= end[0]          # the join waited ~5 ms for the anti-join
    rep, start, end = _run_with(["anti", "join"], 0.0, est)
    assert start[1] < end[0]                                  # without the rule it starts beside it
    rep, start, end = _run_with(["join", "anti"], 10.0, est)
    assert start[1] < end[0]                                  # 40 ms left: the anti-join overlaps
    assert rep["service_ms"]["join"] > 30.0 and rep["service_ms"]["anti"] >= 4.0


def test_overdue_query_does_not_block_admission():
    """A running query past its expected time no longer holds the next one
    back (the estimate was wrong, the wait would be unbounded)."""
    est = {"scan": 30.0, "join": 1.0, "anti": 5.0, "like": 5.0}   # the join really takes 40 ms
    rep, start, end = _run_with(["join", "anti"], 10.0, est)
    assert rep["errors"] == 0 and start[1] < end[0]


def test_overlap_window_of_the_waiting_query():
    """min_overlap_ms < 0: a query waits for a running one that ends within
    its own expected service time — a join behind an anti-join about to end
    waits, an anti-join beside a join with 40 ms to go starts at once."""
    est = {"scan": 30.0, "join": 40.0, "anti": 5.0, "like": 5.0}
    rep, start, end = _run_with(["anti", "join"], -1.0, est)
    assert rep["errors"] == 0 and start[1] >= end[0]
    rep, start, end = _run_with(["join", "anti"], -1.0, est)
    assert start[1] < end[0]


def test_overlap_window_gives_way_to_a_backlog():
    """With more queries waiting than workers, the overlap window holds
    nothing back: the join behind an anti-join about to end starts at once
    when four more queries wait behind it."""
    est = {"scan": 30.0, "join": 40.0, "anti": 5.0, "like": 5.0}
    rep, start, end = _run_with(["anti", "join", "anti", "anti", "anti", "anti"], -1.0, est)
    assert rep["errors"] == 0 and rep["queries"] == 6
    assert start[1] < end[0]
