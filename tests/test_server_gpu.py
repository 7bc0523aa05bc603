"""Workload 4 serving: each query's result is identical alone and in the mix
(bit-exact), admission respects the HBM scratch budget, and the served form
overlaps queries on the device."""

import pytest

import oracle
from gpudbmetalcodegen_amd import server as S
from gpudbmetalcodegen_amd import workloads as W
from tests.helpers import q1_expected

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def db(ctx):
    T = W.workload4(n_f=2_000_000, n_b=500_000, n_a=100_000, n_s=200_000)
    d = S.ResidentDB(ctx, T)
    d.host = T
    yield d
    d.close()


def run_alone(db, stream):
    pool = [S.Worker(db)]
    try:
        S.serve(db, stream, workers=1, budget_bytes=1 << 40, pool=pool)
    finally:
        for w in pool:
            w.close()
    return {q.qid: q.result for q in stream}


def test_results_identical_alone_and_in_mix(db):
    stream = S.make_stream(db, rate=400, duration=0.15, seed=11, backlog=True)
    assert len(stream) >= 20 and {q.kind for q in stream} == set(S.KINDS)
    alone = run_alone(db, stream)
    mix = S.make_stream(db, rate=400, duration=0.15, seed=11, backlog=True)
    rep = S.serve(db, mix, workers=4, budget_bytes=1 << 34)
    assert rep["errors"] == 0
    assert rep["max_inflight"] >= 2          # the served form really overlaps queries
    for q in mix:
        assert q.result == alone[q.qid], (q.qid, q.kind)


def test_scan_results_match_oracle(db):
    stream = [q for q in S.make_stream(db, rate=400, duration=0.2, seed=5, backlog=True) if q.kind == "scan"][:3]
    rep = S.serve(db, stream, workers=2, budget_bytes=1 << 34)
    assert rep["errors"] == 0
    for q in stream:
        exp = q1_expected(oracle.q1(db.host["F"], q.params["cutoff"]))
        lines = q.result.strip().split("\n")
        assert lines[0].split(",")[:2] == ["f_flag", "f_status"]
        got_counts = [int(l.split(",")[7]) for l in lines[1:]]
        assert got_counts == exp["count_order"]


def test_admission_budget_serialises_large_queries(db):
    stream = [q for q in S.make_stream(db, rate=400, duration=0.2, seed=7, backlog=True) if q.kind == "join"][:4]
    assert stream
    budget = min(q.footprint for q in stream)    # room for one join at a time
    rep = S.serve(db, stream, workers=4, budget_bytes=budget)
    assert rep["errors"] == 0 and rep["max_inflight"] == 1
    alone = run_alone(db, [S.Query(q.qid, q.kind, q.params) for q in stream])
    for q in stream:
        assert q.result == alone[q.qid]


def test_poisson_stream_reports_latency(db):
    stream = S.make_stream(db, rate=200, duration=0.3, seed=3)
    rep = S.serve(db, stream, workers=4, budget_bytes=1 << 34)
    assert rep["errors"] == 0 and rep["queries"] == len(stream)
    assert 0 < rep["p50_ms"] <= rep["p99_ms"]
