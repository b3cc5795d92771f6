"""Serving: requests submitted from several threads at staggered times are
batched while others run; every result equals the query run alone on the
plain path (counts and cap flags, bit-exact), and latencies are reported."""
import random
import threading
import time

import pytest

import clustersm_amd as csm

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csm.Context(0, 24 << 30)
    yield c
    c.close()


def test_serve_results_equal_alone(ctx):
    hg = csm.gen_rmat(13, 80000, 8, 1.0, 17)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(23)
    plans = [csm.gen_query(host, r.randint(3, 12), csm.QUERY_AVG_DEGREE, 3000 + i, avg_degree=r.uniform(2, 3.5)).plan()
             for i in range(48)]
    cap = 300_000
    # alone, plain path
    alone = []
    for p in plans:
        q = g.prepare(p)
        res, _ = q.match(cap=cap)
        alone.append((res.count, res.capped))
        q.free()
    ctx.serve_start(g, budget_bytes=8 << 30, max_inflight=16, cap=cap)
    tickets = [None] * len(plans)

    def client(idx):
        rr = random.Random(idx)
        for i in range(idx, len(plans), 4):
            time.sleep(rr.expovariate(400.0))  # staggered arrivals, ~2.5 ms apart
            tickets[i] = ctx.serve_submit(plans[i])

    threads = [threading.Thread(target=client, args=(k,)) for k in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    got = []
    for i, t in enumerate(tickets):
        res = ctx.serve_poll(t, wait=True)
        assert res.status == 0, (i, res.status)
        assert res.ms_latency > 0
        assert res.path in (csm.PATH_BATCH, csm.PATH_HOST, csm.PATH_BATCH_BUDGET)
        assert res.anomalies == 0
        got.append((res.count, res.capped))
    ctx.serve_stop()
    print("served", len(got), "queries; capped", sum(c for _, c in got), flush=True)
    assert got == alone
    g.free()


def test_serve_poll_and_stop_errors(ctx):
    hg = csm.gen_rmat(10, 4000, 4, 1.0, 2)
    g = ctx.upload(hg)
    host = g.to_host()
    plan = csm.gen_query(host, 4, csm.QUERY_TREE, 1).plan()
    ctx.serve_start(g, cap=1000)
    t = ctx.serve_submit(plan)
    with pytest.raises(RuntimeError):
        ctx.serve_poll(t + 12345)  # unknown ticket
    res = ctx.serve_poll(t, wait=True)
    assert res.status == 0
    with pytest.raises(RuntimeError):
        ctx.serve_poll(t)  # already delivered
    ctx.serve_stop()
    with pytest.raises(RuntimeError):
        ctx.serve_submit(plan)  # no server
    g.free()
