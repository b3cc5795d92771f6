"""Serving: requests submitted from several threads at staggered times are
admitted into the resident join kernel while others run; every result equals
the query run alone on the plain path (counts and cap flags, bit-exact), a
late short query returns while an earlier long one still runs, and
latencies are reported."""
import random
import threading
import time

import numpy as np
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


def test_late_query_finishes_while_long_one_runs(ctx):
    """A long query is submitted first; short queries arrive while it runs and
    each returns (equal to its plain-path result) while the long one is still
    pending; the long one's result equals the plain path's too.  The data
    graph is R-MAT scale 16 (16 labels) plus ten triangles labelled 20, 21,
    22, which the short queries match."""
    hg = csm.gen_rmat(16, 1 << 20, 16, 1.0, 1)
    labels, edges = hg.arrays()
    n = len(labels)
    labels = np.concatenate([labels, np.array([20, 21, 22] * 10, np.uint32)])
    tri = [(n + 3 * t + a, n + 3 * t + b) for t in range(10) for a, b in ((0, 1), (1, 2), (0, 2))]
    edges = np.concatenate([edges, np.array(tri, np.uint32)])
    g = ctx.upload_edges(labels, edges)
    long_plan = csm.gen_query(hg, 8, csm.QUERY_EXTRA, 12, extra_edges=3).plan()
    cap = 400_000_000_000
    short_plans = [csm.HostGraph.from_edges(l, e).plan() for l, e in (
        ([20, 21, 22], [(0, 1), (1, 2), (0, 2)]), ([21, 20, 22], [(0, 1), (0, 2)]), ([22, 21], [(0, 1)]))]
    want = []
    for p in [long_plan] + short_plans:
        q = g.prepare(p)
        t = time.perf_counter()
        want.append(q.match(cap=cap)[0])
        print("alone", want[-1].count, want[-1].capped, f"{time.perf_counter() - t:.2f}s", flush=True)
        q.free()
    assert [w.count for w in want[1:]] == [10, 10, 10]
    ctx.serve_start(g, budget_bytes=8 << 30, max_inflight=16, cap=cap)
    try:
        tl = ctx.serve_submit(long_plan)
        time.sleep(0.05)  # the long query is on the device
        overlapped = 0
        for p, w in zip(short_plans, want[1:]):
            ts = ctx.serve_submit(p)
            res = ctx.serve_poll(ts, wait=True)
            pending_long = ctx.serve_poll(tl, wait=False) is None
            print("short", res.count, res.capped, f"latency {res.ms_latency:.1f} ms", "long pending", pending_long,
                  flush=True)
            assert res.status == 0 and res.anomalies == 0
            assert (res.count, res.capped) == (w.count, w.capped)
            overlapped += pending_long
        res = ctx.serve_poll(tl, wait=True)
        print("long", res.count, res.capped, f"latency {res.ms_latency:.1f} ms", flush=True)
        assert res.status == 0 and (res.count, res.capped) == (want[0].count, want[0].capped)
        assert overlapped == len(short_plans), overlapped
    finally:
        ctx.serve_stop()
    g.free()
