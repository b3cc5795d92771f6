"""Batched path (one persistent kernel for many queries): counts and cap
flags identical to the CPU oracle and to the plain per-query path, alone and
in the mix, with and without graph capture, and under pool pressure."""
import random

import numpy as np
import pytest

import clustersm_amd as csm
from helpers import rand_graph, rand_query
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csm.Context(0, 16 << 30)
    yield c
    c.close()


def _oracle_counts(dl, de, queries, cap):
    d = O.Data.from_arrays(dl, de)
    out = []
    for ql, qe in queries:
        oq = O.Query.from_arrays(ql, qe)
        c, capped, _, _, _ = O.match(d, oq, oq.plan(), cap=cap if cap else (1 << 64) - 1, threads=8)
        out.append((c, capped))
    return out


def test_batch_small_random_vs_oracle(ctx):
    r = random.Random(31)
    for round_ in range(6):
        n = r.randint(40, 160)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.03, 0.15), nl)
        queries = [rand_query(r, r.randint(2, 7), r.randint(0, 3), nl) for _ in range(r.randint(1, 24))]
        g = ctx.upload_edges(dl, de)
        qs = [g.prepare(csm.HostGraph.from_edges(ql, qe).plan()) for ql, qe in queries]
        for cap in (0, 50):
            want = _oracle_counts(dl, de, queries, cap)
            for use_graph in (True, False):
                print("batch round", round_, "queries", len(qs), "cap", cap, "graph", use_graph, flush=True)
                # a short time limit turns a scheduling fault into a reported abort
                res = csm.batch_match(g, qs, cap=cap, use_graph=use_graph, time_limit_ms=20000)
                assert all(x.path in (csm.PATH_BATCH, csm.PATH_HOST) and x.anomalies == 0 for x in res), \
                    [(x.path, x.anomalies) for x in res]
                got = [(x.count, bool(x.capped)) for x in res]
                bad = [(i, queries[i], got[i], want[i]) for i in range(len(got)) if got[i] != want[i]]
                assert not bad, (round_, cap, use_graph, bad[:3])
        for q in qs:
            q.free()
        g.free()


def test_batch_equals_plain_and_alone(ctx):
    """Each query's result is the same alone (plain path, batch of one) and in the mix."""
    hg = csm.gen_rmat(12, 40000, 8, 1.0, 9)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(3)
    plans = []
    for i in range(24):
        nv = r.randint(3, 12)
        plans.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 500 + i, avg_degree=r.uniform(2, 3.5)).plan())
    qs = [g.prepare(p) for p in plans]
    cap = 200_000
    mix = csm.batch_match(g, qs, cap=cap, time_limit_ms=20000)
    assert all(x.path in (csm.PATH_BATCH, csm.PATH_HOST) and x.anomalies == 0 for x in mix)
    for i, q in enumerate(qs):
        plain, _ = q.match(cap=cap)
        alone = csm.batch_match(g, [q], cap=cap, time_limit_ms=20000)[0]
        print("query", i, plans[i].num_v, plain.count, mix[i].count, flush=True)
        assert (mix[i].count, mix[i].capped) == (plain.count, plain.capped) == (alone.count, alone.capped), i
    for q in qs:
        q.free()
    g.free()


def test_batch_small_pool_continuations(ctx):
    """A pool of a few blocks forces yields and, if it runs dry, the plain fallback.
    Tuning variants (block size, output bound, lanes per item) give the same counts."""
    hg = csm.gen_rmat(11, 30000, 4, 0.5, 4)
    g = ctx.upload(hg)
    host = g.to_host()
    plans = [csm.gen_query(host, 6, csm.QUERY_TREE, 77 + i).plan() for i in range(6)]
    qs = [g.prepare(p) for p in plans]
    cap = 3_000_000
    want = [q.match(cap=cap)[0] for q in qs]
    for pool in (64 * 128 * 1024, 256 * 128 * 1024):
        res = csm.batch_match(g, qs, cap=cap, pool_bytes=pool, time_limit_ms=20000)
        assert [(x.count, x.capped) for x in res] == [(w.count, w.capped) for w in want], pool
    for q in qs:
        q.free()
    g.free()


def test_batch_tuning_variants_agree(ctx):
    hg = csm.gen_rmat(12, 40000, 6, 1.0, 21)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(8)
    plans = [csm.gen_query(host, r.randint(3, 14), csm.QUERY_AVG_DEGREE, 900 + i, avg_degree=r.uniform(2, 3.5)).plan()
             for i in range(16)]
    qs = [g.prepare(p) for p in plans]
    cap = 500_000
    want = [(x.count, x.capped) for x in (q.match(cap=cap)[0] for q in qs)]
    for kw in (dict(wgs_per_cu=1), dict(wgs_per_cu=3, max_out=1), dict(block_kb=32, ilp=1), dict(max_out=8, ilp=2),
               dict(chunk_rows=64)):
        print("tuning", kw, flush=True)
        res = csm.batch_match(g, qs, cap=cap, time_limit_ms=20000, **kw)
        assert [(x.count, x.capped) for x in res] == want, kw
        assert all(x.path in (csm.PATH_BATCH, csm.PATH_HOST) and x.anomalies == 0 for x in res), kw
    # an item budget stops a query with a lower bound; queries within it are exact
    res = csm.batch_match(g, qs, cap=cap, time_limit_ms=20000, item_budget=1000)
    assert any(x.status == csm.E_BUDGET for x in res)
    for x, w in zip(res, want):
        if x.status == 0:
            assert (x.count, x.capped) == w
        else:
            assert x.status == csm.E_BUDGET and x.count <= w[0] and not x.capped
    plain = [q.match(cap=cap, item_budget=1000)[0] for q in qs]
    assert any(p.status == csm.E_BUDGET for p in plain)
    for p, w in zip(plain, want):
        assert (p.count, p.capped) == w if p.status == 0 else (p.status == csm.E_BUDGET and p.count <= w[0])
    for q in qs:
        q.free()
    g.free()


def test_time_limits_report_timeouts(ctx):
    """A query stopped by its time limit reports E_TIMEOUT and a count that is a
    lower bound; every query finished within the limit is exact (plain and batch)."""
    hg = csm.gen_rmat(14, 200000, 4, 1.0, 5)
    g = ctx.upload(hg)
    host = g.to_host()
    plans = [csm.gen_query(host, 10, csm.QUERY_TREE, 40 + i).plan() for i in range(8)]
    qs = [g.prepare(p) for p in plans]
    cap = 20_000_000
    want = [(x.count, x.capped) for x in (q.match(cap=cap)[0] for q in qs)]
    timeouts = 0
    for q, w in zip(qs, want):
        r, _ = q.match(cap=cap, time_limit_ms=1)
        if r.status == csm.E_TIMEOUT:
            timeouts += 1
            assert r.count <= w[0] and not r.capped
        else:
            assert r.status == 0 and (r.count, r.capped) == w
    for limit in (1, 30, 20000):
        res = csm.batch_match(g, qs, cap=cap, time_limit_ms=limit)
        for x, w in zip(res, want):
            if x.status == csm.E_TIMEOUT:
                timeouts += 1
                assert x.count <= w[0] and not x.capped, limit
            else:
                assert x.status == 0 and (x.count, x.capped) == w, limit
        if limit == 20000:
            assert all(x.status == 0 for x in res)
    print("timeouts seen", timeouts, flush=True)
    for q in qs:
        q.free()
    g.free()


def test_autotune_picks_a_valid_setting(ctx):
    hg = csm.gen_rmat(12, 40000, 6, 1.0, 33)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(44)
    plans = [csm.gen_query(host, r.randint(3, 12), csm.QUERY_AVG_DEGREE, 700 + i, avg_degree=r.uniform(2, 3.5)).plan()
             for i in range(20)]
    qs = [g.prepare(p) for p in plans]
    cap = 300_000
    tuned = csm.batch_autotune(g, qs, cap=cap, sample=8, time_limit_ms=20000)
    assert 1 <= tuned.wgs_per_cu <= 8 and 1 <= tuned.max_out <= 8 and 32 <= tuned.block_kb <= 4096
    assert 1 <= tuned.ilp <= 4 and tuned.cap == cap
    res = csm.batch_match(g, qs, opts=tuned)
    want = [(x.count, x.capped) for x in (q.match(cap=cap)[0] for q in qs)]
    assert [(x.count, x.capped) for x in res] == want
    for q in qs:
        q.free()
    g.free()


def test_reorder_option_keeps_counts(ctx):
    """The reorder option (candidate-aware order from each path's own filter
    counts) gives the default order's counts and cap flags, batched and plain."""
    hg = csm.gen_rmat(12, 40000, 8, 1.0, 9)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(12)
    plans = [csm.gen_query(host, r.randint(3, 14), csm.QUERY_AVG_DEGREE, 1200 + i, avg_degree=r.uniform(2, 3.5)).plan()
             for i in range(16)]
    cap = 500_000
    qs = [g.prepare(p) for p in plans]
    want = [(x.count, x.capped) for x in (q.match(cap=cap)[0] for q in qs)]
    qs2 = [g.prepare(p) for p in plans]
    res = csm.batch_match(g, qs2, cap=cap, reorder=True, time_limit_ms=20000)
    assert [(x.count, x.capped) for x in res] == want
    assert [(x.count, x.capped) for x in (q.match(cap=cap, reorder=True)[0] for q in qs2)] == want
    for q in qs + qs2:
        q.free()
    g.free()
