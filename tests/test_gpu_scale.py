"""The batched and served workloads at full size:
* R-MAT scale 20 (2^24 edge samples), 32 labels Zipf(1.0), random-walk
  queries of 4..32 vertices, cap 1e8;
* a Chung-Lu power-law graph of 16M vertices and 256M edge samples (degree
  exponent 2.1, maximum degree 50,000), 64 labels drawn Zipf(0.8), queries of
  12 to 24 vertices with 2 to 6 non-tree edges, count-only with a 1e9 cap per
  query;
* R-MAT scale 22 (2^26 edge samples), 32 labels Zipf(1.0), random-walk
  queries of 4..32 vertices (geometric, mean 10), cap 1e8.

At these sizes the CPU oracle cannot enumerate the queries, so parity rests
on properties that hold at any size:
* the plain per-query path and one batch call over all the queries give
  identical (count, capped) for every query both finish, and a query run
  alone in a batch gives the same result as in the mix;
* every embedding the plain path lists (first 2000 per query) is one: labels
  equal, ids distinct, every query edge an edge of the device CSR, rows
  distinct and sorted.
The plain path itself is checked against the oracle on the smaller
workloads."""
import random
import time

import numpy as np
import pytest

import clustersm_amd as csm

pytestmark = pytest.mark.gpu

CAP = 1_000_000_000


@pytest.fixture(scope="module")
def ctx():
    c = csm.Context(0, 96 << 30)
    yield c
    c.close()


def check_listed(g, qs, plans, n_list=2000):
    """The plain path's first n_list embeddings of each query are embeddings
    of the device CSR: labels equal, ids distinct, every query edge present,
    rows distinct and sorted."""
    off, adj, lab = g.download()
    checked = 0
    for i, (q, p) in enumerate(zip(qs, plans)):
        res, rows = q.match(cap=n_list, want_list=True, time_limit_ms=3000)
        if res.status != 0 or res.count == 0:
            continue
        rows = rows[: min(res.count, n_list)]
        print("list", i, len(rows), flush=True)
        assert len(rows) == min(res.count, n_list)
        assert all((lab[rows[:, u]] == p.labels[u]).all() for u in range(p.num_v)), i
        srt = np.sort(rows, axis=1)
        assert (srt[:, 1:] != srt[:, :-1]).all(), i
        for e in range(p.num_e):
            for x, y in zip(rows[:, p.edge_src[e]].tolist(), rows[:, p.edge_dst[e]].tolist()):
                lo, hi = int(off[x]), int(off[x + 1])
                k = lo + int(np.searchsorted(adj[lo:hi], y))
                assert k < hi and adj[k] == y, (i, e, x, y)
        keys = [tuple(r) for r in rows.tolist()]
        assert keys == sorted(set(keys)), i
        checked += 1
    return checked


def batch_equals_plain(g, qs, cap, plain_limit_ms):
    plain = []
    for i, q in enumerate(qs):
        res, _ = q.match(cap=cap, time_limit_ms=plain_limit_ms)
        print("plain", i, res.status, res.count, res.capped, f"{res.ms_total:.0f} ms", flush=True)
        plain.append(res)
    batch = csm.batch_match(g, qs, cap=cap, time_limit_ms=30000)
    both = [i for i in range(len(qs)) if plain[i].status == 0 and batch[i].status == 0]
    for i in both:
        assert (batch[i].count, batch[i].capped) == (plain[i].count, plain[i].capped), i
    print("compared", len(both), "of", len(qs), "; uncapped", sum(1 for i in both if not plain[i].capped), flush=True)
    return plain, batch, both


def test_rmat20_batch_equals_plain(ctx):
    hg = csm.gen_rmat(20, 1 << 24, 32, 1.0, 3)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(5)
    plans = []
    for i in range(32):
        nv, avg = r.randint(4, 32), r.uniform(2, 4)
        plans.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 1000 + i, avg_degree=avg).plan())
    qs = [g.prepare(p) for p in plans]
    _, _, both = batch_equals_plain(g, qs, 100_000_000, 2000)
    assert len(both) >= 28
    assert check_listed(g, qs[:8], plans[:8]) >= 4
    for q in qs:
        q.free()
    g.free()


def test_rmat22_batch_equals_plain(ctx):
    hg = csm.gen_rmat(22, 1 << 26, 32, 1.0, 11)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(7)
    plans = []
    for i in range(16):
        nv = 4
        while nv < 32 and r.random() > 1.0 / 7.0:
            nv += 1
        plans.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 7000 + i, avg_degree=r.uniform(2, 4)).plan())
    qs = [g.prepare(p) for p in plans]
    _, _, both = batch_equals_plain(g, qs, 100_000_000, 2000)
    assert len(both) >= 12
    assert check_listed(g, qs[:6], plans[:6]) >= 3
    for q in qs:
        q.free()
    g.free()


def test_chung_lu_16m_batch_equals_plain(ctx):
    t0 = time.time()
    hg = csm.gen_chung_lu(16_000_000, 256_000_000, 2.1, 50_000, 64, 0.8, 41)
    g = ctx.upload(hg)
    del hg
    n, m, maxdeg = g.info()
    print(f"graph |V|={n} |E|={m} max degree {maxdeg} in {time.time() - t0:.1f}s", flush=True)
    assert n == 16_000_000 and m > 200_000_000 and maxdeg <= 50_000
    host = g.to_host()
    r = random.Random(4)
    plans = []
    for i in range(8):
        nv, extra = r.randint(12, 24), r.randint(2, 6)
        plans.append(csm.gen_query(host, nv, csm.QUERY_EXTRA, 300 + i, extra_edges=extra).plan())
        print("query", i, "vertices", nv, "extra edges", extra, flush=True)
    del host
    qs = [g.prepare(p) for p in plans]
    plain = []
    for i, q in enumerate(qs):
        res, _ = q.match(cap=CAP, time_limit_ms=3000)
        print("plain", i, res.status, res.count, res.capped, f"{res.ms_total:.0f} ms", flush=True)
        plain.append(res)
    batch = csm.batch_match(g, qs, cap=CAP, time_limit_ms=30000)
    for i, x in enumerate(batch):
        print("batch", i, x.status, x.count, x.capped, x.path, flush=True)
    both = [i for i in range(len(qs)) if plain[i].status == 0 and batch[i].status == 0]
    assert len(both) >= len(qs) // 2, (both, [p.status for p in plain], [x.status for x in batch])
    for i in both:
        assert (batch[i].count, batch[i].capped) == (plain[i].count, plain[i].capped), i
    # alone in a batch: the same as in the mix
    j = both[0]
    alone = csm.batch_match(g, [qs[j]], cap=CAP, time_limit_ms=30000)[0]
    print("alone", j, alone.status, alone.count, alone.capped, flush=True)
    assert alone.status == 0 and (alone.count, alone.capped) == (batch[j].count, batch[j].capped)
    # listed embeddings are embeddings
    assert check_listed(g, qs, plans) == len(qs)
    for q in qs:
        q.free()
    g.free()
