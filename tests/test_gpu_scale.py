"""The batched and served workloads at full size, against the CPU oracle:
* R-MAT scale 20 (2^24 edge samples), 32 labels Zipf(1.0), random-walk
  queries of 4..32 vertices, cap 1e8;
* a Chung-Lu power-law graph of 16M vertices and 256M edge samples (degree
  exponent 2.1, maximum degree 50,000), 64 labels drawn Zipf(0.8), queries of
  12 to 24 vertices with 2 to 6 non-tree edges, count-only with a 1e9 cap per
  query;
* R-MAT scale 22 (2^26 edge samples), 32 labels Zipf(1.0), random-walk
  queries of 4..32 vertices (geometric, mean 10), cap 1e8.

Parity with the oracle, whose CSR is built from the generator's raw edge
list (not from the device's):
* every query's first 100,000 embeddings in the matching order, listed by
  the plain path (sorted, query-id order), equal the oracle's, with the same
  count and capped flag; a query with fewer embeddings is compared in full;
* a query that stays under the workload's cap on the plain path has its
  exact count compared with the oracle's (16 threads, bounded time);
* the batched path (one persistent kernel over all the queries) gives the
  plain path's (count, capped) at the workload's cap, with no query recounted
  on the plain path and no late or bad task record, and a query alone in a
  batch gives what it gives in the mix.
The oracle runs with a time budget per query; a query it cannot finish is
reported and left out, and each test requires a floor of compared queries."""
import random
import time

import numpy as np
import pytest

import clustersm_amd as csm
from oracle import oracle as O

pytestmark = pytest.mark.gpu

N_LIST = 100_000


@pytest.fixture(scope="module")
def ctx():
    c = csm.Context(0, 96 << 30)
    yield c
    c.close()


def oracle_data(hg):
    t0 = time.time()
    labels, edges = hg.arrays()
    d = O.Data.from_arrays(labels, edges)
    print(f"oracle CSR in {time.time() - t0:.1f}s", flush=True)
    return d


def prefix_vs_oracle(d, items, gpu_limit_ms=5000, oracle_budget_s=20.0):
    """items: (index, device query, host query).  Returns (compared, uncapped)."""
    compared = uncapped = 0
    for i, q, hq in items:
        res, rows = q.match(cap=N_LIST, want_list=True, time_limit_ms=gpu_limit_ms)
        if res.status != 0:
            print("prefix", i, "plain path over its time limit", flush=True)
            continue
        ql, qe = hq.arrays()
        oq = O.Query.from_arrays(ql, qe)
        c, capped, orows, timed_out, secs = O.match(d, oq, oq.plan(), cap=N_LIST, want_list=True, list_cap=N_LIST,
                                                    budget_s=oracle_budget_s, refine=False)
        print("prefix", i, res.count, bool(res.capped), "oracle", c, capped, f"{secs:.1f}s",
              "timed out" if timed_out else "", flush=True)
        if timed_out:
            continue
        assert (res.count, bool(res.capped)) == (c, capped), i
        want = np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, len(ql))
        assert np.array_equal(rows, want), i
        compared += 1
        uncapped += not capped
    return compared, uncapped


def counts_vs_oracle(d, items, plain, cap, budget_s=20.0):
    """Exact counts of the queries the plain path finds under `cap`."""
    compared = 0
    for (i, q, hq), p in zip(items, plain):
        if p.status != 0 or p.capped:
            continue
        ql, qe = hq.arrays()
        oq = O.Query.from_arrays(ql, qe)
        c, capped, _, timed_out, secs = O.match(d, oq, oq.plan(), cap=cap, threads=16, budget_s=budget_s,
                                                refine=False)
        print("count", i, p.count, "oracle", c, capped, f"{secs:.1f}s", "timed out" if timed_out else "",
              flush=True)
        if timed_out:
            continue
        assert (p.count, False) == (c, capped), i
        compared += 1
    return compared


def batch_equals_plain(g, qs, cap, plain_limit_ms):
    plain = []
    for i, q in enumerate(qs):
        res, _ = q.match(cap=cap, time_limit_ms=plain_limit_ms)
        print("plain", i, res.status, res.count, res.capped, f"{res.ms_total:.0f} ms", flush=True)
        plain.append(res)
    batch = csm.batch_match(g, qs, cap=cap, time_limit_ms=30000)
    assert all(x.path != csm.PATH_BATCH_ABORT and x.anomalies == 0 for x in batch), \
        [(x.path, x.anomalies) for x in batch]
    both = [i for i in range(len(qs)) if plain[i].status == 0 and batch[i].status == 0]
    for i in both:
        assert (batch[i].count, batch[i].capped) == (plain[i].count, plain[i].capped), i
    print("batch == plain on", len(both), "of", len(qs), flush=True)
    return plain, batch, both


def test_rmat20_vs_oracle(ctx):
    hg = csm.gen_rmat(20, 1 << 24, 32, 1.0, 3)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(5)
    hqs = []
    for i in range(32):
        nv, avg = r.randint(4, 32), r.uniform(2, 4)
        hqs.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 1000 + i, avg_degree=avg))
    qs = [g.prepare(h.plan()) for h in hqs]
    items = list(zip(range(32), qs, hqs))
    plain, _, both = batch_equals_plain(g, qs, 100_000_000, 2000)
    assert len(both) >= 28
    d = oracle_data(hg)
    compared, uncapped = prefix_vs_oracle(d, items)
    exact = counts_vs_oracle(d, items, plain, 100_000_000, budget_s=90.0)
    print("rmat20: prefixes compared", compared, "uncapped", uncapped, "exact counts", exact, flush=True)
    assert compared >= 24 and exact >= 1
    for q in qs:
        q.free()
    g.free()


def test_rmat22_vs_oracle(ctx):
    hg = csm.gen_rmat(22, 1 << 26, 32, 1.0, 11)
    g = ctx.upload(hg)
    host = g.to_host()
    r = random.Random(7)
    hqs = []
    for i in range(16):
        nv = 4
        while nv < 32 and r.random() > 1.0 / 7.0:
            nv += 1
        hqs.append(csm.gen_query(host, nv, csm.QUERY_AVG_DEGREE, 7000 + i, avg_degree=r.uniform(2, 4)))
    qs = [g.prepare(h.plan()) for h in hqs]
    items = list(zip(range(16), qs, hqs))
    plain, _, both = batch_equals_plain(g, qs, 100_000_000, 2000)
    assert len(both) >= 12
    d = oracle_data(hg)
    compared, uncapped = prefix_vs_oracle(d, items)
    exact = counts_vs_oracle(d, items, plain, 100_000_000, budget_s=60.0)
    print("rmat22: prefixes compared", compared, "uncapped", uncapped, "exact counts", exact, flush=True)
    assert compared >= 12
    for q in qs:
        q.free()
    g.free()


def test_chung_lu_16m_vs_oracle(ctx):
    t0 = time.time()
    hg = csm.gen_chung_lu(16_000_000, 256_000_000, 2.1, 50_000, 64, 0.8, 41)
    g = ctx.upload(hg)
    n, m, maxdeg = g.info()
    print(f"graph |V|={n} |E|={m} max degree {maxdeg} in {time.time() - t0:.1f}s", flush=True)
    assert n == 16_000_000 and m > 200_000_000 and maxdeg <= 50_000
    host = g.to_host()
    r = random.Random(4)
    hqs = []
    for i in range(8):
        nv, extra = r.randint(12, 24), r.randint(2, 6)
        hqs.append(csm.gen_query(host, nv, csm.QUERY_EXTRA, 300 + i, extra_edges=extra))
        print("query", i, "vertices", nv, "extra edges", extra, flush=True)
    del host
    qs = [g.prepare(h.plan()) for h in hqs]
    cap = 1_000_000_000
    plain, batch, both = batch_equals_plain(g, qs, cap, 3000)
    assert len(both) >= len(qs) // 2
    j = both[0]
    alone = csm.batch_match(g, [qs[j]], cap=cap, time_limit_ms=30000)[0]
    print("alone", j, alone.status, alone.count, alone.capped, flush=True)
    assert alone.status == 0 and (alone.count, alone.capped) == (batch[j].count, batch[j].capped)
    d = oracle_data(hg)
    del hg
    compared, uncapped = prefix_vs_oracle(d, list(zip(range(8), qs, hqs)), oracle_budget_s=30.0)
    print("chung-lu: prefixes compared", compared, "uncapped", uncapped, flush=True)
    assert compared >= 6
    for q in qs:
        q.free()
    g.free()
