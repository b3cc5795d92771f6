"""The large cyclic batched workload at full size: a Chung-Lu power-law graph
of 16M vertices and 256M edge samples (degree exponent 2.1, maximum degree
50,000), 64 labels drawn Zipf(0.8), queries of 12 to 24 vertices with 2 to 6
non-tree edges, count-only with a 1e9 cap per query.

At this size the CPU oracle cannot enumerate the queries, so parity rests on
properties that hold at any size:
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
    off, adj, lab = g.download()
    for i, (q, p) in enumerate(zip(qs, plans)):
        res, rows = q.match(cap=2000, want_list=True, time_limit_ms=3000)
        if res.status != 0:
            continue
        rows = rows[: min(res.count, 2000)]
        print("list", i, len(rows), flush=True)
        assert len(rows) == min(res.count, 2000) and len(rows) > 0
        nv = p.num_v
        assert all((lab[rows[:, u]] == p.labels[u]).all() for u in range(nv)), i
        srt = np.sort(rows, axis=1)
        assert (srt[:, 1:] != srt[:, :-1]).all(), i
        for e in range(p.num_e):
            a_, b_ = rows[:, p.edge_src[e]], rows[:, p.edge_dst[e]]
            for x, y in zip(a_.tolist(), b_.tolist()):
                lo, hi = int(off[x]), int(off[x + 1])
                k = lo + int(np.searchsorted(adj[lo:hi], y))
                assert k < hi and adj[k] == y, (i, e, x, y)
        keys = [tuple(r) for r in rows.tolist()]
        assert keys == sorted(set(keys)), i
    for q in qs:
        q.free()
    g.free()
