"""GPU parity of the plain path (device CSR build, filter, refinement, join)
against the CPU oracle and the committed goldens.  Bit-exact: counts, sorted
embedding lists, candidate bitmaps."""
import random

import numpy as np
import pytest

import clustersm_amd as csm
from helpers import (brute_force, example_paths, list_digest, load_example, load_golden, max_size_cases, rand_graph,
                     rand_query)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = csm.Context(0, 8 << 30)
    yield c
    c.close()


def _host_csr(n, edges):
    nbr = [set() for _ in range(n)]
    for a, b in edges:
        if a != b:
            nbr[a].add(b)
            nbr[b].add(a)
    off = np.zeros(n + 1, np.uint64)
    adj = []
    for v in range(n):
        s = sorted(nbr[v])
        adj += s
        off[v + 1] = off[v] + len(s)
    return off, np.array(adj, np.uint32)


def test_device_csr_build(ctx):
    r = random.Random(1)
    for it in range(20):
        n = r.randint(1, 300)
        m = r.randint(0, 3 * n)
        edges = [(r.randrange(n), r.randrange(n)) for _ in range(m)]  # loops and duplicates
        edges += edges[: m // 4]
        labels = [r.randrange(5) for _ in range(n)]
        print("csr case", it, n, len(edges), flush=True)
        g = ctx.upload_edges(labels, edges if edges else np.zeros((0, 2), np.uint32))
        off, adj, lab = g.download()
        eo, ea = _host_csr(n, edges)
        assert (off.astype(np.uint64) == eo).all(), it
        assert (adj == ea).all(), it
        assert (lab == np.array(labels, np.uint32)).all()
        nv, ne, md = g.info()
        assert ne * 2 == len(ea) and md == (int(np.diff(eo).max()) if n else 0)
        g.free()


def test_example_end_to_end(ctx):
    qp, gp = example_paths()
    ex = load_example()
    g = ctx.upload(csm.HostGraph.read(gp))
    q = g.prepare(csm.HostGraph.read(qp).plan())
    counts = q.filter(refine=False)
    assert counts.tolist() == [len(s) for s in ex["nlc_candidates"]]
    bits = q.candidates()
    assert [list(map(int, s)) for s in O.bits_to_sets(bits, 11)] == ex["nlc_candidates"]
    q.filter(refine=True)
    assert [list(map(int, s)) for s in O.bits_to_sets(q.candidates(), 11)] == ex["refined_candidates"]
    r, rows = q.match(want_list=True)
    assert r.count == 4 and not r.capped
    assert [tuple(x) for x in rows.tolist()] == [tuple(e) for e in ex["embeddings"]]
    r, _ = q.match()
    assert r.count == 4
    # levels 2..4 materialised: 3 + 3 + 7 rows (the non-tree check is fused into the last level)
    assert r.partial_rows == 13
    q.free()
    g.free()


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c["name"])
def test_golden_cases(ctx, case):
    g = ctx.upload_edges(case["data_labels"], case["data_edges"])
    q = g.prepare(csm.HostGraph.from_edges(case["query_labels"], case["query_edges"]).plan())
    for refine in (False, True):
        r, rows = q.match(want_list=True, list_capacity=1 << 16, refine=refine)
        assert r.count == case["count"] and not r.capped
        assert list_digest([tuple(x) for x in rows.tolist()]) == case["sha256"]
        r, _ = q.match(refine=refine)
        assert r.count == case["count"]
    q.free()
    g.free()


def _oracle_all(dl, de, ql, qe, cap=(1 << 64) - 1, list_cap=1 << 20):
    d = O.Data.from_arrays(dl, de)
    oq = O.Query.from_arrays(ql, qe)
    op = oq.plan()
    return d, oq, op, O.match(d, oq, op, cap=cap, want_list=True, list_cap=list_cap)


def test_filter_and_refine_bitmaps_vs_oracle(ctx):
    r = random.Random(2)
    for it in range(25):
        n = r.randint(50, 3000)
        nl = r.randint(1, 6)
        dl, de = rand_graph(r, min(n, 200), r.uniform(0.02, 0.2), nl) if it % 2 else (None, None)
        if dl is None:
            hg = csm.gen_erdos_renyi(n, min(4 * n, n * (n - 1) // 2), nl, it)
            dl, de = hg.arrays()
            dl, de = dl.tolist(), de.tolist()
        nv = len(dl)
        nq = r.randint(2, 10)
        ql, qe = rand_query(r, nq, r.randint(0, nq), nl)
        print("bitmap case", it, nv, len(de), nq, flush=True)
        d = O.Data.from_arrays(dl, de)
        oq = O.Query.from_arrays(ql, qe)
        op = oq.plan()
        g = ctx.upload_edges(dl, de)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        q.filter(refine=False)
        assert (q.candidates() == O.filter_nlc(d, nv, oq)).all(), it
        q.filter(refine=True)
        assert (q.candidates() == O.refine(d, nv, oq, op)).all(), it
        q.free()
        g.free()


def test_filter_labels_beyond_table_vs_oracle(ctx):
    """Query labels >= 64 are not covered by the neighbour-label table: the
    adjacency-scanning filter runs.  Both paths are checked in one graph."""
    r = random.Random(12)
    for it in range(8):
        n = r.randint(100, 400)
        nl = 90  # labels 0..89: some queries fall outside the table
        dl, de = rand_graph(r, n, r.uniform(0.05, 0.2), 1)
        dl = [r.choice((r.randint(0, 5), r.randint(60, 70))) for _ in dl]
        d = O.Data.from_arrays(dl, de)
        g = ctx.upload_edges(dl, de)
        for hi in (False, True):
            nq = r.randint(2, 6)
            ql, qe = rand_query(r, nq, r.randint(0, nq), 1)
            ql = [r.randint(60, 70) if hi else r.randint(0, 5) for _ in ql]
            print("label case", it, hi, n, nq, flush=True)
            oq = O.Query.from_arrays(ql, qe)
            q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
            q.filter(refine=False)
            assert (q.candidates() == O.filter_nlc(d, len(dl), oq)).all(), (it, hi)
            m, _ = q.match(cap=0)
            c, _, _, _, _ = O.match(d, oq, oq.plan(), cap=(1 << 64) - 1, threads=4)
            assert m.count == c, (it, hi)
            q.free()
        g.free()


def test_random_matches_vs_oracle_with_small_chunks(ctx):
    """Many chunks per level (DFS across chunks) must not change the result or its order."""
    r = random.Random(4)
    for it in range(40):
        n = r.randint(20, 120)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.05, 0.2), nl)
        nq = r.randint(3, 6)
        ql, qe = rand_query(r, nq, r.randint(0, 3), nl)
        d, oq, op, (oc, ocap, orows, _, _) = _oracle_all(dl, de, ql, qe, list_cap=200000)
        if oc > 200000:
            continue
        print("match case", it, n, len(de), nq, oc, flush=True)
        g = ctx.upload_edges(dl, de)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        maxdeg = g.info()[2]
        for chunk in (0, max(maxdeg, 256)):
            res, rows = q.match(want_list=True, list_capacity=200000, chunk_items=chunk)
            assert res.count == oc, (it, chunk)
            assert np.array_equal(rows, np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, nq))
            res, _ = q.match(chunk_items=chunk)
            assert res.count == oc
        q.free()
        g.free()


def test_caps_match_oracle(ctx):
    r = random.Random(6)
    done = 0
    for it in range(60):
        dl, de = rand_graph(r, 40, 0.25, 2)
        ql, qe = rand_query(r, 4, r.randint(0, 2), 2)
        d = O.Data.from_arrays(dl, de)
        oq = O.Query.from_arrays(ql, qe)
        op = oq.plan()
        total = O.match(d, oq, op)[0]
        if total < 10:
            continue
        g = ctx.upload_edges(dl, de)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        for cap in (1, total // 3, total - 1, total, total + 5):
            oc, ocap, orows, _, _ = O.match(d, oq, op, cap=cap, want_list=True)
            res, rows = q.match(cap=cap, want_list=True, chunk_items=256)
            assert (res.count, bool(res.capped)) == (oc, ocap), (it, cap)
            assert np.array_equal(rows, np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, 4))
            res, _ = q.match(cap=cap)
            assert (res.count, bool(res.capped)) == (oc, ocap)
        q.free()
        g.free()
        done += 1
    assert done > 5


def test_er_1000_4vertex_full_list(ctx):
    """ER 1000/4000 (seed 1), 4 uniform labels, 4-vertex random-walk query (seed 2), full list."""
    hg = csm.gen_erdos_renyi(1000, 4000, 4, 1)
    hq = csm.gen_query(hg, 4, csm.QUERY_INDUCED, 2)
    dl, de = hg.arrays()
    ql, qe = hq.arrays()
    d, oq, op, (oc, _, orows, _, _) = _oracle_all(dl, de, ql, qe)
    g = ctx.upload(hg)
    q = g.prepare(hq.plan())
    res, rows = q.match(want_list=True, list_capacity=max(oc, 1))
    print("er 1000 count", res.count, flush=True)
    assert res.count == oc
    assert np.array_equal(rows, np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, 4))
    q.free()
    g.free()


def test_rmat16_8vertex_first_100k_vs_oracle(ctx):
    """R-MAT scale 16, 16 Zipf labels: two 8-vertex queries, tree-like and with
    3 extra cycle edges.  Both have more than 1e10 embeddings, so the first
    100,000 in the matching order (listed, sorted) must be the oracle's, with
    the same capped flag.  The oracle builds its CSR from the generator's raw
    edge list, not from the device's."""
    hg = csm.gen_rmat(16, 1 << 20, 16, 1.0, 1)
    g = ctx.upload(hg)
    host = g.to_host()
    dl, de = hg.arrays()
    d = O.Data.from_arrays(dl, de)
    n_list = 100_000
    for mode, extra, seed in ((csm.QUERY_TREE, 0, 11), (csm.QUERY_EXTRA, 3, 12)):
        hq = csm.gen_query(host, 8, mode, seed, extra_edges=extra)
        ql, qe = hq.arrays()
        oq = O.Query.from_arrays(ql, qe)
        oc, ocap, orows, _, secs = O.match(d, oq, oq.plan(), cap=n_list, want_list=True, list_cap=n_list,
                                           refine=False)
        q = g.prepare(hq.plan())
        res, rows = q.match(cap=n_list, want_list=True)
        print("rmat16 8-vertex", mode, res.count, bool(res.capped), f"oracle {secs:.1f}s", flush=True)
        assert (res.count, bool(res.capped)) == (oc, ocap) == (n_list, True)
        want = np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, len(ql))
        assert np.array_equal(rows, want)
        q.free()
    g.free()


def test_hub_probes_and_fused_counts(ctx):
    """Graphs with hub vertices (dense-bitmap probes) and queries with and
    without non-tree edges at the last position (fused count vs count kernel)."""
    r = random.Random(8)
    n = 700
    edges = set()
    for h in range(6):
        for v in r.sample(range(6, n), 250 + 40 * h):
            edges.add((h, v))
    for _ in range(1500):
        a, b = r.sample(range(n), 2)
        edges.add((min(a, b), max(a, b)))
    de = list(edges)
    dl = [r.randrange(2) for _ in range(n)]
    d = O.Data.from_arrays(dl, de)
    g = ctx.upload_edges(dl, de)
    assert g.info()[2] >= 128
    seen_fused = seen_back = 0
    for it in range(30):
        nq = r.randint(2, 5)
        ql, qe = rand_query(r, nq, r.randint(0, 3), 2)
        oq = O.Query.from_arrays(ql, qe)
        op = oq.plan()
        oc, ocap, _, _, _ = O.match(d, oq, op, cap=50_000_000, threads=8)
        plan = csm.HostGraph.from_edges(ql, qe).plan()
        last = plan.num_v - 1
        if plan.num_back[last]:
            seen_back += 1
        else:
            seen_fused += 1
        q = g.prepare(plan)
        res, _ = q.match(cap=50_000_000)
        print("hub case", it, nq, len(qe), oc, flush=True)
        assert (res.count, bool(res.capped)) == (oc, ocap), it
        res, _ = q.match(cap=50_000_000, chunk_items=4096)
        assert (res.count, bool(res.capped)) == (oc, ocap), it
        res, _ = q.match(cap=50_000_000, lds_lists=True)
        assert (res.count, bool(res.capped)) == (oc, ocap), it
        q.free()
    assert seen_fused and seen_back
    g.free()


def test_lds_lists_vs_oracle(ctx):
    """The count level with its rows' short candidate lists copied into LDS
    (lds_lists) gives the oracle's counts: a dense graph without hubs (degrees
    near 48, two labels), so rows have many items and short non-hub lists,
    queries with non-tree edges and with fused last positions, capped and
    uncapped; and the default count on R-MAT 16's cyclic query."""
    r = random.Random(31)
    n = 1200
    edges = set()
    while len(edges) < n * 24:
        a, b = r.sample(range(n), 2)
        edges.add((min(a, b), max(a, b)))
    de = sorted(edges)
    dl = [r.randrange(2) for _ in range(n)]
    d = O.Data.from_arrays(dl, de)
    g = ctx.upload_edges(dl, de)
    assert g.info()[2] < 128  # no hubs: every anchor is searched
    for it in range(12):
        ql, qe = rand_query(r, r.randint(3, 5), r.randint(0, 3), 2)
        oq = O.Query.from_arrays(ql, qe)
        cap = 20_000_000 if it % 2 else 0
        oc, ocap, _, _, _ = O.match(d, oq, oq.plan(), cap=cap if cap else (1 << 64) - 1, threads=8)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        res, _ = q.match(cap=cap, lds_lists=True)
        print("lds lists", it, len(ql), len(qe), oc, bool(ocap), flush=True)
        assert (res.count, bool(res.capped)) == (oc, bool(ocap)), it
        res0, _ = q.match(cap=cap)
        assert (res0.count, bool(res0.capped)) == (oc, bool(ocap)), it
        q.free()
    g.free()
    hg = csm.gen_rmat(16, 1 << 20, 16, 1.0, 1)
    g = ctx.upload(hg)
    q = g.prepare(csm.gen_query(hg, 8, csm.QUERY_EXTRA, 12, extra_edges=3).plan())
    a, _ = q.match(cap=3_000_000_000)
    b, _ = q.match(cap=3_000_000_000, lds_lists=True)
    print("rmat16 cycles+3", a.count, b.count, a.items, b.items, flush=True)
    assert (a.count, a.capped, a.items) == (b.count, b.capped, b.items)
    q.free()
    g.free()


def test_reorder_keeps_counts_vs_oracle(ctx):
    """csm_query_reorder changes the matching order only: counts and cap
    flags stay the oracle's, on random graphs and queries with non-tree
    edges, and equal the default order's on R-MAT queries."""
    r = random.Random(77)
    for round_ in range(6):
        n = r.randint(40, 140)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.04, 0.12), nl)
        g = ctx.upload_edges(dl, de)
        d = O.Data.from_arrays(dl, de)
        for _ in range(6):
            ql, qe = rand_query(r, r.randint(3, 7), r.randint(0, 3), nl)
            oq = O.Query.from_arrays(ql, qe)
            q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
            q.reorder()
            for cap in (0, 50):
                want, wcap, _, _, _ = O.match(d, oq, oq.plan(), cap=cap if cap else (1 << 64) - 1, threads=4)
                res, _ = q.match(cap=cap)
                assert (res.count, bool(res.capped)) == (want, bool(wcap)), (round_, ql, qe, cap)
            q.free()
        g.free()
    hg = csm.gen_rmat(12, 40000, 8, 1.0, 9)
    g = ctx.upload(hg)
    host = g.to_host()
    for i in range(12):
        plan = csm.gen_query(host, 4 + i, csm.QUERY_EXTRA, 400 + i, extra_edges=i % 4).plan()
        a, b = g.prepare(plan), g.prepare(plan)
        b.reorder()
        ra, _ = a.match(cap=2_000_000)
        rb, _ = b.match(cap=2_000_000)
        print("reorder", i, ra.count, rb.count, f"{ra.ms_total:.1f} ms -> {rb.ms_total:.1f} ms", flush=True)
        assert (ra.count, ra.capped) == (rb.count, rb.capped), i
        a.free()
        b.free()
    g.free()


def test_symmetry_breaking_counts_vs_oracle(ctx):
    """With the symmetry option, count x |Aut(Q)| equals the oracle's full
    count: random graphs and queries (cyclic, few labels, so that many
    queries have automorphisms), the worked example and the golden cases."""
    r = random.Random(77)
    symmetric = 0
    cases = []
    for it in range(30):
        n = r.randint(30, 150)
        nl = r.randint(1, 2)
        dl, de = rand_graph(r, n, r.uniform(0.04, 0.2), nl)
        cases.append((dl, de) + rand_query(r, r.randint(3, 6), r.randint(0, 4), nl))
    for c in load_golden()[:20]:
        cases.append((c["data_labels"], c["data_edges"], c["query_labels"], c["query_edges"]))
    for dl, de, ql, qe in cases:
        oc, ocap, _, _, _ = O.match(O.Data.from_arrays(dl, de), O.Query.from_arrays(ql, qe),
                                    O.Query.from_arrays(ql, qe).plan(), threads=4)
        g = ctx.upload_edges(dl, de)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        res, _ = q.match(symmetry=True)
        assert res.status == 0 and not res.capped
        assert res.count * res.automorphisms == oc, (ql, qe, res.count, res.automorphisms, oc)
        symmetric += res.automorphisms > 1 and oc > 0
        q.free()
        g.free()
    print("symmetric queries with embeddings:", symmetric, flush=True)
    assert symmetric >= 5



def test_max_size_queries_closed_forms(ctx):
    """32-vertex queries (the largest a plan takes) on labelled cycles, whose
    counts follow from the cycles' rotations and reflections (listed in
    helpers.max_size_cases), on the plain path (count and list) and the
    batched path, each also against the oracle; the one-label 32-cycle has
    64 automorphisms, so the symmetry option counts 1."""
    for name, dl, de, ql, qe, want in max_size_cases():
        print("case", name, flush=True)
        oq = O.Query.from_arrays(ql, qe)
        oc, ocap, orows, _, _ = O.match(O.Data.from_arrays(dl, de), oq, oq.plan(), want_list=True)
        assert (oc, bool(ocap)) == (want, False), name
        g = ctx.upload_edges(dl, de)
        q = g.prepare(csm.HostGraph.from_edges(ql, qe).plan())
        res, rows = q.match(want_list=True, list_capacity=max(want, 1))
        assert (res.count, bool(res.capped), res.status) == (want, False, 0), name
        assert np.array_equal(rows, np.array(sorted(map(tuple, orows.tolist())), np.uint32).reshape(-1, 32)), name
        (b,) = csm.batch_match(g, [q])
        assert (b.count, bool(b.capped), b.status) == (want, False, 0), name
        if name == "C32 in C32, 1 label":
            s, _ = q.match(symmetry=True)
            assert (s.count, s.automorphisms) == (1, 64), name
        q.free()
        g.free()
