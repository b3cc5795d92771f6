"""csm_match_noalloc: count-only matching with every workspace buffer in the
library's static device scratch.  Its results must equal csm_match's on two
calls in a row (the second with a smaller graph and query), a query whose
workspace does not fit must fail with CSM_E_NOMEM and a message naming the
limit, and the context must stay usable after that."""
import pytest

import clustersm_amd as csm

pytestmark = pytest.mark.gpu


def test_noalloc_equals_match_then_rejects_oversize():
    ctx = csm.Context(0, 8 << 30)
    try:
        # call 1: R-MAT scale 14, a 7-vertex query with 2 non-tree edges
        hg1 = csm.gen_rmat(14, 1 << 18, 8, 1.0, 21)
        g1 = ctx.upload(hg1)
        q1 = g1.prepare(csm.gen_query(hg1, 7, csm.QUERY_EXTRA, 22, extra_edges=2).plan())
        # call 2, smaller: the 1,000-vertex uniform graph, a 4-vertex query
        hg2 = csm.gen_erdos_renyi(1000, 4000, 4, 1)
        g2 = ctx.upload(hg2)
        q2 = g2.prepare(csm.gen_query(hg2, 4, csm.QUERY_INDUCED, 2).plan())
        for caps in ((100_000_000, 0), (5000, 50)):
            for g, q, cap in ((g1, q1, caps[0]), (g2, q2, caps[1])):
                print("noalloc vs match, cap", cap, "graph", g.info(), flush=True)
                want, _ = q.match(cap=cap)
                got = q.match_noalloc(cap=cap)
                assert want.status == 0 and got.status == 0
                assert (got.count, got.capped) == (want.count, want.capped), (cap, got.count, want.count)
        # oversize: complete bipartite 2000 x 2000 and a 10-vertex path over
        # its two labels: nine candidate-list slots of 2000 x 2000 ids each
        # (144 MiB) cannot fit the 128 MiB scratch
        n = 4000
        labels = [0] * 2000 + [1] * 2000
        edges = [(a, 2000 + b) for a in range(2000) for b in range(2000)]
        g3 = ctx.upload_edges(labels, edges)
        path = csm.HostGraph.from_edges([i % 2 for i in range(10)], [(i, i + 1) for i in range(9)])
        q3 = g3.prepare(path.plan())
        print("oversize path query on", g3.info(), flush=True)
        with pytest.raises(csm.CsmError) as e:
            q3.match_noalloc(cap=1000)
        assert e.value.status == csm.E_NOMEM and "static scratch" in str(e.value), str(e.value)
        # still usable, and still equal to csm_match
        want, _ = q2.match(cap=0)
        got = q2.match_noalloc(cap=0)
        assert (got.count, got.capped) == (want.count, want.capped)
        assert n == g3.info()[0]
        for q in (q1, q2, q3):
            q.free()
        for g in (g1, g2, g3):
            g.free()
    finally:
        ctx.close()
