"""Failure paths forced from the host: a small arena makes csm_match,
csm_batch_match and a served request report CSM_E_NOMEM for a query whose
candidate lists do not fit, and the same context then answers a small query
exactly (closed form).  Data graph: the complete bipartite graph K(2000,
2000), labels 0 and 1 by side; the large query is a 32-vertex path over the
two labels, whose 31 candidate-list slots need 31 x 2000 x 2000 ids (about
496 MB) in a 512 MiB arena that the graph's CSR (built in about 260 MB of
temporaries) already partly holds."""
import numpy as np
import pytest

import clustersm_amd as csm

pytestmark = pytest.mark.gpu

SIDE = 2000
SMALL_COUNT = SIDE * SIDE * (SIDE - 1)  # path 0-1-0: a label-1 middle, an ordered pair of distinct label-0 ends


def test_arena_exhaustion_is_reported_and_the_context_stays_usable():
    ctx = csm.Context(0, 512 << 20)
    try:
        labels = np.array([0] * SIDE + [1] * SIDE, np.uint32)
        a = np.repeat(np.arange(SIDE, dtype=np.uint32), SIDE)
        b = SIDE + np.tile(np.arange(SIDE, dtype=np.uint32), SIDE)
        g = ctx.upload_edges(labels, np.stack([a, b], axis=1))
        big = csm.HostGraph.from_edges([i % 2 for i in range(32)], [(i, i + 1) for i in range(31)]).plan()
        small = csm.HostGraph.from_edges([0, 1, 0], [(0, 1), (1, 2)]).plan()
        qb, qs = g.prepare(big), g.prepare(small)
        print("plain path, large query", flush=True)
        with pytest.raises(csm.CsmError) as e:
            qb.match(cap=1000)
        assert e.value.status == csm.E_NOMEM, str(e.value)
        r, _ = qs.match()
        assert (r.count, r.capped) == (SMALL_COUNT, 0)
        print("batched path, large query", flush=True)
        with pytest.raises(csm.CsmError) as e:
            csm.batch_match(g, [qb, qs], cap=1000)
        assert e.value.status == csm.E_NOMEM, str(e.value)
        res = csm.batch_match(g, [qs])
        assert (res[0].count, res[0].capped) == (SMALL_COUNT, 0)
        print("served, large query then a small one", flush=True)
        ctx.serve_start(g, budget_bytes=64 << 20, max_inflight=4)
        try:
            rb = ctx.serve_poll(ctx.serve_submit(big))
            assert rb.status == csm.E_NOMEM, rb.status
            rs = ctx.serve_poll(ctx.serve_submit(small))
            assert rs.status == 0 and (rs.count, rs.capped) == (SMALL_COUNT, 0)
        finally:
            ctx.serve_stop()
        r, _ = qs.match()
        assert r.count == SMALL_COUNT
        qb.free()
        qs.free()
        g.free()
    finally:
        ctx.close()
