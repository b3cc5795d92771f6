"""Pins the CPU oracle (oracle/csm_oracle.c) before anything is checked
against it: the reference's worked example, networkx goldens, brute force,
closed-form invariants, and the reference's own decycle.cpp (oracle/_ref)."""
import itertools
import random

import numpy as np
import pytest

from helpers import (brute_force, example_paths, load_example, load_golden, max_size_cases, rand_graph, rand_query,
                     list_digest)
from oracle import oracle as O


def test_example_trace():
    qp, gp = example_paths()
    ex = load_example()
    q = O.Query.read(qp)
    p = q.plan()
    assert p.p.tree_root == ex["decycle_tree_root"]
    assert p.cover == ex["vertex_cover"]
    assert p.clusters == ex["clusters"]
    assert p.compact == ex["compact"]
    assert p.cluster_offsets == ex["cluster_offsets"]
    assert p.root == ex["root"]
    assert p.order == [tuple(e) for e in ex["edge_order"]]
    d = O.Data.read(gp)
    nlc = [list(map(int, s)) for s in O.bits_to_sets(O.filter_nlc(d, 11, q), 11)]
    assert nlc == ex["nlc_candidates"]
    ref = [list(map(int, s)) for s in O.bits_to_sets(O.refine(d, 11, q, p), 11)]
    assert ref == ex["refined_candidates"]
    for refine in (False, True):
        c, capped, rows, _, _ = O.match(d, q, p, want_list=True, refine=refine)
        assert c == ex["count"] and not capped
        assert sorted(map(tuple, rows.tolist())) == [tuple(e) for e in ex["embeddings"]]


@pytest.mark.parametrize("case", load_golden(), ids=lambda c: c["name"])
def test_oracle_vs_networkx(case):
    q = O.Query.from_arrays(case["query_labels"], case["query_edges"])
    p = q.plan()
    d = O.Data.from_arrays(case["data_labels"], case["data_edges"])
    for refine in (False, True):
        c, capped, rows, _, _ = O.match(d, q, p, want_list=True, list_cap=1 << 20, refine=refine)
        assert c == case["count"] and not capped
        srt = sorted(map(tuple, rows.tolist()))
        assert list_digest(srt) == case["sha256"]
        if "embeddings" in case:
            assert srt == [tuple(e) for e in case["embeddings"]]


def test_oracle_vs_brute_force_random():
    r = random.Random(7)
    for it in range(150):
        n = r.randint(4, 11)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.15, 0.7), nl)
        nq = r.randint(2, min(5, n))
        ql, qe = rand_query(r, nq, r.randint(0, 4), nl)
        want = brute_force(dl, de, ql, qe)
        q = O.Query.from_arrays(ql, qe)
        p = q.plan()
        d = O.Data.from_arrays(dl, de)
        c, capped, rows, _, _ = O.match(d, q, p, want_list=True, list_cap=1 << 16)
        assert c == len(want), it
        assert sorted(map(tuple, rows.tolist())) == want
        c4, _, _, _, _ = O.match(d, q, p, threads=4)
        assert c4 == len(want)


def test_cap_is_lexicographic_prefix_in_matching_order():
    """Capped lists are the first `cap` embeddings of the matching-order tuple order."""
    r = random.Random(11)
    checked = 0
    for it in range(80):
        dl, de = rand_graph(r, 10, 0.5, 2)
        ql, qe = rand_query(r, 4, 1, 2)
        want = brute_force(dl, de, ql, qe)
        if len(want) < 4:
            continue
        q = O.Query.from_arrays(ql, qe)
        p = q.plan()
        d = O.Data.from_arrays(dl, de)
        mo = p.match_order
        by_match = sorted(want, key=lambda t: tuple(t[u] for u in mo))
        cap = len(want) // 2
        c, capped, rows, _, _ = O.match(d, q, p, cap=cap, want_list=True)
        assert c == cap and capped
        assert sorted(map(tuple, rows.tolist())) == sorted(by_match[:cap])
        c, capped, _, _, _ = O.match(d, q, p, cap=len(want))
        assert c == len(want) and not capped
        checked += 1
    assert checked > 10


def test_filter_predicate_brute_force():
    """NLC predicate restated per (u, v) in Python (filter.cu:486-507)."""
    r = random.Random(3)
    for it in range(60):
        n = r.randint(5, 30)
        nl = r.randint(1, 4)
        dl, de = rand_graph(r, n, r.uniform(0.1, 0.5), nl)
        nq = r.randint(2, 6)
        ql, qe = rand_query(r, nq, r.randint(0, 3), nl)
        q = O.Query.from_arrays(ql, qe)
        d = O.Data.from_arrays(dl, de)
        got = [set(map(int, s)) for s in O.bits_to_sets(O.filter_nlc(d, n, q), n)]
        nbr = [set() for _ in range(n)]
        for a, b in de:
            nbr[a].add(b)
            nbr[b].add(a)
        qn = [set() for _ in range(nq)]
        for a, b in qe:
            qn[a].add(b)
            qn[b].add(a)
        nvlq = max(ql) + 1
        for u in range(nq):
            want = set()
            for v in range(n):
                if dl[v] != ql[u] or len(nbr[v]) < len(qn[u]):
                    continue
                if all(sum(1 for w in nbr[v] if dl[w] == l) >= sum(1 for x in qn[u] if ql[x] == l)
                       for l in range(nvlq)):
                    want.add(v)
            assert got[u] == want, (it, u)


def test_refinement_sound_and_subset():
    r = random.Random(5)
    for it in range(100):
        n = r.randint(5, 12)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.2, 0.6), nl)
        nq = r.randint(2, 6)
        ql, qe = rand_query(r, nq, r.randint(0, 3), nl)
        q = O.Query.from_arrays(ql, qe)
        p = q.plan()
        d = O.Data.from_arrays(dl, de)
        base = [set(map(int, s)) for s in O.bits_to_sets(O.filter_nlc(d, n, q), n)]
        ref = [set(map(int, s)) for s in O.bits_to_sets(O.refine(d, n, q, p), n)]
        for u in range(nq):
            assert ref[u] <= base[u]
        for t in brute_force(dl, de, ql, qe):
            for u in range(nq):
                assert t[u] in ref[u], (it, u, t)


def _refine_by_spec(nbr, sets, clusters):
    """The refinement spec (csm.h, csm_filter) with Python sets: for every
    flattened cluster [k, m1, m2, ...] in order, first the core step
    C(k) <- {v in C(k) : N(v) meets C(m) for every member m}, then the member
    step C(m) <- {w in C(m) : N(w) meets C(k)} for each member."""
    cur = [set(s) for s in sets]
    for cl in clusters:
        k, members = cl[0], cl[1:]
        cur[k] = {v for v in cur[k] if all(nbr[v] & cur[m] for m in members)}
        for m in members:
            cur[m] = {w for w in cur[m] if nbr[w] & cur[k]}
    return cur


def test_refinement_equals_spec():
    """oc_refine equals the spec computed with sets, exactly, starting from the
    NLC sets; and the refinement removes candidates on these inputs, so a
    refinement that did nothing (or skipped a step) would fail."""
    r = random.Random(17)
    removed = 0
    for it in range(150):
        n = r.randint(8, 40)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.05, 0.3), nl)
        nq = r.randint(3, 9)
        ql, qe = rand_query(r, nq, r.randint(0, 3), nl)
        q = O.Query.from_arrays(ql, qe)
        p = q.plan()
        d = O.Data.from_arrays(dl, de)
        nbr = [set() for _ in range(n)]
        for a, b in de:
            nbr[a].add(b)
            nbr[b].add(a)
        base = [set(map(int, s)) for s in O.bits_to_sets(O.filter_nlc(d, n, q), n)]
        got = [set(map(int, s)) for s in O.bits_to_sets(O.refine(d, n, q, p), n)]
        want = _refine_by_spec(nbr, base, p.clusters)
        assert got == want, it
        removed += sum(len(b) - len(w) for b, w in zip(base, want))
    assert removed > 100, removed


def test_refinement_spec_on_the_worked_example():
    """The refined sets of the reference's worked example (example_q_01 over
    example_g_01; u0,u2: {0,4,5,6,7,8}; u1: {1,6,8}; u3: {2,9}; u4: {3,10}),
    traced by hand through its clusters, from the spec."""
    qp, gp = example_paths()
    q = O.Query.read(qp)
    d = O.Data.read(gp)
    got = [sorted(map(int, s)) for s in O.bits_to_sets(O.refine(d, 11, q, q.plan()), 11)]
    assert got == [[0, 4, 5, 6, 7, 8], [1, 6, 8], [0, 4, 5, 6, 7, 8], [2, 9], [3, 10]]


def _brute_min_cover(n, tree_edges):
    for k in range(n + 1):
        for S in itertools.combinations(range(n), k):
            s = set(S)
            if all(a in s or b in s for a, b in tree_edges):
                return k
    return n


def test_decycle_and_vertex_cover_invariants():
    r = random.Random(9)
    for it in range(200):
        nq = r.randint(2, 12)
        ql, qe = rand_query(r, nq, r.randint(0, 6), 3)
        p = O.Query.from_arrays(ql, qe).plan()
        off, adj, deg = p.tree_csr
        tree_edges = {(min(a, b), max(a, b)) for a in range(nq) for b in adj[off[a]:off[a + 1]]}
        assert len(tree_edges) == nq - 1
        assert sum(deg) == 2 * (nq - 1)
        assert tree_edges <= {(min(a, b), max(a, b)) for a, b in qe}
        # the BFS root is the first vertex of maximum degree
        full_deg = [0] * nq
        for a, b in qe:
            full_deg[a] += 1
            full_deg[b] += 1
        assert p.p.tree_root == full_deg.index(max(full_deg))
        cover = p.cover
        assert all(a in cover or b in cover for a, b in tree_edges)
        assert len(cover) == _brute_min_cover(nq, tree_edges)
        # matching order covers every vertex once, parents precede children
        mo = p.match_order
        assert sorted(mo) == list(range(nq)) and mo[0] == p.root
        for i in range(1, nq):
            assert p.parents[i] < i
        assert sum(len(b) for b in p.backs) == len(qe) - (nq - 1)


@pytest.mark.skipif(not O.ref_decycle_available(), reason="oracle/_ref not built (needs the reference tree)")
def test_decycle_matches_reference_source():
    """The reference's own decycle() (decycle.cpp, compiled into oracle/_ref)."""
    r = random.Random(13)
    for it in range(500):
        nq = r.randint(2, 32)
        ql, qe = rand_query(r, nq, r.randint(0, nq), 4)
        off, adj, deg, tm = O.ref_decycle(ql, qe)
        p = O.Query.from_arrays(ql, qe).plan()
        o2, a2, d2 = p.tree_csr
        assert (list(map(int, off)), list(map(int, adj)), list(map(int, deg))) == (o2, a2, d2), it
        assert tm == nq - 1


def test_oracle_rejects_malformed(tmp_path):
    bad = tmp_path / "bad.in"
    bad.write_text("x 3 2\n")
    with pytest.raises(O.OracleError):
        O.Query.read(str(bad))
    bad.write_text("t 2 1\nv 0 0 2\nv 1 0 1\ne 0 1\n")  # declared degree mismatch
    with pytest.raises(O.OracleError):
        O.Query.read(str(bad))
    with pytest.raises(O.OracleError):  # disconnected query
        O.Query.from_arrays([0, 0, 0, 0], [(0, 1), (2, 3)]).plan()


@pytest.mark.parametrize("case", max_size_cases(), ids=lambda c: c[0])
def test_max_size_queries_closed_forms(case):
    """32-vertex queries on labelled cycles: the oracle's count and list
    against counts fixed by the cycles' rotations and reflections; every
    listed row is an embedding (labels kept, edges mapped to edges,
    injective)."""
    name, dl, de, ql, qe, want = case
    oq = O.Query.from_arrays(ql, qe)
    c, capped, rows, _, _ = O.match(O.Data.from_arrays(dl, de), oq, oq.plan(), want_list=True)
    assert (c, bool(capped)) == (want, False)
    adj = {(a, b) for a, b in de} | {(b, a) for a, b in de}
    assert len({tuple(r) for r in rows.tolist()}) == want
    for r in rows.tolist():
        assert len(set(r)) == 32
        assert all(dl[r[i]] == ql[i] for i in range(32))
        assert all((r[a], r[b]) in adj for a, b in qe)
