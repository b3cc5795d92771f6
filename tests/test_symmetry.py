"""Symmetry breaking for counting (csm_query_symmetry): |Aut(Q)| equals a
brute-force count of label- and edge-preserving permutations, and the
conditions f(v) < f(w) keep exactly one embedding of every Aut(Q) class,
checked by brute-force enumeration of all embeddings on small graphs (so the
conditioned count times |Aut(Q)| is the full count)."""
import itertools
import random

import clustersm_amd as csm
from helpers import brute_force, rand_graph, rand_query


def brute_automorphisms(labels, edges):
    n = len(labels)
    es = set()
    for a, b in edges:
        es.add((a, b))
        es.add((b, a))
    out = []
    for p in itertools.permutations(range(n)):
        if all(labels[p[i]] == labels[i] for i in range(n)) and all((p[a], p[b]) in es for a, b in edges):
            out.append(p)
    return out


def test_automorphism_count_matches_brute_force():
    r = random.Random(1)
    for it in range(400):
        n = r.randint(2, 8)
        ql, qe = rand_query(r, n, r.randint(0, 5), r.randint(1, 3))
        aut, _ = csm.query_symmetry(csm.HostGraph.from_edges(ql, qe).plan())
        assert aut == len(brute_automorphisms(ql, qe)), (it, ql, qe)


def test_symmetric_shapes():
    """Closed forms: a labelled-uniform k-clique has k! automorphisms, a
    k-cycle 2k, a star with k equal leaves k!, a path 2."""
    import math
    for k in range(3, 7):
        clique = [(a, b) for a in range(k) for b in range(a + 1, k)]
        cycle = [(i, (i + 1) % k) for i in range(k)]
        star = [(0, i) for i in range(1, k + 1)]
        path = [(i, i + 1) for i in range(k - 1)]
        for edges, n, want in ((clique, k, math.factorial(k)), (cycle, k, 2 * k), (star, k + 1, math.factorial(k)),
                               (path, k, 2)):
            aut, pairs = csm.query_symmetry(csm.HostGraph.from_edges([0] * n, edges).plan())
            assert aut == want, (edges, aut, want)


def test_conditions_keep_one_embedding_per_class():
    r = random.Random(7)
    checked = 0
    for it in range(120):
        n = r.randint(6, 11)
        nl = r.randint(1, 2)
        dl, de = rand_graph(r, n, r.uniform(0.3, 0.7), nl)
        nq = r.randint(2, 5)
        ql, qe = rand_query(r, nq, r.randint(0, 3), nl)
        aut, pairs = csm.query_symmetry(csm.HostGraph.from_edges(ql, qe).plan())
        autos = brute_automorphisms(ql, qe)
        assert aut == len(autos)
        emb = brute_force(dl, de, ql, qe)
        kept = [t for t in emb if all(t[v] < t[w] for v, w in pairs)]
        assert len(kept) * aut == len(emb), (it, ql, qe, pairs)
        # every class meets the conditions exactly once
        es = set(emb)
        for t in kept:
            images = {tuple(t[p[i]] for i in range(nq)) for p in autos}
            assert images <= es and len([x for x in images if x in set(kept)]) == 1
        checked += len(emb) > 0 and aut > 1
    assert checked > 10
