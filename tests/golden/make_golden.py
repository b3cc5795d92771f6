"""Writes tests/golden/mono_small.json: small labelled (data, query) pairs with
every subgraph monomorphism listed, computed by networkx's VF2
GraphMatcher.subgraph_monomorphisms_iter (an implementation independent of
this repository).  Run in the build container: python tests/golden/make_golden.py

Each case: data labels/edges, query labels/edges (file order), the count, the
sha256 of the sorted embedding list (query-id order) and, for up to 400
embeddings, the list itself.

The reference's bundled example pair (dataset/example_{q,g}_01.in) is kept
beside this script as data; example_expected.json holds its answer and the
planner trace read from the reference source (filter.cu, order.cpp, main.cu).
"""
import hashlib
import json
import os
import random

import networkx as nx
from networkx.algorithms import isomorphism as iso

HERE = os.path.dirname(os.path.abspath(__file__))


def monomorphisms(dl, de, ql, qe):
    G = nx.Graph()
    G.add_nodes_from((v, {"l": l}) for v, l in enumerate(dl))
    G.add_edges_from(de)
    Q = nx.Graph()
    Q.add_nodes_from((v, {"l": l}) for v, l in enumerate(ql))
    Q.add_edges_from(qe)
    gm = iso.GraphMatcher(G, Q, node_match=lambda a, b: a["l"] == b["l"])
    out = set()
    for m in gm.subgraph_monomorphisms_iter():
        t = [None] * len(ql)
        for dv, qv in m.items():
            t[qv] = dv
        out.add(tuple(t))
    return sorted(out)


def list_digest(embs):
    """sha256 of the sorted embedding list, one line per embedding: "v0 v1 ...\n"."""
    h = hashlib.sha256()
    for t in embs:
        h.update((" ".join(map(str, t)) + "\n").encode())
    return h.hexdigest()


def rand_graph(r, n, p, nl):
    labels = [r.randrange(nl) for _ in range(n)]
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if r.random() < p]
    r.shuffle(edges)
    edges = [(a, b) if r.random() < 0.5 else (b, a) for a, b in edges]
    return labels, edges


def rand_connected_query(r, n, extra, nl):
    order = list(range(n))
    r.shuffle(order)
    edges = set()
    for i in range(1, n):
        a, b = order[i], order[r.randrange(i)]
        edges.add((min(a, b), max(a, b)))
    for _ in range(extra):
        a, b = r.sample(range(n), 2)
        edges.add((min(a, b), max(a, b)))
    edges = list(edges)
    r.shuffle(edges)
    edges = [(a, b) if r.random() < 0.5 else (b, a) for a, b in edges]
    return [r.randrange(nl) for _ in range(n)], edges


def main():
    r = random.Random(20261203)
    cases = []
    for i in range(60):
        n = r.randint(6, 14)
        nl = r.randint(1, 3)
        dl, de = rand_graph(r, n, r.uniform(0.2, 0.6), nl)
        nq = r.randint(2, min(6, n))
        ql, qe = rand_connected_query(r, nq, r.randint(0, 3), nl)
        embs = monomorphisms(dl, de, ql, qe)
        case = {"name": f"rand{i}", "data_labels": dl, "data_edges": de,
                "query_labels": ql, "query_edges": qe, "count": len(embs),
                "sha256": list_digest(embs)}
        if len(embs) <= 400:
            case["embeddings"] = embs
        cases.append(case)
    out = {"generator": "networkx.algorithms.isomorphism.GraphMatcher.subgraph_monomorphisms_iter "
                        f"(networkx {nx.__version__})", "cases": cases}
    with open(os.path.join(HERE, "mono_small.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(len(cases), "cases, max count", max(c["count"] for c in cases))


if __name__ == "__main__":
    main()
