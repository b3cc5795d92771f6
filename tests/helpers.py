"""Independent checkers used by the tests: brute-force enumeration and
random small graphs (Python's own RNG, not the product's generators)."""
import itertools
import json
import os
import random

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def brute_force(dl, de, ql, qe):
    """All injective label-preserving maps with every query edge on a data edge."""
    adj = set()
    for a, b in de:
        if a != b:
            adj.add((a, b))
            adj.add((b, a))
    nq = len(ql)
    by_label = {}
    for v, l in enumerate(dl):
        by_label.setdefault(l, []).append(v)
    out = []
    pools = [by_label.get(l, []) for l in ql]
    for t in itertools.product(*pools):
        if len(set(t)) != nq:
            continue
        if all((t[a], t[b]) in adj for a, b in qe):
            out.append(tuple(t))
    return sorted(out)


def rand_graph(r, n, p, nl):
    labels = [r.randrange(nl) for _ in range(n)]
    edges = [(a, b) for a in range(n) for b in range(a + 1, n) if r.random() < p]
    r.shuffle(edges)
    return labels, [(a, b) if r.random() < 0.5 else (b, a) for a, b in edges]


def rand_query(r, n, extra, nl):
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
    return [r.randrange(nl) for _ in range(n)], [(a, b) if r.random() < 0.5 else (b, a) for a, b in edges]


def load_golden():
    with open(os.path.join(GOLDEN, "mono_small.json")) as f:
        return json.load(f)["cases"]


def load_example():
    with open(os.path.join(GOLDEN, "example_expected.json")) as f:
        return json.load(f)


def example_paths():
    return os.path.join(GOLDEN, "example_q_01.in"), os.path.join(GOLDEN, "example_g_01.in")


def list_digest(rows):
    import hashlib
    h = hashlib.sha256()
    for t in rows:
        h.update((" ".join(map(str, t)) + "\n").encode())
    return h.hexdigest()


def _cycle(n):
    return [(i, (i + 1) % n) for i in range(n)]


def max_size_cases():
    """(name, data labels, data edges, query labels, query edges, count) for
    32-vertex queries on labelled cycles, the counts by closed form:
    * a path labelled i mod 4 in a 64-cycle labelled j mod 4: only the
      forward walk keeps the label order, one per start labelled 0: 16;
    * the same path closed into a 32-cycle, in the 64-cycle: vertices 31
      steps apart are never adjacent there: 0;
    * that 32-cycle in a 32-cycle labelled j mod 4: the rotations by a
      multiple of 4: 8;
    * one label everywhere, the 32-cycle in the 32-cycle: the 64 dihedral
      images; the 32-vertex path in a 40-cycle: 40 starts x 2 directions: 80;
    * a query with a label the data graph does not hold: 0."""
    path = [(i, i + 1) for i in range(31)]
    lab4 = [i % 4 for i in range(32)]
    return [
        ("path in C64, 4 labels", [j % 4 for j in range(64)], _cycle(64), lab4, path, 16),
        ("C32 in C64, 4 labels", [j % 4 for j in range(64)], _cycle(64), lab4, _cycle(32), 0),
        ("C32 in C32, 4 labels", [j % 4 for j in range(32)], _cycle(32), lab4, _cycle(32), 8),
        ("C32 in C32, 1 label", [0] * 32, _cycle(32), [0] * 32, _cycle(32), 64),
        ("path in C40, 1 label", [0] * 40, _cycle(40), [0] * 32, path, 80),
        ("absent label", [j % 4 for j in range(64)], _cycle(64), lab4[:31] + [7], path, 0),
    ]
