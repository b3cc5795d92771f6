"""The device code on the CPU, under AddressSanitizer.

tests/emu/build.py compiles the library's HIP sources for the host: each
workgroup runs on an OS thread and its lanes on fibers, with block barriers
and wave operations as meeting points (tests/emu/emu_rt.cpp).  Every global
allocation is its own heap block, so an out-of-bounds access in a kernel is
reported with its stack.  The plain path (also through csm_match_noalloc,
its workspace in the static scratch, and with the count level's lists
copied into LDS) and the persistent batched kernel
run on small random graphs and queries (with non-tree edges) and must give
the oracle's counts, capped and uncapped, also with the candidate-aware
order (reorder option), and the plain path's list under a cap of 64 (sorted
on the device) must be the oracle's.  Test infrastructure only; the product
is the gfx950 build."""
import os
import random
import subprocess
import sys

import pytest

import clustersm_amd as csm
from helpers import rand_graph, rand_query
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def driver():
    sys.path.insert(0, os.path.join(HERE, "emu"))
    import build
    try:
        return str(build.build())
    except OSError as e:
        pytest.skip(f"g++ is not available here: {e}")


def _oracle(dl, de, queries, cap):
    d = O.Data.from_arrays(dl, de)
    out = []
    for ql, qe in queries:
        oq = O.Query.from_arrays(ql, qe)
        c, capped, _, _, _ = O.match(d, oq, oq.plan(), cap=cap if cap else (1 << 64) - 1, threads=4)
        out.append((c, bool(capped)))
    return out


def _oracle_lists(dl, de, queries, cap=64):
    d = O.Data.from_arrays(dl, de)
    out = []
    for ql, qe in queries:
        oq = O.Query.from_arrays(ql, qe)
        _, _, rows, _, _ = O.match(d, oq, oq.plan(), cap=cap, want_list=True, list_cap=cap, threads=1)
        out.append(sorted(tuple(int(x) for x in r) for r in rows.tolist()))  # the oracle lists in matching order
    return out


def _parse(stdout):
    """Per-kind results and the listed rows of each query, from the driver."""
    got = {"plain": [], "noalloc": [], "plainl": [], "plains": [], "batch": [], "plainr": [], "batchr": [],
           "served": []}
    lists = []
    for ln in stdout.split("\n"):
        f = ln.split()
        if not f:
            continue
        if f[0] in got:
            got[f[0]].append((int(f[1]), f[2] == "1"))
        elif f[0] == "list":
            lists.append([])
        elif f[0] == "row":
            lists[-1].append(tuple(int(x) for x in f[1:]))
    return got, lists


@pytest.mark.parametrize("seed", [3, 4])
def test_device_code_on_cpu_matches_oracle(driver, tmp_path, seed):
    r = random.Random(seed)
    n = r.randint(60, 140)
    nl = r.randint(1, 3)
    dl, de = rand_graph(r, n, r.uniform(0.05, 0.12), nl)
    queries = [rand_query(r, r.randint(3, 7), r.randint(0, 3), nl) for _ in range(6)]
    dpath = tmp_path / "data.graph"
    csm.HostGraph.from_edges(dl, de).write(str(dpath))
    qpaths = []
    for i, (ql, qe) in enumerate(queries):
        p = tmp_path / f"q{i}.graph"
        csm.HostGraph.from_edges(ql, qe).write(str(p))
        qpaths.append(str(p))
    env = dict(os.environ, EMU_THREADS="4", EMU_CUS="2", ASAN_OPTIONS="detect_leaks=0")
    for cap in (0, 50):
        out = subprocess.run([driver, str(dpath), str(cap), *qpaths], env=env, capture_output=True, text=True,
                             timeout=600)
        assert out.returncode == 0, out.stderr[-3000:]
        assert "AddressSanitizer" not in out.stderr, out.stderr[-3000:]
        got, lists = _parse(out.stdout)
        want = _oracle(dl, de, queries, cap)
        for kind in ("plain", "noalloc", "plainl", "batch", "plainr", "batchr", "served") + (("plains",) if cap == 0 else ()):
            assert got[kind] == want, (cap, kind, got[kind], want)
        assert lists == _oracle_lists(dl, de, queries), cap


def test_heavy_candidate_lists_on_cpu(driver, tmp_path):
    """Vertices of degree >= 1024 have their candidate lists built by a
    workgroup each, outside the per-word walk: a graph with three such hubs
    (label 0, the only vertices of that label) and queries through them give
    the oracle's counts, uncapped and capped."""
    r = random.Random(12)
    n = 2600
    dl = [1 + r.randrange(3) for _ in range(n)]
    hubs = [0, 700, 1901]
    for h in hubs:
        dl[h] = 0
    edges = set()
    for h in hubs:
        for v in r.sample(range(n), 1000 + r.randrange(600)):
            if v != h:
                edges.add((min(h, v), max(h, v)))
    while len(edges) < 9000:
        a, b = r.randrange(n), r.randrange(n)
        if a != b:
            edges.add((min(a, b), max(a, b)))
    de = sorted(edges)
    queries = [
        ([0, 1], [(0, 1)]),
        ([0, 1, 2], [(0, 1), (1, 2)]),
        ([2, 0, 3], [(0, 1), (1, 2), (0, 2)]),
        ([1, 0, 2, 3], [(0, 1), (0, 2), (2, 3), (1, 3)]),
    ]
    dpath = tmp_path / "data.graph"
    csm.HostGraph.from_edges(dl, de).write(str(dpath))
    qpaths = []
    for i, (ql, qe) in enumerate(queries):
        p = tmp_path / f"q{i}.graph"
        csm.HostGraph.from_edges(ql, qe).write(str(p))
        qpaths.append(str(p))
    env = dict(os.environ, EMU_THREADS="4", EMU_CUS="2", ASAN_OPTIONS="detect_leaks=0")
    for cap in (0, 50):
        out = subprocess.run([driver, str(dpath), str(cap), *qpaths], env=env, capture_output=True, text=True,
                             timeout=900)
        assert out.returncode == 0, out.stderr[-3000:]
        assert "AddressSanitizer" not in out.stderr, out.stderr[-3000:]
        got, lists = _parse(out.stdout)
        want = _oracle(dl, de, queries, cap)
        for kind in ("plain", "noalloc", "plainl", "batch", "plainr", "batchr", "served") + (("plains",) if cap == 0 else ()):
            assert got[kind] == want, (cap, kind, got[kind], want)
        assert lists == _oracle_lists(dl, de, queries), cap


def test_dense_rows_lds_lists_on_cpu(driver, tmp_path):
    """A dense graph without hubs (160 vertices of degree near 24, two labels)
    and queries with non-tree edges: the count level's rows have many items
    and short lists, so the lds_lists option copies lists into LDS; every path
    gives the oracle's counts."""
    r = random.Random(1)
    n = 160
    edges = set()
    while len(edges) < n * 12:
        a, b = r.sample(range(n), 2)
        edges.add((min(a, b), max(a, b)))
    de = sorted(edges)
    dl = [r.randrange(2) for _ in range(n)]
    queries = [rand_query(r, r.randint(3, 5), r.randint(1, 3), 2) for _ in range(4)]
    dpath = tmp_path / "data.graph"
    csm.HostGraph.from_edges(dl, de).write(str(dpath))
    qpaths = []
    for i, (ql, qe) in enumerate(queries):
        p = tmp_path / f"q{i}.graph"
        csm.HostGraph.from_edges(ql, qe).write(str(p))
        qpaths.append(str(p))
    env = dict(os.environ, EMU_THREADS="4", EMU_CUS="2", ASAN_OPTIONS="detect_leaks=0")
    out = subprocess.run([driver, str(dpath), "0", *qpaths], env=env, capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-3000:]
    assert "AddressSanitizer" not in out.stderr, out.stderr[-3000:]
    got, _ = _parse(out.stdout)
    want = _oracle(dl, de, queries, 0)
    for kind in ("plain", "noalloc", "plainl", "plains", "batch", "plainr", "batchr", "served"):
        assert got[kind] == want, (kind, got[kind], want)
