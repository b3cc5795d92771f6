"""Host layer of the product (libcsm.so, no GPU calls): planner against the
oracle's independent restatement, parser errors, generators, and the C ABI's
exported symbols."""
import os
import random
import re
import subprocess

import numpy as np
import pytest

import clustersm_amd as csm
from helpers import example_paths, load_example, rand_query
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "csm.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(csm_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 30
    L = csm.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert names == set(csm.SIGNATURES), sorted(names ^ set(csm.SIGNATURES))


def test_plan_struct_layout_matches_oracle_fields():
    assert csm.QueryPlan.num_v.offset == 0
    # the device never sees the Python struct; sanity-check the size is what C sees
    assert csm.C.sizeof(csm.QueryPlan) % 4 == 0


def _plan_tuple(p):
    return (p.tree_csr, p.cover if hasattr(p, "cover_size") is False else list(p.cover[: p.cover_size]),
            p.clusters, p.root, p.order, p.match_order, p.parents, p.backs)


def _oracle_tuple(p):
    return (p.tree_csr, p.cover, p.clusters, p.root, p.order, p.match_order, p.parents, p.backs)


def test_example_plan():
    qp, _ = example_paths()
    ex = load_example()
    p = csm.HostGraph.read(qp).plan()
    assert list(p.cover[: p.cover_size]) == ex["vertex_cover"]
    assert p.clusters == ex["clusters"]
    assert p.compact == ex["compact"]
    assert p.root == ex["root"]
    assert p.order == [tuple(e) for e in ex["edge_order"]]
    assert p.tree_root == ex["decycle_tree_root"]


def test_planner_matches_oracle_on_random_queries():
    r = random.Random(21)
    for it in range(3000):
        nq = r.randint(2, 32)
        ql, qe = rand_query(r, nq, r.randint(0, nq), r.randint(1, 6))
        pp = csm.HostGraph.from_edges(ql, qe).plan()
        po = O.Query.from_arrays(ql, qe).plan()
        assert _plan_tuple(pp) == _oracle_tuple(po), it
        assert list(pp.cluster_valid[: pp.num_clusters]) == list(po.p.cluster_valid[: po.p.n_clusters])
        assert pp.num_layers == po.p.n_layers


def test_planner_rejects_bad_queries():
    with pytest.raises(csm.CsmError) as e:
        csm.HostGraph.from_edges([0, 0, 0, 0], [(0, 1), (2, 3)]).plan()
    assert e.value.status == csm.E_QUERY
    with pytest.raises(csm.CsmError):
        csm.HostGraph.from_edges([0, 0], [(0, 1), (1, 0)]).plan()  # duplicate edge
    with pytest.raises(csm.CsmError):
        csm.HostGraph.from_edges([0], []).plan()
    with pytest.raises(csm.CsmError):
        csm.HostGraph.from_edges(list(range(33)), [(i, i + 1) for i in range(32)]).plan()


def test_parser_errors(tmp_path):
    f = tmp_path / "g.in"
    with pytest.raises(csm.CsmError) as e:
        csm.HostGraph.read(str(tmp_path / "missing.in"))
    assert e.value.status == csm.E_IO
    for text in ["x 3 2\n", "t 2 1\nv 0 0 2\nv 1 0 1\ne 0 1\n", "t 2 1\nv 1 0 1\nv 0 0 1\ne 0 1\n",
                 "t 2 1\nv 0 0 1\nv 1 0 1\ne 0 5\n", "t 2 2\nv 0 0 1\nv 1 0 1\ne 0 1\n"]:
        f.write_text(text)
        with pytest.raises(csm.CsmError) as e:
            csm.HostGraph.read(str(f))
        assert e.value.status == csm.E_FORMAT, text


def test_parser_rejects_headers_larger_than_the_file(tmp_path):
    """A header promising more records than the file can hold is an error
    before anything is sized by it (no multi-GB allocation for a 30-byte file)."""
    f = tmp_path / "g.in"
    for text in ["t 4000000000 1\n", "t 2 900000000000000000\nv 0 0 0\nv 1 0 0\n", "t 3 0\nv 0 0 0\nv 1 0 0\n"]:
        f.write_text(text)
        with pytest.raises(csm.CsmError) as e:
            csm.HostGraph.read(str(f))
        assert e.value.status == csm.E_FORMAT and "bytes after it can hold" in str(e.value), (text, str(e.value))
    f.write_text("t 2 1\nv 0 0 1\nv 1 0 1\ne 0 1")  # the smallest records, no final newline
    assert csm.HostGraph.read(str(f)).arrays()[1].tolist() == [[0, 1]]


def test_parser_roundtrip_and_edge_order(tmp_path):
    qp, gp = example_paths()
    g = csm.HostGraph.read(gp)
    labels, edges = g.arrays()
    assert g.num_v == 11 and g.num_e == 13
    assert edges[0].tolist() == [0, 1] and edges[-1].tolist() == [9, 10]
    out = tmp_path / "rt.in"
    g.write(str(out))
    g2 = csm.HostGraph.read(str(out))
    l2, e2 = g2.arrays()
    assert (labels == l2).all() and (edges == e2).all()
    # declared degrees sum to 2M (io.cu:75)
    text = out.read_text().split("\n")
    assert sum(int(l.split()[3]) for l in text if l.startswith("v")) == 2 * 13


def test_generators_deterministic():
    a = csm.gen_rmat(10, 8000, 8, 1.0, 5)
    b = csm.gen_rmat(10, 8000, 8, 1.0, 5)
    la, ea = a.arrays()
    lb, eb = b.arrays()
    assert (la == lb).all() and (ea == eb).all()
    assert la.max() < 8 and ea.max() < 1024
    er = csm.gen_erdos_renyi(1000, 4000, 4, 1)
    l, e = er.arrays()
    assert e.shape == (4000, 2)
    assert len({(min(x, y), max(x, y)) for x, y in e.tolist()}) == 4000
    assert (e[:, 0] != e[:, 1]).all()
    assert set(l.tolist()) == {0, 1, 2, 3}
    # Zipf(1): label 0 is the most frequent
    z = csm.gen_rmat(12, 20000, 16, 1.0, 3).arrays()[0]
    cnt = np.bincount(z, minlength=16)
    assert cnt[0] == cnt.max() and cnt[0] > 2 * cnt[8]


def test_query_generator_shapes():
    g = csm.gen_erdos_renyi(1000, 4000, 4, 1)
    q = csm.gen_query(g, 4, csm.QUERY_INDUCED, 2)
    assert q.num_v == 4
    p = q.plan()
    assert p.num_v == 4
    t = csm.gen_query(g, 8, csm.QUERY_TREE, 3)
    assert t.num_e == 7
    x = csm.gen_query(g, 8, csm.QUERY_EXTRA, 3, extra_edges=3)
    assert 7 <= x.num_e <= 10
    # labels are data labels of connected walk vertices: the plan must build
    for s in range(50):
        csm.gen_query(g, 12, csm.QUERY_AVG_DEGREE, s, avg_degree=3.0).plan()


CLI = os.path.join(ROOT, "build", "ClusterSM")


@pytest.mark.skipif(not os.path.exists(CLI), reason="CLI not built")
def test_cli_argument_errors():
    qp, gp = example_paths()
    run = lambda *a: subprocess.run([CLI, *a], capture_output=True, text=True)
    assert run("-d", gp).returncode == 106
    assert run("-q", qp).returncode == 106
    assert run("-q", qp, "-d", gp, "--bogus").returncode == 109
    assert run("-q", qp, "-d", gp, "--gpu", "x").returncode == 104
    r = run("-h")
    assert r.returncode == 0 and "-q" in r.stdout
    assert run("-q", qp, "-d", "/nonexistent/g.in").returncode == 1


def test_abi_struct_layouts_match_ctypes(tmp_path):
    """Every struct the C ABI passes by pointer has the same size (and the
    later fields the same offsets) in include/csm.h and in the ctypes mirror."""
    import ctypes as C
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("gcc not found")
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "csm.h"\n'
        'int main(void) {\n'
        ' printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n", sizeof(csm_result), offsetof(csm_result, path),\n'
        '  offsetof(csm_result, kernels), sizeof(csm_batch_opts), offsetof(csm_batch_opts, item_budget),\n'
        '  sizeof(csm_serve_opts), sizeof(csm_match_opts), sizeof(csm_query_plan), sizeof(csm_kernel_stat),\n'
        '  sizeof(csm_ctx_info));\n return 0;\n}\n')
    exe = tmp_path / "sizes"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(csm.Result), csm.Result.path.offset, csm.Result.kernels.offset, C.sizeof(csm.BatchOpts),
            csm.BatchOpts.item_budget.offset, C.sizeof(csm.ServeOpts), C.sizeof(csm.MatchOpts),
            C.sizeof(csm.QueryPlan), C.sizeof(csm.KernelStat), C.sizeof(csm.CtxInfo)]
    assert got == want
