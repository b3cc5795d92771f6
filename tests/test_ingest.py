"""Graph ingest at scale (host code, CPU only).

Text files of 32 MB and more are parsed by several threads (records counted
per newline-aligned chunk, then parsed into place); anything irregular falls
back to the sequential parser.  Both must give the same graph as the
generator that wrote the file, and the same error as the sequential parser on
malformed input.  The binary graph file must round-trip exactly and reject
bad headers, sizes and endpoints."""
import os
import random
import time

import numpy as np
import pytest

import clustersm_amd as csm

CSM_E_FORMAT = 3


@pytest.fixture(scope="module")
def big(tmp_path_factory):
    """An Erdos-Renyi graph of 300,000 vertices and 3,000,000 edges written as
    text (about 52 MB), with the generator's own arrays to compare against."""
    hg = csm.gen_erdos_renyi(300_000, 3_000_000, 8, 17)
    path = tmp_path_factory.mktemp("ingest") / "big.graph"
    hg.write(str(path))
    labels, edges = hg.arrays()
    assert os.path.getsize(path) >= 32 << 20
    return path, labels, edges


def _lines(path):
    with open(path, "rb") as f:
        return f.read().split(b"\n")


def test_parallel_parse_equals_generator(big):
    path, labels, edges = big
    t = time.perf_counter()
    g = csm.HostGraph.read(str(path))
    dt = time.perf_counter() - t
    print(f"parsed {os.path.getsize(path) / 1e6:.0f} MB in {dt:.2f} s", flush=True)
    gl, ge = g.arrays()
    assert np.array_equal(gl, labels) and np.array_equal(ge, edges)


def test_records_split_across_lines_parse_the_same(big, tmp_path):
    """Whitespace, newlines included, separates tokens: an edge record split
    over two lines is the same record, whether or not a chunk boundary falls
    inside it."""
    path, labels, edges = big
    lines = _lines(path)
    r = random.Random(3)
    first_edge = next(i for i, ln in enumerate(lines) if ln.startswith(b"e "))
    for i in r.sample(range(first_edge, len(lines) - 1), 400):
        lines[i] = lines[i].replace(b" ", b"\n", 1) if r.random() < 0.5 else lines[i].replace(b" ", b" \n", 2)
    p2 = tmp_path / "split.graph"
    p2.write_bytes(b"\n".join(lines))
    gl, ge = csm.HostGraph.read(str(p2)).arrays()
    assert np.array_equal(gl, labels) and np.array_equal(ge, edges)


@pytest.mark.parametrize("kind", ["bad_token", "vertex_after_edges", "degree", "endpoint"])
def test_malformed_large_files_report_the_sequential_error(big, tmp_path, kind):
    path, labels, edges = big
    lines = _lines(path)
    n = len(labels)
    first_edge = next(i for i, ln in enumerate(lines) if ln.startswith(b"e "))
    k = 1_234_567  # edge index
    if kind == "bad_token":
        lines[first_edge + k] = b"e 12 x"
        want = f"malformed edge line {k}"
    elif kind == "vertex_after_edges":
        lines.insert(first_edge + k, b"v 5 0 0")
        want = f"malformed vertex line {n}"
    elif kind == "degree":
        a, b = (int(x) for x in lines[first_edge + k].split()[1:3])
        c = next(v for v in range(n) if v not in (a, b))
        lines[first_edge + k] = f"e {a} {c}".encode()
        want = "declared degree of vertex"
    else:
        lines[first_edge + k] = f"e 1 {n}".encode()
        want = f"malformed edge line {k}"
    p2 = tmp_path / "bad.graph"
    p2.write_bytes(b"\n".join(lines))
    with pytest.raises(csm.CsmError) as e:
        csm.HostGraph.read(str(p2))
    assert e.value.status == CSM_E_FORMAT and want in str(e.value), str(e.value)


def test_binary_graph_round_trip(big, tmp_path):
    path, labels, edges = big
    g = csm.HostGraph.read(str(path))
    pb = tmp_path / "big.csmbin"
    g.write_bin(str(pb))
    assert os.path.getsize(pb) == 24 + 4 * len(labels) + 8 * len(edges)
    t = time.perf_counter()
    gb = csm.HostGraph.read_bin(str(pb))
    print(f"binary read in {time.perf_counter() - t:.2f} s", flush=True)
    bl, be = gb.arrays()
    assert np.array_equal(bl, labels) and np.array_equal(be, edges)


def test_binary_graph_rejects_bad_files(tmp_path):
    g = csm.HostGraph.from_edges([0, 1, 2], [(0, 1), (1, 2)])
    p = tmp_path / "g.csmbin"
    g.write_bin(str(p))
    raw = p.read_bytes()
    cases = {
        "magic": b"XSMGBIN1" + raw[8:],
        "truncated": raw[:-4],
        "longer": raw + b"\0\0\0\0",
        "endpoint": raw[:-4] + (7).to_bytes(4, "little"),
    }
    for name, data in cases.items():
        q = tmp_path / f"{name}.csmbin"
        q.write_bytes(data)
        with pytest.raises(csm.CsmError) as e:
            csm.HostGraph.read_bin(str(q))
        assert e.value.status == CSM_E_FORMAT, name
    gl, ge = csm.HostGraph.read_bin(str(p)).arrays()
    assert gl.tolist() == [0, 1, 2] and ge.tolist() == [[0, 1], [1, 2]]
