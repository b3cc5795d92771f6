"""The drop-in CLI on the GPU: `ClusterSM -q <query> -d <data>` as
runClusterSM.sh:9-13 calls it, with the last stdout line `num_res: N`
(main.cu:172).  The worked example's count and sorted list; a binary data
file gives the same result as its text; `--batch` gives each query's
single-run count."""
import json
import os
import subprocess

import pytest

import clustersm_amd as csm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "build", "ClusterSM")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(*args):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} is not built")
    print("ClusterSM", *args, flush=True)
    return subprocess.run([CLI, *args, "--arena-gb", "4"], capture_output=True, text=True, timeout=120)


def last_count(out):
    line = out.strip().splitlines()[-1]
    assert line.startswith("num_res: "), out[-500:]
    return int(line.split()[1])


def test_cli_worked_example_count_and_list():
    qp, gp = os.path.join(GOLDEN, "example_q_01.in"), os.path.join(GOLDEN, "example_g_01.in")
    with open(os.path.join(GOLDEN, "example_expected.json")) as f:
        want = json.load(f)
    r = run("-q", qp, "-d", gp)
    assert r.returncode == 0, r.stderr
    assert last_count(r.stdout) == want["count"]
    r = run("-q", qp, "-d", gp, "--list")
    assert r.returncode == 0, r.stderr
    rows = [[int(x) for x in ln.split()] for ln in r.stdout.strip().splitlines()[:-1]]
    assert rows == sorted(want["embeddings"])


def test_cli_binary_data_and_batch(tmp_path):
    hg = csm.gen_rmat(12, 40000, 8, 1.0, 9)
    text, binary = tmp_path / "data.graph", tmp_path / "data.csmbin"
    hg.write(str(text))
    hg.write_bin(str(binary))
    qpaths = []
    for i in range(4):
        q = csm.gen_query(hg, 4 + 2 * i, csm.QUERY_EXTRA, 60 + i, extra_edges=1)
        p = tmp_path / f"q{i}.graph"
        q.write(str(p))
        qpaths.append(str(p))
    cap = "200000"
    single = []
    for p in qpaths:
        a = run("-q", p, "-d", str(text), "--cap", cap)
        b = run("-q", p, "-d", str(binary), "--cap", cap)
        assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
        assert last_count(a.stdout) == last_count(b.stdout)
        single.append(last_count(a.stdout))
    listing = tmp_path / "queries.txt"
    listing.write_text("\n".join(qpaths) + "\n")
    r = run("-q", qpaths[0], "-d", str(binary), "--batch", str(listing), "--cap", cap)
    assert r.returncode == 0, r.stderr
    got = [int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith("num_res: ")]
    assert got == single
    # the candidate-aware order counts the same
    r = run("-q", qpaths[0], "-d", str(binary), "--batch", str(listing), "--cap", cap, "--reorder")
    assert r.returncode == 0, r.stderr
    assert [int(ln.split()[1]) for ln in r.stdout.splitlines() if ln.startswith("num_res: ")] == single
    for p, want in zip(qpaths, single):
        r = run("-q", p, "-d", str(text), "--cap", cap, "--reorder")
        assert r.returncode == 0 and last_count(r.stdout) == want, r.stderr
