"""The planner against fixtures traced by hand from the reference's own
source (tests/golden/planner_traces.json): decycle (decycle.cpp:6-134), tree
vertex cover (filter.cu:23-112), clustering with merges and combines
(filter.cu:127-401), flattening (structure.cu:219-239), root (main.cu:144)
and BFS edge order (order.cpp:150-205).  Together the cases cover a merge
through a core (j_ptr == 0), merges through shared members with different
cores, a combine that creates a cluster, a combine with a holder cluster,
two combines in one layer, and runs of two layers after layer 0.  Both the
CPU oracle and the product's planner (csm_plan_build) must reproduce every
field exactly."""
import json
import os

import pytest

import clustersm_amd as csm
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "planner_traces.json")
CASES = json.load(open(GOLDEN))["cases"]


def _flatten(clusters):
    compact, off = [], [0]
    for c in clusters:
        compact += c
        off.append(len(compact))
    return compact, off


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_planner_follows_trace(case):
    q = O.Query.from_arrays([0] * case["num_v"], case["edges"])
    p = q.plan()
    compact, off = _flatten(case["clusters"])
    assert p.cover == case["cover"]
    assert p.compact == compact and p.cluster_offsets == off
    assert list(p.p.cluster_valid[: p.p.n_clusters]) == case["valid"]
    assert p.p.n_layers == case["layers"]
    assert p.root == case["root"]
    assert p.order == [tuple(e) for e in case["order"]]
    assert p.match_order == case["match_order"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_product_planner_follows_trace(case):
    p = csm.HostGraph.from_edges([0] * case["num_v"], case["edges"]).plan()
    compact, off = _flatten(case["clusters"])
    assert list(p.cover[: p.cover_size]) == case["cover"]
    assert p.compact == compact
    assert list(p.cluster_offsets[: p.num_clusters + 1]) == off
    assert list(p.cluster_valid[: p.num_clusters]) == case["valid"]
    assert p.num_layers == case["layers"]
    assert p.root == case["root"]
    assert p.order == [tuple(e) for e in case["order"]]
    assert p.match_order == case["match_order"]


def test_traces_cover_the_planner_branches():
    """The fixture set exercises each reachable branch of the clustering."""
    ex = " ".join(" ".join(c["exercises"]) for c in CASES)
    for needle in ("j_ptr == 0", "j_ptr > 0, different cores", "creates a cluster", "holder cluster",
                   "two layers"):
        assert needle in ex, needle
