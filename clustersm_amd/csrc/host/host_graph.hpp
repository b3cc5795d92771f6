// Host-side graph model shared by the planner, the generators, the C ABI and
// the CLI.  Replaces cpuGraph (cpuGraph.h:13-45) without its std::map edge
// indexes: a query keeps its file-order CSR with the reference's edge ids,
// a data graph only needs its labels and edge list (the device builds the
// sorted CSR), plus an optional host CSR used by the query generator.
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "csm.h"

struct csm_host_graph {
  uint32_t n = 0;
  uint64_t m = 0;
  std::vector<uint32_t> labels;     // n
  std::vector<uint32_t> edges;      // 2m, file order (src, dst)

  // Sorted, duplicate-free, loop-free CSR built on demand.
  bool csr_ready = false;
  std::vector<uint64_t> off;        // n + 1
  std::vector<uint32_t> adj;
};

namespace csm {

// Reads the t/v/e format with the semantics of readGraphToCPU2 (io.cu:24-108).
csm_status read_graph_file(const std::string &path, csm_host_graph &g, std::string *err);

// Validates ids, builds nothing else.
csm_status graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t m,
                            const uint32_t *edges, csm_host_graph &g);

csm_status write_graph_file(const csm_host_graph &g, const std::string &path);

// Binary form of the same graph (labels and file-order edge list), for
// graphs too large to re-parse as text on every run.
csm_status write_graph_bin(const csm_host_graph &g, const std::string &path);
csm_status read_graph_bin(const std::string &path, csm_host_graph &g, std::string *err);

// Builds g.off / g.adj (threaded for large graphs).
void build_sorted_csr(csm_host_graph &g);

// File-order CSR of a query with the reference's edge ids (io.cu:82-94):
// file edge k is eid 2k for (src,dst) and 2k+1 for (dst,src).
struct QueryCsr {
  uint32_t n = 0, m = 0;
  std::vector<uint32_t> label, deg, off, adj, eid;
  std::vector<uint32_t> esrc, edst;  // endpoints of file edge k
};

csm_status make_query_csr(const csm_host_graph &g, QueryCsr &q);

csm_status build_plan(const csm_host_graph &g, csm_query_plan &plan, std::string *err);

// Symmetry-breaking conditions of the labelled query: pairs (v, w) meaning
// f(v) < f(w), and |Aut(Q)|.  CSM_E_LIMIT (with no pairs and 1) when the
// automorphism search needs more than step_budget steps.
csm_status query_symmetry(const csm_query_plan &plan, uint64_t &automorphisms,
                          std::vector<std::pair<uint32_t, uint32_t>> &less, uint64_t step_budget = 1u << 22);

}  // namespace csm
