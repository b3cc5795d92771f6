// C ABI of the host layer: graphs, planner, generators, status strings.
#include <memory>
#include <new>
#include <stdexcept>
#include <string>

#include "host_graph.hpp"

namespace csm {
csm_status gen_erdos_renyi(uint32_t n, uint64_t m, uint32_t num_labels, double zipf_s, uint64_t seed,
                           csm_host_graph &g);
csm_status gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint32_t num_labels,
                    double zipf_s, uint64_t seed, csm_host_graph &g);
csm_status gen_chung_lu(uint32_t n, uint64_t samples, double exponent, uint32_t max_degree,
                        uint32_t num_labels, double zipf_s, uint64_t seed, csm_host_graph &g);
csm_status gen_query(csm_host_graph &data, uint32_t nq, uint32_t mode, uint32_t extra, double avg_degree,
                     uint64_t seed, csm_host_graph &q);
}  // namespace csm

namespace {
thread_local std::string g_host_error;

// Body of an exported host entry point: an exception thrown inside (host
// memory exhausted) becomes a status and a message instead of crossing the
// C ABI.
template <class F>
csm_status host_guarded(F &&body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc &) {
  } catch (const std::length_error &) {
  } catch (const std::exception &e) {
    try {
      g_host_error = std::string("runtime error: ") + e.what();
    } catch (...) {
    }
    return CSM_E_ARG;
  } catch (...) {
    return CSM_E_ARG;
  }
  try {
    g_host_error = "host memory exhausted";
  } catch (...) {
  }
  return CSM_E_NOMEM;
}

template <class F>
csm_status make_graph(csm_host_graph **out, F f) {
  if (!out) return CSM_E_ARG;
  *out = nullptr;
  return host_guarded([&]() -> csm_status {
    std::unique_ptr<csm_host_graph> g(new csm_host_graph());
    csm_status st = f(*g);
    if (st == CSM_OK) *out = g.release();
    return st;
  });
}
}  // namespace

extern "C" {

const char *csm_status_string(csm_status s) {
  switch (s) {
    case CSM_OK: return "ok";
    case CSM_E_ARG: return "invalid argument";
    case CSM_E_IO: return "cannot open file";
    case CSM_E_FORMAT: return "malformed graph file";
    case CSM_E_QUERY: return "query outside the supported envelope";
    case CSM_E_PLAN: return "planner limits exceeded";
    case CSM_E_DEVICE: return "HIP runtime error";
    case CSM_E_NOMEM: return "device arena or host memory exhausted";
    case CSM_E_NO_GPU: return "no gfx950 device";
    case CSM_E_BUSY: return "serving queue full";
    case CSM_E_NOT_FOUND: return "unknown ticket";
    case CSM_E_PENDING: return "result pending";
    case CSM_E_LIMIT: return "graph exceeds the uint32 CSR layout";
    case CSM_E_TIMEOUT: return "time limit reached";
    case CSM_E_BUDGET: return "item budget reached";
  }
  return "unknown status";
}

// Last host-layer error message of the calling thread (parser, planner).
const char *csm_host_last_error(void) { return g_host_error.c_str(); }

csm_status csm_host_graph_read(const char *path, csm_host_graph **out) {
  if (!path) return CSM_E_ARG;
  return make_graph(out, [&](csm_host_graph &g) { return csm::read_graph_file(path, g, &g_host_error); });
}

csm_status csm_host_graph_from_edges(uint32_t num_v, const uint32_t *labels, uint64_t num_e, const uint32_t *edges,
                                     csm_host_graph **out) {
  if ((num_v && !labels) || (num_e && !edges)) return CSM_E_ARG;
  return make_graph(out, [&](csm_host_graph &g) { return csm::graph_from_edges(num_v, labels, num_e, edges, g); });
}

void csm_host_graph_free(csm_host_graph *g) { delete g; }

csm_status csm_host_graph_info(const csm_host_graph *g, uint32_t *num_v, uint64_t *num_e) {
  if (!g) return CSM_E_ARG;
  if (num_v) *num_v = g->n;
  if (num_e) *num_e = g->m;
  return CSM_OK;
}

csm_status csm_host_graph_views(const csm_host_graph *g, const uint32_t **labels, const uint32_t **edges) {
  if (!g) return CSM_E_ARG;
  if (labels) *labels = g->labels.data();
  if (edges) *edges = g->edges.data();
  return CSM_OK;
}

csm_status csm_host_graph_write(const csm_host_graph *g, const char *path) {
  if (!g || !path) return CSM_E_ARG;
  return host_guarded([&]() { return csm::write_graph_file(*g, path); });
}

csm_status csm_host_graph_write_bin(const csm_host_graph *g, const char *path) {
  if (!g || !path) return CSM_E_ARG;
  return host_guarded([&]() { return csm::write_graph_bin(*g, path); });
}

csm_status csm_host_graph_read_bin(const char *path, csm_host_graph **out) {
  if (!path) return CSM_E_ARG;
  return make_graph(out, [&](csm_host_graph &g) { return csm::read_graph_bin(path, g, &g_host_error); });
}

csm_status csm_plan_build(const csm_host_graph *query, csm_query_plan *out) {
  if (!query || !out) return CSM_E_ARG;
  return host_guarded([&]() { return csm::build_plan(*query, *out, &g_host_error); });
}

csm_status csm_query_symmetry(const csm_query_plan *plan, uint64_t *automorphisms, uint32_t *num_pairs,
                              uint32_t *pairs) {
  if (!plan || !automorphisms || !num_pairs) return CSM_E_ARG;
  return host_guarded([&]() -> csm_status {
    std::vector<std::pair<uint32_t, uint32_t>> less;
    const csm_status st = csm::query_symmetry(*plan, *automorphisms, less);
    *num_pairs = static_cast<uint32_t>(less.size());
    if (pairs)
      for (size_t i = 0; i < less.size(); ++i) {
        pairs[2 * i] = less[i].first;
        pairs[2 * i + 1] = less[i].second;
      }
    return st;
  });
}

csm_status csm_gen_erdos_renyi(uint32_t num_v, uint64_t num_e, uint32_t num_labels, double zipf_s, uint64_t seed,
                               csm_host_graph **out) {
  return make_graph(out, [&](csm_host_graph &g) {
    return csm::gen_erdos_renyi(num_v, num_e, num_labels, zipf_s, seed, g);
  });
}

csm_status csm_gen_rmat(uint32_t scale, uint64_t num_samples, double a, double b, double c, uint32_t num_labels,
                        double zipf_s, uint64_t seed, csm_host_graph **out) {
  return make_graph(out, [&](csm_host_graph &g) {
    return csm::gen_rmat(scale, num_samples, a, b, c, num_labels, zipf_s, seed, g);
  });
}

csm_status csm_gen_chung_lu(uint32_t num_v, uint64_t num_samples, double exponent, uint32_t max_degree,
                            uint32_t num_labels, double zipf_s, uint64_t seed, csm_host_graph **out) {
  return make_graph(out, [&](csm_host_graph &g) {
    return csm::gen_chung_lu(num_v, num_samples, exponent, max_degree, num_labels, zipf_s, seed, g);
  });
}

csm_status csm_gen_query(const csm_host_graph *data, uint32_t num_v, uint32_t edges_mode, uint32_t extra_edges,
                         double avg_degree, uint64_t seed, csm_host_graph **out) {
  if (!data) return CSM_E_ARG;
  // the sorted CSR is a cache on the data graph: building it does not change the graph
  csm_host_graph &d = const_cast<csm_host_graph &>(*data);
  return make_graph(out, [&](csm_host_graph &q) {
    return csm::gen_query(d, num_v, edges_mode, extra_edges, avg_degree, seed, q);
  });
}

}  // extern "C"
