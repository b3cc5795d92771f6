// Batched filter interface shared by the plain, batched and serving paths.
#pragma once

#include "common.hpp"

namespace csm {

// Statistics block per query: |C(u)| counters then scanned-entry counters.
enum : uint32_t { kStatFilterScanned = CSM_MAX_QV, kStatRefineScanned = CSM_MAX_QV + 1, kStatSize = CSM_MAX_QV + 2 };

struct FilterQ {
  const QueryDesc *desc;
  uint64_t *bits;
  const uint32_t *cl_us;
  const uint32_t *cl_off;
  uint32_t num_clusters;
  uint32_t nq;
  unsigned long long *stats;  // kStatSize entries
  uint32_t use_table;         // every query label is covered by the graph's neighbour-label table
  uint32_t pad;
};

FilterQ filter_entry(const csm_graph *g, csm_query *q, uint64_t *stats);

// Largest member count of cluster c over the queries, per refinement step c.
std::vector<uint32_t> step_members_of(const std::vector<const csm_query *> &qs);

// Filter, refinement steps and candidate counts for nqueries queries (device
// array d_qs), all on stream s, no host syncs.
void launch_filter_batch(hipStream_t s, const csm_graph *g, const FilterQ *d_qs, uint32_t nqueries, uint32_t words,
                         const std::vector<uint32_t> &step_members, uint32_t max_nq, bool refine, KernelTimer *kt);

// Adds one query's filter, refinement and counting bytes to ks, from the
// query's statistics block (host copy).
void add_filter_bytes(const csm_graph *g, const csm_query *q, const uint64_t *stats, bool refine, csm_kernel_stat *ks);

// Filter + refinement for one query on ctx->stream; its scratch comes from
// `mem` (nullptr: the context arena).
csm_status run_filter(csm_ctx *ctx, const csm_graph *g, csm_query *q, int refine, WorkMem *mem = nullptr);

}  // namespace csm
