// Batched matching (one persistent join kernel for many queries), in three
// stages so a caller can overlap one batch's join with the next batch's
// filter and index build on another stream.
#pragma once

#include <chrono>

#include "cand_index.hpp"

namespace csm {

class BatchRun {
 public:
  BatchRun(csm_ctx *ctx, const csm_graph *g, std::vector<csm_query *> queries, const csm_batch_opts &o);
  ~BatchRun();
  BatchRun(const BatchRun &) = delete;
  BatchRun &operator=(const BatchRun &) = delete;

  // Filter, candidate indexes and descriptors, on stream s (host waits on s).
  csm_status prepare(hipStream_t s);
  // Pool, deques and the persistent join kernel, queued on stream s.
  csm_status launch(hipStream_t s);
  // Waits for the join, fills results[i] for queries[i]; after an early stop
  // of the kernel its unfinished queries are finished by the plain path on
  // ctx->stream.
  csm_status finish(csm_result *results);

  const std::vector<csm_query *> &queries() const { return queries_; }

 private:
  void *get(uint64_t bytes);
  void release_graph();

  csm_ctx *ctx_;
  const csm_graph *g_;
  std::vector<csm_query *> queries_;
  csm_batch_opts o_;
  WorkMem mem_;  // arena blocks of this batch
  hipStream_t join_stream_ = nullptr;
  std::chrono::steady_clock::time_point t0_;
  double ms_filter_ = 0;
  uint32_t words_ = 0, max_nq_ = 0, grid_ = 0;
  uint64_t num_blocks_ = 0;
  uint32_t launch_epoch_ = 0, tasks_per_query_ = 0;
  std::vector<csm_kernel_stat> kstat_;
  std::vector<CandIndexHost> cih_;
  std::vector<uint8_t> active_;
  std::vector<uint64_t> host_count_;
  void *d_bq_ = nullptr;   // per-query join descriptors (device)
  void *d_st_ = nullptr;   // per-query counts and state (device)
  void *d_ctl_ = nullptr;  // scheduler counters (device)
  std::unique_ptr<KernelTimer> join_kt_;  // the join's events, read in finish()
  hipGraph_t graph_ = nullptr;             // captured init + join, alive until its launch has completed
  hipGraphExec_t graph_exec_ = nullptr;
  bool prepared_ = false, launched_ = false;
};

// prepare + launch + finish on ctx->stream.
csm_status run_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<csm_query *> &queries,
                     const csm_batch_opts &o, csm_result *results);

}  // namespace csm
