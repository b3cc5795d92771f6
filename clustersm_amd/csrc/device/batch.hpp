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
  // After prepare: whether query i needs the join, and its count when not
  // (answered on the host: an empty root candidate set or two vertices)
  bool active(uint32_t i) const { return active_[i] != 0; }
  uint64_t host_count(uint32_t i) const { return host_count_[i]; }
  // Copies query i's join descriptor (which points into this batch's memory)
  // to device address dst on stream s.
  csm_status copy_descriptor(uint32_t i, void *dst, hipStream_t s) const;
  double ms_filter() const { return ms_filter_; }

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
  std::vector<uint8_t> bq_host_;  // the join descriptors, host copy
  void *d_bq_ = nullptr;   // per-query join descriptors (device)
  void *d_st_ = nullptr;   // per-query counts and state (device)
  void *d_ctl_ = nullptr;  // scheduler counters (device)
  std::unique_ptr<KernelTimer> join_kt_;  // the join's events, read in finish()
  hipGraph_t graph_ = nullptr;             // captured init + join, alive until its launch has completed
  hipGraphExec_t graph_exec_ = nullptr;
  bool prepared_ = false, launched_ = false;
};

// The persistent join kept resident for serving: queries prepared by a
// BatchRun are admitted into slots while the kernel runs (pinned doorbell),
// each query's result is published as soon as its last task ends, and the
// kernel exits when idle and is relaunched for later admissions.
class ResidentJoin {
 public:
  struct Done {
    uint32_t slot;
    uint64_t count, items, rows;
    bool recount;  // the kernel stopped early with the query unfinished: recount it exactly
  };
  ResidentJoin(csm_ctx *ctx, const csm_graph *g, uint32_t slots, uint64_t pool_bytes, uint32_t chunk_rows,
               uint32_t wgs_per_cu);
  ~ResidentJoin();
  ResidentJoin(const ResidentJoin &) = delete;
  ResidentJoin &operator=(const ResidentJoin &) = delete;
  csm_status init(hipStream_t join_stream);
  // Admission in three steps, so that a caller serialising the slot
  // bookkeeping with poll() need not hold its lock across the copy:
  // reserve() takes a free slot, fill() writes active query i of a prepared
  // batch into the slot's descriptor (copied on stream s, which is
  // synchronised), ring() hands the slot to the kernel through the
  // doorbell; unreserve() returns a slot whose fill failed.  reserve, ring,
  // unreserve, free_slots and poll share the slot state: one caller at a time.
  uint32_t free_slots() const { return static_cast<uint32_t>(free_.size()); }
  bool reserve(uint32_t &slot);
  csm_status fill(const BatchRun &b, uint32_t i, hipStream_t s, uint32_t slot);
  void ring(uint32_t slot);
  void unreserve(uint32_t slot) { free_.push_back(slot); }
  // Launches the kernel when admitted work waits and it is not running;
  // collects completed slots (their slots become free).
  csm_status poll(std::vector<Done> &done);
  // Asks the kernel to exit once idle and waits for it.
  void shutdown();
  bool running() const { return running_; }

 private:
  csm_status launch();
  csm_ctx *ctx_;
  const csm_graph *g_;
  uint32_t slots_, root_words_, wgs_per_cu_;
  uint64_t pool_bytes_;
  WorkMem mem_;
  hipStream_t s_ = nullptr;
  hipEvent_t ev_ = nullptr;
  bool running_ = false;
  void *sh_ = nullptr;  // pinned ServeShared
  void *args_ = nullptr;  // host copy of the kernel arguments
  uint64_t doorbell_ = 0;
  std::vector<uint32_t> free_;
  std::vector<uint32_t> gen_;
  std::vector<uint8_t> busy_;
};

// prepare + launch + finish on ctx->stream.
csm_status run_batch(csm_ctx *ctx, const csm_graph *g, const std::vector<csm_query *> &queries,
                     const csm_batch_opts &o, csm_result *results);

}  // namespace csm
