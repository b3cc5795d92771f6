// Serving: continuous batching of queries against a resident graph.
//
// The reference has no server; it runs one process per query
// (runClusterSM.sh:9-13), re-reading and re-uploading the data graph each
// time (main.cu:55-76).  Here the CSR, the neighbour-label table and the hub
// bitmaps stay on the device, and one worker thread runs a two-stage
// pipeline over two streams:
//
//   while batch k's persistent join kernel runs on the join stream, the
//   requests that arrived meanwhile are admitted (in arrival order, up to
//   max_inflight, while their estimated footprint fits the budget) and
//   batch k+1's filter and candidate indexes are built on the prep stream;
//   then batch k's results are published and batch k+1's join is launched.
//
// Admission estimate per query: its candidate bitmaps plus, for every query
// edge (s, t), |label(s)| * (8 + 4 * avg_degree * share(label(t))) bytes of
// candidate-neighbour lists: the expected index size before filtering.  A
// batch whose build still runs out of memory is split in halves and retried.
// Each batch also holds a frontier pool of min(budget / 4, 8 GiB).
//
// Threading: submit/poll/stop are safe from any thread; while a server runs,
// its worker owns the context's streams, so the context must not be used for
// other matching calls until csm_serve_stop.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <thread>
#include <unordered_map>

#include "batch.hpp"

extern "C" csm_status csm_query_prepare(csm_ctx *ctx, const csm_graph *g, const csm_query_plan *plan, csm_query **out);
extern "C" csm_status csm_query_free(csm_ctx *ctx, csm_query *q);

namespace csm {

using Clock = std::chrono::steady_clock;

class Server {
 public:
  Server(csm_ctx *ctx, const csm_graph *g, const csm_serve_opts &o) : ctx_(ctx), g_(g), o_(o) {}
  ~Server() { stop(); }

  csm_status start();
  csm_status submit(const csm_query_plan *plan, uint64_t *ticket);
  csm_status poll(uint64_t ticket, int wait, csm_result *out);
  void stop();

 private:
  struct Req {
    uint64_t ticket;
    csm_query_plan plan;
    Clock::time_point arrival;
    uint64_t est;
  };
  struct Batch {
    std::vector<Req> reqs;
    std::vector<csm_query *> queries;
    std::unique_ptr<BatchRun> run;
  };

  void loop();
  std::vector<Req> admit(bool block);
  std::unique_ptr<Batch> build(std::vector<Req> reqs);
  void complete(Batch &b, csm_status st, const csm_result *res);
  void release(Batch &b);
  uint64_t estimate(const csm_query_plan &p) const;

  csm_ctx *ctx_;
  const csm_graph *g_;
  csm_serve_opts o_;
  csm_batch_opts bo_{};
  hipStream_t s_prep_ = nullptr, s_join_ = nullptr;
  std::vector<uint64_t> label_count_;  // data vertices per label
  double avg_degree_ = 0;
  uint64_t pool_bytes_ = 0;

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Req> pending_;
  std::unordered_map<uint64_t, csm_result> done_;
  std::unordered_map<uint64_t, bool> known_;  // submitted and not yet polled
  uint64_t next_ticket_ = 1;
  bool stopping_ = false, running_ = false;
  std::thread worker_;
};

uint64_t Server::estimate(const csm_query_plan &p) const {
  const uint64_t words = (g_->n + 63) / 64;
  double bytes = 8.0 * words * p.num_v;
  const double n = std::max<double>(1.0, g_->n);
  for (uint32_t e = 0; e < p.num_e; ++e) {
    const uint32_t ls = p.labels[p.edge_src[e]], lt = p.labels[p.edge_dst[e]];
    const double cs = ls < label_count_.size() ? static_cast<double>(label_count_[ls]) : 0.0;
    const double ct = lt < label_count_.size() ? static_cast<double>(label_count_[lt]) : 0.0;
    bytes += cs * (8.0 + 4.0 * avg_degree_ * ct / n);  // edge s -> t
    bytes += ct * (8.0 + 4.0 * avg_degree_ * cs / n);  // and t -> s: which direction is used depends on the order
  }
  return static_cast<uint64_t>(bytes) + 4096;
}

csm_status Server::start() {
  hipSetDevice(ctx_->device);
  CSM_HIP_TRY(ctx_, hipStreamCreateWithFlags(&s_prep_, hipStreamNonBlocking));
  CSM_HIP_TRY(ctx_, hipStreamCreateWithFlags(&s_join_, hipStreamNonBlocking));
  // label histogram of the data graph, for the admission estimate
  std::vector<uint32_t> labels(g_->n);
  if (g_->n) {
    CSM_HIP_TRY(ctx_, hipMemcpy(labels.data(), g_->label, 4ull * g_->n, hipMemcpyDeviceToHost));
  }
  label_count_.assign(g_->max_label + 1ull, 0);
  for (uint32_t l : labels) label_count_[l]++;
  avg_degree_ = g_->n ? 2.0 * static_cast<double>(g_->m) / g_->n : 0.0;
  if (!o_.budget_bytes) o_.budget_bytes = std::min<uint64_t>(ctx_->arena.largest_free() / 2, 64ull << 30);
  if (!o_.max_inflight) o_.max_inflight = 256;
  if (!o_.queue_capacity) o_.queue_capacity = 1u << 16;
  pool_bytes_ = std::min<uint64_t>(o_.budget_bytes / 4, 8ull << 30);
  bo_.cap = o_.cap;
  bo_.refine = o_.refine;
  bo_.use_graph = 1;
  bo_.chunk_rows = o_.chunk_rows;
  bo_.pool_bytes = pool_bytes_;
  running_ = true;
  worker_ = std::thread([this] { loop(); });
  return CSM_OK;
}

csm_status Server::submit(const csm_query_plan *plan, uint64_t *ticket) {
  if (!plan || !ticket || plan->num_v < 2 || plan->num_v > CSM_MAX_QV) return CSM_E_ARG;
  const uint64_t est = estimate(*plan);
  std::lock_guard<std::mutex> lk(mu_);
  if (stopping_ || !running_) return CSM_E_ARG;
  if (pending_.size() >= o_.queue_capacity) return CSM_E_BUSY;
  const uint64_t t = next_ticket_++;
  pending_.push_back(Req{t, *plan, Clock::now(), est});
  known_[t] = true;
  *ticket = t;
  cv_work_.notify_one();
  return CSM_OK;
}

csm_status Server::poll(uint64_t ticket, int wait, csm_result *out) {
  std::unique_lock<std::mutex> lk(mu_);
  if (!known_.count(ticket)) return CSM_E_NOT_FOUND;
  for (;;) {
    auto it = done_.find(ticket);
    if (it != done_.end()) {
      if (out) *out = it->second;
      done_.erase(it);
      known_.erase(ticket);
      return CSM_OK;
    }
    if (!wait) return CSM_E_PENDING;
    if (!running_) return CSM_E_NOT_FOUND;
    cv_done_.wait(lk);
  }
}

void Server::stop() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!running_ && !worker_.joinable()) return;
    stopping_ = true;
    cv_work_.notify_all();
  }
  if (worker_.joinable()) worker_.join();
  {
    std::lock_guard<std::mutex> lk(mu_);
    running_ = false;
    cv_done_.notify_all();
  }
  if (s_prep_) hipStreamDestroy(s_prep_);
  if (s_join_) hipStreamDestroy(s_join_);
  s_prep_ = s_join_ = nullptr;
}

// Requests for the next batch, in arrival order.  Blocks only when nothing
// is running (`block`); always takes at least one request when any waits.
std::vector<Server::Req> Server::admit(bool block) {
  std::unique_lock<std::mutex> lk(mu_);
  if (block)
    while (pending_.empty() && !stopping_) cv_work_.wait(lk);
  std::vector<Req> out;
  const uint64_t limit = o_.budget_bytes / 2 > pool_bytes_ ? o_.budget_bytes / 2 - pool_bytes_ : 0;
  uint64_t sum = 0;
  while (!pending_.empty() && out.size() < o_.max_inflight) {
    const Req &r = pending_.front();
    if (!out.empty() && sum + r.est > limit) break;
    sum += r.est;
    out.push_back(r);
    pending_.pop_front();
  }
  return out;
}

void Server::release(Batch &b) {
  b.run.reset();
  for (csm_query *q : b.queries) csm_query_free(ctx_, q);
  b.queries.clear();
}

void Server::complete(Batch &b, csm_status st, const csm_result *res) {
  const Clock::time_point now = Clock::now();
  std::lock_guard<std::mutex> lk(mu_);
  for (size_t i = 0; i < b.reqs.size(); ++i) {
    csm_result r{};
    if (st == CSM_OK && res) r = res[i];
    r.status = st;
    r.ms_latency = std::chrono::duration<double, std::milli>(now - b.reqs[i].arrival).count();
    done_[b.reqs[i].ticket] = r;
  }
  cv_done_.notify_all();
}

// Prepares the queries and the batch's filter and indexes on the prep
// stream.  Out of memory: the batch is split and each half built alone
// (the second half goes back to the front of the queue).
std::unique_ptr<Server::Batch> Server::build(std::vector<Req> reqs) {
  while (!reqs.empty()) {
    std::unique_ptr<Batch> b(new Batch());
    b->reqs = reqs;
    csm_status st = CSM_OK;
    {
      hipStream_t keep = ctx_->stream;
      ctx_->stream = s_prep_;
      for (Req &r : b->reqs) {
        csm_query *q = nullptr;
        st = csm_query_prepare(ctx_, g_, &r.plan, &q);
        if (st != CSM_OK) break;
        b->queries.push_back(q);
      }
      ctx_->stream = keep;
    }
    if (st == CSM_OK) {
      b->run.reset(new BatchRun(ctx_, g_, b->queries, bo_));
      st = b->run->prepare(s_prep_);
    }
    if (st == CSM_OK) return b;
    release(*b);
    if (st == CSM_E_NOMEM && reqs.size() > 1) {
      std::vector<Req> back(reqs.begin() + reqs.size() / 2, reqs.end());
      reqs.resize(reqs.size() / 2);
      std::lock_guard<std::mutex> lk(mu_);
      for (auto it = back.rbegin(); it != back.rend(); ++it) pending_.push_front(*it);
      continue;
    }
    Batch failed;
    failed.reqs = reqs;
    complete(failed, st, nullptr);
    return nullptr;
  }
  return nullptr;
}

void Server::loop() {
  hipSetDevice(ctx_->device);
  std::unique_ptr<Batch> cur;  // launched on the join stream
  for (;;) {
    std::vector<Req> reqs = admit(!cur);
    std::unique_ptr<Batch> next = reqs.empty() ? nullptr : build(std::move(reqs));
    if (cur) {
      std::vector<csm_result> res(cur->reqs.size());
      const csm_status st = cur->run->finish(res.data());
      complete(*cur, st, res.data());
      release(*cur);
      cur.reset();
    }
    if (next) {
      const csm_status st = next->run->launch(s_join_);
      if (st != CSM_OK) {
        complete(*next, st, nullptr);
        release(*next);
      } else {
        cur = std::move(next);
      }
    }
    if (!cur) {
      std::lock_guard<std::mutex> lk(mu_);
      if (stopping_ && pending_.empty()) break;
    }
  }
}

}  // namespace csm

using namespace csm;

extern "C" {

csm_status csm_serve_start(csm_ctx *ctx, const csm_graph *g, const csm_serve_opts *opts) {
  if (!ctx || !g) return CSM_E_ARG;
  if (ctx->server) return set_error(ctx, CSM_E_BUSY, "a server is already running on this context");
  csm_serve_opts o{};
  if (opts) o = *opts;
  Server *srv = new Server(ctx, g, o);
  csm_status st = srv->start();
  if (st != CSM_OK) {
    delete srv;
    return st;
  }
  ctx->server = srv;
  return CSM_OK;
}

csm_status csm_serve_submit(csm_ctx *ctx, const csm_query_plan *plan, uint64_t *ticket) {
  if (!ctx || !ctx->server) return CSM_E_ARG;
  return ctx->server->submit(plan, ticket);
}

csm_status csm_serve_poll(csm_ctx *ctx, uint64_t ticket, int wait, csm_result *out) {
  if (!ctx || !ctx->server) return CSM_E_ARG;
  return ctx->server->poll(ticket, wait, out);
}

// Finishes every submitted request, then stops the worker.  Results not yet
// polled are dropped with the server.
csm_status csm_serve_stop(csm_ctx *ctx) {
  if (!ctx) return CSM_E_ARG;
  if (!ctx->server) return CSM_OK;
  ctx->server->stop();
  delete ctx->server;
  ctx->server = nullptr;
  return CSM_OK;
}

}  // extern "C"
