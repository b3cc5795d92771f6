// Serving: continuous batching of queries against a resident graph.
//
// The reference has no server; it runs one process per query
// (runClusterSM.sh:9-13), re-reading and re-uploading the data graph each
// time (main.cu:55-76).  Here the CSR, the neighbour-label table and the hub
// bitmaps stay on the device, and one persistent join kernel stays resident
// (ResidentJoin): a worker thread admits arriving requests in arrival order
// while slots and the memory budget allow, builds their filter and candidate
// indexes on a preparation stream beside the running kernel, and rings a
// doorbell in pinned memory; the running kernel takes the new queries' root
// tasks between its other tasks, so their work overlaps with the queries
// already in flight, and publishes each query's result as soon as its last
// task ends.  A query is never held back by the others admitted with it.
//
// Admission estimate per query: its candidate bitmaps plus, for every query
// edge (s, t), |label(s)| * (8 + 4 * avg_degree * share(label(t))) bytes of
// candidate-neighbour lists: the expected index size before filtering.  A
// group whose build runs out of memory is split in halves and retried.  The
// resident kernel's frontier pool is min(budget / 4, 8 GiB).
//
// Threading: submit/poll/stop are safe from any thread.  Two threads work
// for the server: the preparation thread admits requests, prepares them and
// rings the doorbell (it also frees finished groups and recounts what an
// early kernel stop left unfinished); the completion thread polls the
// kernel's completion records, relaunches the kernel when admitted work waits
// and publishes results, so a result is never held back by a preparation in
// progress.  The slot bookkeeping (ResidentJoin's slot state, owners, groups)
// is under one lock, taken before the request-queue lock where both are held.
// While a server runs, its threads own the context's streams, so the context
// must not be used for other matching calls until csm_serve_stop.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
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
    Clock::time_point admitted, rung;  // taken from the queue; handed to the kernel
  };
  // Requests prepared together; their memory is released when the last one
  // has its result.
  struct Group {
    std::vector<Req> reqs;
    std::vector<csm_query *> queries;
    std::unique_ptr<BatchRun> run;
    uint32_t outstanding = 0;
    uint64_t est = 0;
  };
  struct Owner {
    Group *g = nullptr;
    uint32_t i = 0;
  };

  void prep_loop();
  void done_loop();
  std::vector<Req> admit();
  bool build(std::vector<Req> reqs);
  void publish(const Req &r, csm_result res);
  void finish_one(Group *g, uint32_t i, csm_result res);
  void release(Group *g);
  uint64_t estimate(const csm_query_plan &p) const;

  csm_ctx *ctx_;
  const csm_graph *g_;
  csm_serve_opts o_;
  csm_batch_opts bo_{};
  hipStream_t s_prep_ = nullptr, s_join_ = nullptr;
  std::vector<uint64_t> label_count_;  // data vertices per label
  double avg_degree_ = 0;
  uint64_t pool_bytes_ = 0;
  // slot bookkeeping, under bk_
  std::mutex bk_;
  std::unique_ptr<ResidentJoin> rj_;
  std::vector<Owner> owner_;               // per slot
  std::list<std::unique_ptr<Group>> groups_;
  uint64_t inflight_est_ = 0;
  std::vector<Group *> to_release_;        // groups whose last result is published
  std::vector<Owner> to_recount_;          // left unfinished by an early kernel stop
  bool prep_done_ = false;                 // the preparation thread has exited

  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Req> pending_;
  std::unordered_map<uint64_t, csm_result> done_;
  std::unordered_map<uint64_t, bool> known_;  // submitted and not yet polled
  uint64_t next_ticket_ = 1;
  bool stopping_ = false, running_ = false;
  std::thread worker_, completer_;
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
  o_.max_inflight = std::min<uint32_t>(o_.max_inflight, 256);
  if (!o_.queue_capacity) o_.queue_capacity = 1u << 16;
  pool_bytes_ = std::min<uint64_t>(o_.budget_bytes / 4, 8ull << 30);
  bo_.cap = o_.cap;
  bo_.refine = o_.refine;
  bo_.use_graph = 0;
  bo_.chunk_rows = o_.chunk_rows;
  rj_.reset(new ResidentJoin(ctx_, g_, o_.max_inflight, pool_bytes_, o_.chunk_rows, o_.wgs_per_cu));
  csm_status st = rj_->init(s_join_);
  if (st != CSM_OK) {
    rj_.reset();
    return st;
  }
  owner_.assign(o_.max_inflight, Owner{});
  running_ = true;
  worker_ = std::thread([this] { prep_loop(); });
  completer_ = std::thread([this] { done_loop(); });
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
    if (!running_ && !worker_.joinable() && !completer_.joinable()) return;
    stopping_ = true;
    cv_work_.notify_all();
  }
  if (worker_.joinable()) worker_.join();
  if (completer_.joinable()) completer_.join();
  if (rj_) rj_->shutdown();
  {
    std::lock_guard<std::mutex> lk(mu_);
    running_ = false;
    cv_done_.notify_all();
  }
  rj_.reset();
  if (s_prep_) hipStreamDestroy(s_prep_);
  if (s_join_) hipStreamDestroy(s_join_);
  s_prep_ = s_join_ = nullptr;
}

// Requests to admit now, in arrival order, while slots and the budget allow
// (the first always when nothing is in flight).
std::vector<Server::Req> Server::admit() {
  std::lock_guard<std::mutex> bk(bk_);
  std::lock_guard<std::mutex> lk(mu_);
  std::vector<Req> out;
  const uint64_t limit = o_.budget_bytes / 2 > pool_bytes_ ? o_.budget_bytes / 2 - pool_bytes_ : 0;
  uint64_t sum = inflight_est_;
  const uint32_t slots = rj_->free_slots();
  while (!pending_.empty() && out.size() < slots) {
    const Req &r = pending_.front();
    if ((!out.empty() || !groups_.empty()) && sum + r.est > limit) break;
    sum += r.est;
    out.push_back(r);
    out.back().admitted = Clock::now();
    pending_.pop_front();
  }
  return out;
}

void Server::publish(const Req &r, csm_result res) {
  res.ms_latency = std::chrono::duration<double, std::milli>(Clock::now() - r.arrival).count();
  std::lock_guard<std::mutex> lk(mu_);
  done_[r.ticket] = res;
  cv_done_.notify_all();
}

// Preparation thread: frees a group whose last result is published.
void Server::release(Group *g) {
  g->run.reset();
  for (csm_query *q : g->queries) csm_query_free(ctx_, q);
  g->queries.clear();
  std::lock_guard<std::mutex> bk(bk_);
  inflight_est_ -= g->est;
  for (auto it = groups_.begin(); it != groups_.end(); ++it)
    if (it->get() == g) {
      groups_.erase(it);
      break;
    }
}

// Publishes query i's result; the group goes to the preparation thread for
// freeing once all of its results are out.
void Server::finish_one(Group *g, uint32_t i, csm_result res) {
  publish(g->reqs[i], res);
  std::lock_guard<std::mutex> bk(bk_);
  if (--g->outstanding == 0) to_release_.push_back(g);
}

// Prepares the requests' queries, filter and candidate indexes on the
// preparation stream (beside the running kernel) and admits them.  Out of
// memory: the group is split and the second half goes back to the queue.
bool Server::build(std::vector<Req> reqs) {
  while (!reqs.empty()) {
    std::unique_ptr<Group> g(new Group());
    g->reqs = reqs;
    csm_status st = CSM_OK;
    {
      hipStream_t keep = ctx_->stream;
      ctx_->stream = s_prep_;
      for (Req &r : g->reqs) {
        csm_query *q = nullptr;
        st = csm_query_prepare(ctx_, g_, &r.plan, &q);
        if (st != CSM_OK) break;
        g->queries.push_back(q);
      }
      ctx_->stream = keep;
    }
    if (st == CSM_OK) {
      g->run.reset(new BatchRun(ctx_, g_, g->queries, bo_));
      st = g->run->prepare(s_prep_);
    }
    if (st == CSM_E_NOMEM && reqs.size() > 1) {
      release(g.get());
      std::vector<Req> back(reqs.begin() + reqs.size() / 2, reqs.end());
      reqs.resize(reqs.size() / 2);
      std::lock_guard<std::mutex> lk(mu_);
      for (auto it = back.rbegin(); it != back.rend(); ++it) pending_.push_front(*it);
      continue;
    }
    Group *gp = g.get();
    for (const Req &r : reqs) gp->est += r.est;
    {
      std::lock_guard<std::mutex> bk(bk_);
      inflight_est_ += gp->est;
      groups_.push_back(std::move(g));
      // one count more than the results: the group stays while it is admitted
      gp->outstanding = static_cast<uint32_t>(reqs.size()) + 1;
    }
    for (uint32_t i = 0; i < reqs.size(); ++i) {
      csm_result r{};
      if (st != CSM_OK) {  // the whole group fails with the status
        r.status = st;
        finish_one(gp, i, r);
        continue;
      }
      r.ms_filter = gp->run->ms_filter();
      if (!gp->run->active(i)) {  // answered without a join
        const uint64_t cap = o_.cap ? o_.cap : ~0ull;
        r.count = std::min(gp->run->host_count(i), cap);
        r.capped = gp->run->host_count(i) > cap;
        r.path = CSM_PATH_HOST;
        finish_one(gp, i, r);
        continue;
      }
      uint32_t slot = 0;
      bool got;
      {
        std::lock_guard<std::mutex> bk(bk_);
        got = rj_->reserve(slot);
      }
      // admit() counted free slots, and only completions free more
      const csm_status fs = got ? rj_->fill(*gp->run, i, s_prep_, slot) : set_error(ctx_, CSM_E_BUSY, "no free slot");
      if (fs != CSM_OK) {
        if (got) {
          std::lock_guard<std::mutex> bk(bk_);
          rj_->unreserve(slot);
        }
        r.status = fs;
        finish_one(gp, i, r);
        continue;
      }
      std::lock_guard<std::mutex> bk(bk_);
      owner_[slot] = Owner{gp, i};  // before the kernel can finish it
      gp->reqs[i].rung = Clock::now();
      rj_->ring(slot);
    }
    std::lock_guard<std::mutex> bk(bk_);
    if (--gp->outstanding == 0) to_release_.push_back(gp);
    return true;
  }
  return false;
}

// Preparation thread: frees finished groups, recounts what an early kernel
// stop left unfinished, admits and prepares arrived requests.
void Server::prep_loop() {
  hipSetDevice(ctx_->device);
  const char *tr = std::getenv("CSM_TRACE");
  const bool trace = tr && *tr == '1';
  const Clock::time_point t_start = Clock::now();
  auto ms = [&]() { return std::chrono::duration<double, std::milli>(Clock::now() - t_start).count(); };
  for (;;) {
    bool worked = false;
    std::vector<Group *> rel;
    std::vector<Owner> rec;
    {
      std::lock_guard<std::mutex> bk(bk_);
      rel.swap(to_release_);
      rec.swap(to_recount_);
    }
    for (const Owner &o : rec) {  // the plain path, exact
      csm_match_opts mo{};
      mo.cap = o_.cap;
      mo.refine = o_.refine;
      mo.reuse_filter = 1;
      csm_result pr{};
      const csm_status rs = run_join(ctx_, g_, o.g->queries[o.i], mo, pr);
      csm_result r{};
      r.ms_filter = o.g->run->ms_filter();
      r.count = pr.count;
      r.capped = pr.capped;
      r.items = pr.items;
      r.status = rs;
      r.path = CSM_PATH_BATCH_ABORT;
      finish_one(o.g, o.i, r);
      worked = true;
    }
    for (Group *gp : rel) {
      release(gp);
      worked = true;
    }
    std::vector<Req> reqs = admit();
    if (!reqs.empty()) {
      const double t0 = ms();
      const size_t n = reqs.size();
      worked |= build(std::move(reqs));
      if (trace) std::fprintf(stderr, "csm serve %.2f ms: %zu admitted, prepared in %.2f ms\n", ms(), n, ms() - t0);
    }
    {
      std::lock_guard<std::mutex> bk(bk_);
      std::lock_guard<std::mutex> lk(mu_);
      if (stopping_ && pending_.empty() && groups_.empty() && to_release_.empty() && to_recount_.empty()) {
        prep_done_ = true;
        break;
      }
    }
    if (!worked) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_work_.wait_for(lk, std::chrono::milliseconds(1));
    }
  }
}

// Completion thread: polls the kernel's completion records (relaunching the
// kernel when admitted work waits) and publishes results.
void Server::done_loop() {
  hipSetDevice(ctx_->device);
  const uint64_t cap = o_.cap ? o_.cap : ~0ull;
  const char *tr = std::getenv("CSM_TRACE");
  const bool trace = tr && *tr == '1';
  const Clock::time_point t_start = Clock::now();
  auto ms = [&]() { return std::chrono::duration<double, std::milli>(Clock::now() - t_start).count(); };
  std::vector<ResidentJoin::Done> done;
  std::vector<Owner> owners, failed;
  for (;;) {
    done.clear();
    owners.clear();
    failed.clear();
    bool was_running, now_running, exit_now;
    csm_status st;
    {
      std::lock_guard<std::mutex> bk(bk_);
      was_running = rj_->running();
      st = rj_->poll(done);
      for (const ResidentJoin::Done &d : done) {
        owners.push_back(owner_[d.slot]);
        owner_[d.slot] = Owner{};
      }
      if (st != CSM_OK)  // the kernel cannot be launched: fail what is in flight
        for (Owner &o : owner_)
          if (o.g) {
            failed.push_back(o);
            o = Owner{};
          }
      now_running = rj_->running();
      exit_now = prep_done_;
    }
    if (trace && (done.size() || was_running != now_running))
      std::fprintf(stderr, "csm serve %.2f ms: %zu done, kernel %s\n", ms(), done.size(),
                   now_running ? (was_running ? "running" : "launched") : "stopped");
    for (size_t j = 0; j < done.size(); ++j) {
      const ResidentJoin::Done &d = done[j];
      const Owner o = owners[j];
      if (d.recount) {  // the resident kernel stopped early: recounted by the preparation thread
        std::lock_guard<std::mutex> bk(bk_);
        to_recount_.push_back(o);
        continue;
      }
      csm_result r{};
      r.ms_filter = o.g->run->ms_filter();
      r.count = std::min(d.count, cap);
      r.capped = d.count > cap;
      r.items = d.items;
      r.partial_rows = d.rows;
      r.path = CSM_PATH_BATCH;
      if (trace) {
        const Req &q = o.g->reqs[o.i];
        auto d_ms = [](Clock::time_point a, Clock::time_point b) {
          return std::chrono::duration<double, std::milli>(b - a).count();
        };
        std::fprintf(stderr, "csm serve %.2f ms: ticket %llu queued %.2f prepared %.2f joined %.2f ms, %llu items\n",
                     ms(), static_cast<unsigned long long>(q.ticket), d_ms(q.arrival, q.admitted),
                     d_ms(q.admitted, q.rung), d_ms(q.rung, Clock::now()), static_cast<unsigned long long>(d.items));
      }
      finish_one(o.g, o.i, r);
    }
    for (const Owner &o : failed) {
      csm_result r{};
      r.status = st;
      finish_one(o.g, o.i, r);
    }
    if (!done.empty() || !failed.empty()) cv_work_.notify_one();  // slots and memory freed
    if (exit_now) break;
    if (done.empty()) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace csm

using namespace csm;

extern "C" {

csm_status csm_serve_start(csm_ctx *ctx, const csm_graph *g, const csm_serve_opts *opts) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !g) return CSM_E_ARG;
    if (ctx->server) return set_error(ctx, CSM_E_BUSY, "a server is already running on this context");
    csm_serve_opts o{};
    if (opts) o = *opts;
    std::unique_ptr<Server> srv(new Server(ctx, g, o));
    csm_status st = srv->start();
    if (st != CSM_OK) return st;
    ctx->server = srv.release();
    return CSM_OK;
  });
}

csm_status csm_serve_submit(csm_ctx *ctx, const csm_query_plan *plan, uint64_t *ticket) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !ctx->server) return CSM_E_ARG;
    return ctx->server->submit(plan, ticket);
  });
}

csm_status csm_serve_poll(csm_ctx *ctx, uint64_t ticket, int wait, csm_result *out) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx || !ctx->server) return CSM_E_ARG;
    return ctx->server->poll(ticket, wait, out);
  });
}

// Finishes every submitted request, then stops the worker.  Results not yet
// polled are dropped with the server.
csm_status csm_serve_stop(csm_ctx *ctx) {
  return csm::guarded(ctx, [&]() -> csm_status {
    if (!ctx) return CSM_E_ARG;
    if (!ctx->server) return CSM_OK;
    ctx->server->stop();
    delete ctx->server;
    ctx->server = nullptr;
    return CSM_OK;
  });
}

}  // extern "C"
