// Runs the library, built for the CPU emulation, on graph files:
//   driver <data.graph> <cap> <query.graph>...
// prints per query "plain <count> <capped>", the same from csm_match_noalloc
// ("noalloc ...") and with the count level's lists in LDS ("plainl ..."), without a cap the count with symmetry breaking times
// |Aut(Q)| ("plains ..."), then its list under a cap of 64
// ("list <rows>" and one "row ..." line per embedding, sorted, query-id
// order), "batch <count> <capped>", and with the candidate-aware order
// (reorder option) "plainr ..." and "batchr ...", then through the server
// ("served ...").
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "csm.h"

static void check(csm_status s, const char *what, csm_ctx *ctx) {
  if (s != CSM_OK) {
    std::fprintf(stderr, "%s: status %d %s\n", what, s, ctx ? csm_last_error(ctx) : csm_host_last_error());
    std::exit(1);
  }
}

int main(int argc, char **argv) {
  if (argc < 4) return 2;
  const uint64_t cap = std::strtoull(argv[2], nullptr, 10);
  csm_ctx *ctx = nullptr;
  csm_host_graph *hd = nullptr;
  check(csm_host_graph_read(argv[1], &hd), "read data", nullptr);
  check(csm_open(0, 1ull << 30, &ctx), "open", nullptr);
  csm_graph *g = nullptr;
  check(csm_graph_upload(ctx, hd, &g), "upload", ctx);
  std::vector<csm_query *> qs;
  std::vector<csm_query_plan> plans(static_cast<size_t>(argc - 3));
  for (int i = 3; i < argc; ++i) {
    csm_host_graph *hq = nullptr;
    check(csm_host_graph_read(argv[i], &hq), "read query", nullptr);
    check(csm_plan_build(hq, &plans[i - 3]), "plan", nullptr);
    csm_query *q = nullptr;
    check(csm_query_prepare(ctx, g, &plans[i - 3], &q), "prepare", ctx);
    qs.push_back(q);
    csm_host_graph_free(hq);
  }
  for (size_t i = 0; i < qs.size(); ++i) {
    csm_query *q = qs[i];
    csm_match_opts o{};
    o.cap = cap;
    o.refine = 1;
    csm_result r{};
    check(csm_match(ctx, g, q, &o, &r), "match", ctx);
    std::printf("plain %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
    check(csm_match_noalloc(ctx, g, q, &o, &r), "match noalloc", ctx);
    std::printf("noalloc %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
    csm_match_opts lo2 = o;  // the count level's lists copied into LDS
    lo2.lds_lists = 1;
    check(csm_match(ctx, g, q, &lo2, &r), "match lds lists", ctx);
    std::printf("plainl %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
    if (cap == 0) {  // symmetry breaking: one embedding per automorphism class
      csm_match_opts so = o;
      so.symmetry = 1;
      check(csm_match(ctx, g, q, &so, &r), "match symmetry", ctx);
      std::printf("plains %llu %u\n", static_cast<unsigned long long>(r.count * r.automorphisms), r.capped);
    }
    const uint32_t nq = plans[i].num_v;
    std::vector<uint32_t> rows(64ull * nq);
    csm_match_opts lo{};
    lo.cap = 64;
    lo.refine = 1;
    lo.list = rows.data();
    lo.list_capacity = 64;
    check(csm_match(ctx, g, q, &lo, &r), "list", ctx);
    std::printf("list %llu\n", static_cast<unsigned long long>(r.listed));
    for (uint64_t k = 0; k < r.listed; ++k) {
      std::printf("row");
      for (uint32_t j = 0; j < nq; ++j) std::printf(" %u", rows[k * nq + j]);
      std::printf("\n");
    }
  }
  csm_batch_opts bo{};
  bo.cap = cap;
  bo.use_graph = 1;
  std::vector<csm_result> res(qs.size());
  check(csm_batch_match(ctx, g, qs.data(), static_cast<uint32_t>(qs.size()), &bo, res.data()), "batch", ctx);
  for (const csm_result &r : res) std::printf("batch %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
  // the candidate-aware order, on fresh copies of the queries
  std::vector<csm_query *> qr;
  for (const csm_query_plan &p : plans) {
    csm_query *q = nullptr;
    check(csm_query_prepare(ctx, g, &p, &q), "prepare", ctx);
    qr.push_back(q);
  }
  bo.reorder = 1;
  check(csm_batch_match(ctx, g, qr.data(), static_cast<uint32_t>(qr.size()), &bo, res.data()), "batch reorder", ctx);
  for (const csm_result &r : res) std::printf("batchr %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
  for (csm_query *q : qr) {
    csm_match_opts o{};
    o.cap = cap;
    o.refine = 1;
    o.reorder = 1;
    csm_result r{};
    check(csm_match(ctx, g, q, &o, &r), "match reorder", ctx);
    std::printf("plainr %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
    csm_query_free(ctx, q);
  }
  // served: the resident join, every query submitted, results polled
  csm_serve_opts so{};
  so.budget_bytes = 256ull << 20;
  so.max_inflight = 4;  // fewer slots than queries: slots are reused
  so.cap = cap;
  check(csm_serve_start(ctx, g, &so), "serve start", ctx);
  std::vector<uint64_t> tickets(plans.size());
  for (size_t i = 0; i < plans.size(); ++i) check(csm_serve_submit(ctx, &plans[i], &tickets[i]), "submit", ctx);
  for (size_t i = 0; i < plans.size(); ++i) {
    csm_result r{};
    check(csm_serve_poll(ctx, tickets[i], 1, &r), "serve poll", ctx);
    if (r.status != CSM_OK) check(static_cast<csm_status>(r.status), "served result", ctx);
    std::printf("served %llu %u\n", static_cast<unsigned long long>(r.count), r.capped);
  }
  check(csm_serve_stop(ctx), "serve stop", ctx);
  for (csm_query *q : qs) csm_query_free(ctx, q);
  csm_graph_free(ctx, g);
  csm_host_graph_free(hd);
  csm_close(ctx);
  return 0;
}
