// ClusterSM command line: the reference's flags and output (main.cu:24-175).
//
//   ClusterSM -q <query> -d <data> [--gpu N]      prints "num_res: N"
//
// Parse errors follow CLI11's exit codes (RequiredError 106, ConversionError
// 104, ExtrasError 109, --help 0); unreadable or malformed graph files exit
// 1 like readGraphToCPU2 (io.cu:29-42).  Options beyond the reference's:
//   --cap N           stop after N embeddings (count = min(total, N))
//   --list            print the embeddings (query-id order, sorted) before num_res
//   --no-refine       skip the cluster refinement
//   --reorder         candidate-aware matching order (same counts)
//   --arena-gb G      device arena size
//   --batch FILE      FILE lists query paths; all run in one batched launch
//   --stats           timing and kernel accounting as one JSON line on stderr
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "csm.h"

namespace {

[[noreturn]] void usage_exit(int code, const std::string &msg) {
  if (!msg.empty()) std::fprintf(stderr, "%s\nRun with --help for more information.\n", msg.c_str());
  std::exit(code);
}

void print_help() {
  std::printf(
      "App description\n"
      "Usage: ClusterSM [OPTIONS]\n\n"
      "Options:\n"
      "  -h,--help                   Print this help message and exit\n"
      "  -q TEXT REQUIRED            query graph path\n"
      "  -d TEXT REQUIRED            data graph path (t/v/e text, or a binary graph file)\n"
      "  --gpu UINT                  gpu number\n"
      "  --cap UINT                  result cap (0 = none)\n"
      "  --list                      print the embeddings\n"
      "  --no-refine                 skip the cluster refinement\n"
      "  --reorder                   candidate-aware matching order (same counts; capped lists differ)\n"
      "  --arena-gb FLOAT            device arena size in GiB (0 = 85%% of free HBM)\n"
      "  --batch TEXT                file with one query path per line (batched launch)\n"
      "  --stats                     JSON statistics on stderr\n");
}

bool parse_uint(const char *s, uint64_t &v) {
  if (!s || !*s || *s == '-') return false;
  errno = 0;
  char *end = nullptr;
  unsigned long long x = std::strtoull(s, &end, 10);
  if (errno || *end) return false;
  v = x;
  return true;
}

[[noreturn]] void die(csm_status st, const char *what, const char *detail) {
  std::fprintf(stderr, "Error: %s: %s%s%s\n", what, csm_status_string(st), detail && *detail ? ": " : "",
               detail ? detail : "");
  std::exit(st == CSM_E_IO || st == CSM_E_FORMAT ? 1 : 2);
}

void print_stats(const csm_result &r) {
  static const char *names[CSM_K_NUM] = {"filter", "refine", "expand", "materialize", "count", "aux", "batch"};
  std::fprintf(stderr, "{\"count\": %" PRIu64 ", \"capped\": %u, \"ms_filter\": %.3f, \"ms_join\": %.3f, "
                       "\"ms_total\": %.3f, \"partial_rows\": %" PRIu64 ", \"items\": %" PRIu64 ", \"kernels\": {",
               r.count, r.capped, r.ms_filter, r.ms_join, r.ms_total, r.partial_rows, r.items);
  for (int c = 0; c < CSM_K_NUM; ++c)
    std::fprintf(stderr, "%s\"%s\": [%" PRIu64 ", %" PRIu64 ", %" PRIu64 "]", c ? ", " : "", names[c],
                 r.kernels[c].launches, r.kernels[c].ns, r.kernels[c].bytes);
  std::fprintf(stderr, "}}\n");
}

}  // namespace

int main(int argc, char **argv) {
  std::string qpath, dpath, batch_path;
  uint64_t gpu = 0, cap = 0;
  double arena_gb = 0;
  bool have_q = false, have_d = false, list = false, refine = true, stats = false, reorder = false;
  std::vector<std::string> extras;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto value = [&](const char *name) -> const char * {
      if (i + 1 >= argc) usage_exit(107, std::string(name) + ": 1 required TEXT missing");  // ArgumentMismatch
      return argv[++i];
    };
    if (a == "-h" || a == "--help") {
      print_help();
      return 0;
    } else if (a == "-q") {
      qpath = value("-q");
      have_q = true;
    } else if (a == "-d") {
      dpath = value("-d");
      have_d = true;
    } else if (a == "--gpu") {
      const char *v = value("--gpu");
      if (!parse_uint(v, gpu) || gpu > 0xFFFFFFFFull) usage_exit(104, std::string("--gpu: Value ") + v + " could not be converted");
    } else if (a == "--cap") {
      const char *v = value("--cap");
      if (!parse_uint(v, cap)) usage_exit(104, std::string("--cap: Value ") + v + " could not be converted");
    } else if (a == "--arena-gb") {
      const char *v = value("--arena-gb");
      char *end = nullptr;
      arena_gb = std::strtod(v, &end);
      if (!end || *end || arena_gb < 0) usage_exit(104, std::string("--arena-gb: Value ") + v + " could not be converted");
    } else if (a == "--batch") {
      batch_path = value("--batch");
    } else if (a == "--list") {
      list = true;
    } else if (a == "--reorder") {
      reorder = true;
    } else if (a == "--no-refine") {
      refine = false;
    } else if (a == "--stats") {
      stats = true;
    } else {
      extras.push_back(a);
    }
  }
  if (!have_q && batch_path.empty()) usage_exit(106, "-q is required");
  if (!have_d) usage_exit(106, "-d is required");
  if (!extras.empty()) {
    std::string m = "The following arguments were not expected:";
    for (auto &e : extras) m += " " + e;
    usage_exit(109, m);
  }

  csm_host_graph *hg = nullptr;
  // a binary graph file (csm_host_graph_write_bin) is recognised by its magic
  bool bin = false;
  if (FILE *f = std::fopen(dpath.c_str(), "rb")) {
    char magic[8] = {};
    bin = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "CSMGBIN1", 8) == 0;
    std::fclose(f);
  }
  csm_status st = bin ? csm_host_graph_read_bin(dpath.c_str(), &hg) : csm_host_graph_read(dpath.c_str(), &hg);
  if (st != CSM_OK) die(st, dpath.c_str(), csm_host_last_error());
  std::vector<std::string> qpaths;
  if (!batch_path.empty()) {
    std::ifstream f(batch_path);
    if (!f) die(CSM_E_IO, batch_path.c_str(), "cannot open");
    for (std::string line; std::getline(f, line);)
      if (!line.empty()) qpaths.push_back(line);
  } else {
    qpaths.push_back(qpath);
  }
  std::vector<csm_query_plan> plans(qpaths.size());
  for (size_t i = 0; i < qpaths.size(); ++i) {
    csm_host_graph *hq = nullptr;
    st = csm_host_graph_read(qpaths[i].c_str(), &hq);
    if (st != CSM_OK) die(st, qpaths[i].c_str(), csm_host_last_error());
    st = csm_plan_build(hq, &plans[i]);
    csm_host_graph_free(hq);
    if (st != CSM_OK) die(st, qpaths[i].c_str(), csm_host_last_error());
  }

  csm_ctx *ctx = nullptr;
  st = csm_open(static_cast<int>(gpu), static_cast<uint64_t>(arena_gb * (1ull << 30)), &ctx);
  if (st != CSM_OK) die(st, "device", "");
  if (stats) csm_set_profiling(ctx, 1);
  csm_graph *g = nullptr;
  st = csm_graph_upload(ctx, hg, &g);
  if (st != CSM_OK) die(st, "graph upload", csm_last_error(ctx));
  csm_host_graph_free(hg);

  std::vector<csm_query *> qs(plans.size(), nullptr);
  for (size_t i = 0; i < plans.size(); ++i) {
    st = csm_query_prepare(ctx, g, &plans[i], &qs[i]);
    if (st != CSM_OK) die(st, qpaths[i].c_str(), csm_last_error(ctx));
    if (reorder) {
      st = csm_query_reorder(ctx, g, qs[i], refine ? 1 : 0);
      if (st != CSM_OK) die(st, qpaths[i].c_str(), csm_last_error(ctx));
    }
  }

  if (!batch_path.empty()) {
    csm_batch_opts bo{};
    bo.cap = cap;
    bo.refine = refine;
    bo.use_graph = 1;
    std::vector<csm_result> rs(qs.size());
    st = csm_batch_match(ctx, g, qs.data(), static_cast<uint32_t>(qs.size()), &bo, rs.data());
    if (st != CSM_OK) die(st, "batch", csm_last_error(ctx));
    for (size_t i = 0; i < qs.size(); ++i) {
      std::printf("query: %s\nnum_res: %" PRIu64 "\n", qpaths[i].c_str(), rs[i].count);
      if (stats) print_stats(rs[i]);
    }
  } else {
    const uint32_t nq = plans[0].num_v;
    csm_match_opts mo{};
    mo.cap = cap;
    mo.refine = refine;
    std::vector<uint32_t> rows;
    if (list) {
      const uint64_t lim = cap ? cap : (1ull << 22);
      rows.resize(lim * nq);
      mo.list = rows.data();
      mo.list_capacity = lim;
    }
    csm_result r{};
    st = csm_match(ctx, g, qs[0], &mo, &r);
    if (st != CSM_OK) die(st, "match", csm_last_error(ctx));
    for (uint64_t i = 0; i < r.listed; ++i)
      for (uint32_t j = 0; j < nq; ++j) std::printf("%u%c", rows[i * nq + j], j + 1 == nq ? '\n' : ' ');
    if (list && !cap && r.count > r.listed)
      std::fprintf(stderr, "warning: listed %" PRIu64 " of %" PRIu64 " embeddings (use --cap)\n", r.listed, r.count);
    std::printf("num_res: %" PRIu64 "\n", r.count);  // main.cu:172
    if (stats) print_stats(r);
  }
  for (auto *q : qs) csm_query_free(ctx, q);
  csm_graph_free(ctx, g);
  csm_close(ctx);
  return 0;
}
