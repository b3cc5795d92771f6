// Graph text I/O and host CSR construction.
#include "host_graph.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <thread>

namespace csm {

namespace {

struct Reader {
  const char *p;
  const char *end;
  bool skip_ws() {
    while (p < end && static_cast<unsigned char>(*p) <= ' ') ++p;
    return p < end;
  }
  bool next_type(char &c) {
    if (!skip_ws()) return false;
    c = *p++;
    return true;
  }
  bool next_uint(uint64_t &v) {
    if (!skip_ws() || *p < '0' || *p > '9') return false;
    uint64_t x = 0;
    while (p < end && *p >= '0' && *p <= '9') x = x * 10 + static_cast<uint64_t>(*p++ - '0');
    v = x;
    return true;
  }
};

bool slurp(const std::string &path, std::vector<char> &buf) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(static_cast<size_t>(len > 0 ? len : 0));
  size_t got = len > 0 ? std::fread(buf.data(), 1, buf.size(), f) : 0;
  std::fclose(f);
  buf.resize(got);
  return true;
}

}  // namespace

namespace {

// Text files at least this large are parsed by several threads.
constexpr size_t kParallelParseBytes = 32u << 20;

// Parallel parse of the t/v/e text: the body after the header is cut into
// chunks that start after a newline; pass 1 counts each chunk's vertex and
// edge records, pass 2 parses them again into their places (prefix sums of
// the counts).  Anything irregular (a record split across chunks, a vertex
// record after an edge record, a bad token or id, totals that differ from
// the header) returns CSM_E_FORMAT without a message, and the caller then
// runs the sequential parser, which reports the exact error.
csm_status parse_parallel(const std::vector<char> &buf, csm_host_graph &g) {
  Reader r{buf.data(), buf.data() + buf.size()};
  char type = 0;
  uint64_t n = 0, m = 0;
  if (!r.next_type(type) || type != 't' || !r.next_uint(n) || !r.next_uint(m) || n > 0xFFFFFFFFull)
    return CSM_E_FORMAT;
  const char *body = r.p, *end = buf.data() + buf.size();
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<const char *> cut(nt + 1, end);
  cut[0] = body;
  for (unsigned i = 1; i < nt; ++i) {
    const char *q = body + static_cast<size_t>(end - body) * i / nt;
    if (q < cut[i - 1]) q = cut[i - 1];
    while (q < end && *q != '\n') ++q;
    cut[i] = q < end ? q + 1 : end;
  }
  struct Count {
    uint64_t v = 0, e = 0;
    bool bad = false;
  };
  std::vector<Count> cnt(nt);
  auto count = [&](unsigned i) {
    Reader c{cut[i], cut[i + 1]};
    Count &k = cnt[i];
    char t = 0;
    uint64_t x = 0;
    while (!k.bad && c.next_type(t)) {
      if (t == 'v') {
        k.bad = k.e != 0 || !c.next_uint(x) || !c.next_uint(x) || !c.next_uint(x);
        ++k.v;
      } else {
        k.bad = !c.next_uint(x) || !c.next_uint(x);
        ++k.e;
      }
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; ++i) th.emplace_back(count, i);
    for (auto &x : th) x.join();
  }
  std::vector<uint64_t> vbase(nt + 1, 0), ebase(nt + 1, 0);
  bool edges_begun = false;
  for (unsigned i = 0; i < nt; ++i) {
    if (cnt[i].bad || (edges_begun && cnt[i].v)) return CSM_E_FORMAT;
    edges_begun |= cnt[i].e != 0;
    vbase[i + 1] = vbase[i] + cnt[i].v;
    ebase[i + 1] = ebase[i] + cnt[i].e;
  }
  if (vbase[nt] != n || ebase[nt] != m) return CSM_E_FORMAT;
  g = csm_host_graph();
  g.n = static_cast<uint32_t>(n);
  g.m = m;
  g.labels.assign(n, 0);
  g.edges.resize(2 * m);
  std::vector<uint32_t> declared(n, 0), seen(n, 0);
  std::vector<char> bad(nt, 0);
  auto fill = [&](unsigned i) {
    Reader c{cut[i], cut[i + 1]};
    uint64_t nv = vbase[i], ne = ebase[i];
    char t = 0;
    while (c.next_type(t)) {
      if (t == 'v') {
        uint64_t id = 0, l = 0, d = 0;  // pass 1 parsed every token of this chunk
        c.next_uint(id);
        c.next_uint(l);
        c.next_uint(d);
        if (id != nv || l > 0xFFFFFFFFull || d > 0xFFFFFFFFull) {
          bad[i] = 1;
          return;
        }
        g.labels[id] = static_cast<uint32_t>(l);
        declared[id] = static_cast<uint32_t>(d);
        ++nv;
      } else {
        uint64_t a = 0, b = 0;
        c.next_uint(a);
        c.next_uint(b);
        if (a >= n || b >= n) {
          bad[i] = 1;
          return;
        }
        g.edges[2 * ne] = static_cast<uint32_t>(a);
        g.edges[2 * ne + 1] = static_cast<uint32_t>(b);
        __atomic_fetch_add(&seen[a], 1u, __ATOMIC_RELAXED);
        __atomic_fetch_add(&seen[b], 1u, __ATOMIC_RELAXED);
        ++ne;
      }
    }
  };
  {
    std::vector<std::thread> th;
    for (unsigned i = 0; i < nt; ++i) th.emplace_back(fill, i);
    for (auto &x : th) x.join();
  }
  for (unsigned i = 0; i < nt; ++i)
    if (bad[i]) return CSM_E_FORMAT;
  for (uint64_t v = 0; v < n; ++v)
    if (declared[v] != seen[v]) return CSM_E_FORMAT;
  return CSM_OK;
}

}  // namespace

csm_status read_graph_file(const std::string &path, csm_host_graph &g, std::string *err) {
  std::vector<char> buf;
  if (!slurp(path, buf)) {
    if (err) *err = "Error: Unable to open file " + path;  // io.cu:29-33
    return CSM_E_IO;
  }
  if (buf.size() >= kParallelParseBytes && parse_parallel(buf, g) == CSM_OK) return CSM_OK;
  Reader r{buf.data(), buf.data() + buf.size()};
  char type = 0;
  uint64_t n = 0, m = 0;
  if (!r.next_type(type) || type != 't' || !r.next_uint(n) || !r.next_uint(m)) {
    if (err) *err = std::string("Error: Expected 't' at the beginning, found '") + type + "'";  // io.cu:36-42
    return CSM_E_FORMAT;
  }
  if (n > 0xFFFFFFFFull) {
    if (err) *err = "vertex count exceeds uint32";
    return CSM_E_FORMAT;
  }
  // A vertex record takes at least 8 bytes ("v 0 0 0" and a separator), an
  // edge record at least 6 ("e 0 1"); a header that promises more records
  // than the rest of the file can hold is rejected before anything is sized
  // by it.
  const uint64_t rest = static_cast<uint64_t>(buf.data() + buf.size() - r.p);
  if (n > rest / 7 || m > rest / 5 || 7 * n + 5 * m > rest + 1) {
    if (err)
      *err = "header declares " + std::to_string(n) + " vertices and " + std::to_string(m) +
             " edges, more than the " + std::to_string(rest) + " bytes after it can hold";
    return CSM_E_FORMAT;
  }
  g = csm_host_graph();
  g.n = static_cast<uint32_t>(n);
  g.m = m;
  g.labels.assign(n, 0);
  std::vector<uint32_t> declared(n, 0), seen(n, 0);
  g.edges.resize(2 * m);
  uint64_t nv = 0, ne = 0;
  while (r.next_type(type)) {
    if (type == 'v') {
      uint64_t id, l, d;
      if (!r.next_uint(id) || !r.next_uint(l) || !r.next_uint(d) || ne != 0 || id != nv || id >= n ||
          l > 0xFFFFFFFFull || d > 0xFFFFFFFFull) {
        if (err) *err = "malformed vertex line " + std::to_string(nv);
        return CSM_E_FORMAT;
      }
      g.labels[id] = static_cast<uint32_t>(l);
      declared[id] = static_cast<uint32_t>(d);
      ++nv;
    } else {  // any other type is an edge line (io.cu:77-95)
      uint64_t s, d;
      if (!r.next_uint(s) || !r.next_uint(d) || s >= n || d >= n || ne >= m) {
        if (err) *err = "malformed edge line " + std::to_string(ne);
        return CSM_E_FORMAT;
      }
      g.edges[2 * ne] = static_cast<uint32_t>(s);
      g.edges[2 * ne + 1] = static_cast<uint32_t>(d);
      ++seen[s];
      ++seen[d];
      ++ne;
    }
  }
  if (nv != n || ne != m) {
    if (err) *err = "expected " + std::to_string(n) + " vertices and " + std::to_string(m) +
                    " edges, found " + std::to_string(nv) + " and " + std::to_string(ne);
    return CSM_E_FORMAT;
  }
  for (uint64_t v = 0; v < n; ++v)
    if (declared[v] != seen[v]) {  // offsets come from declared degrees (io.cu:75)
      if (err) *err = "declared degree of vertex " + std::to_string(v) + " does not match its edges";
      return CSM_E_FORMAT;
    }
  return CSM_OK;
}

csm_status graph_from_edges(uint32_t n, const uint32_t *labels, uint64_t m, const uint32_t *edges,
                            csm_host_graph &g) {
  g = csm_host_graph();
  g.n = n;
  g.m = m;
  g.labels.assign(labels, labels + n);
  g.edges.assign(edges, edges + 2 * m);
  for (uint64_t i = 0; i < 2 * m; ++i)
    if (edges[i] >= n) return CSM_E_ARG;
  return CSM_OK;
}

csm_status write_graph_file(const csm_host_graph &g, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return CSM_E_IO;
  std::vector<uint32_t> deg(g.n, 0);
  for (uint64_t i = 0; i < 2 * g.m; ++i) ++deg[g.edges[i]];
  std::fprintf(f, "t %u %llu\n", g.n, static_cast<unsigned long long>(g.m));
  for (uint32_t v = 0; v < g.n; ++v) std::fprintf(f, "v %u %u %u\n", v, g.labels[v], deg[v]);
  for (uint64_t k = 0; k < g.m; ++k) std::fprintf(f, "e %u %u\n", g.edges[2 * k], g.edges[2 * k + 1]);
  return std::fclose(f) == 0 ? CSM_OK : CSM_E_IO;
}

// Binary graph file: "CSMGBIN1", u32 n, u32 0, u64 m, u32 labels[n],
// u32 edges[2m] (file order), little-endian.  Read back by a single fread
// per array instead of a text parse.
namespace {
constexpr char kBinMagic[8] = {'C', 'S', 'M', 'G', 'B', 'I', 'N', '1'};
}

csm_status write_graph_bin(const csm_host_graph &g, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return CSM_E_IO;
  const uint32_t hdr[2] = {g.n, 0};
  const uint64_t m = g.m;
  bool ok = std::fwrite(kBinMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(&m, 8, 1, f) == 1;
  ok = ok && std::fwrite(g.labels.data(), 4, g.n, f) == g.n;
  ok = ok && std::fwrite(g.edges.data(), 4, 2 * m, f) == 2 * m;
  return std::fclose(f) == 0 && ok ? CSM_OK : CSM_E_IO;
}

csm_status read_graph_bin(const std::string &path, csm_host_graph &g, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    if (err) *err = "Error: Unable to open file " + path;
    return CSM_E_IO;
  }
  char magic[8];
  uint32_t hdr[2];
  uint64_t m = 0;
  auto fail = [&](const std::string &msg) {
    std::fclose(f);
    if (err) *err = msg;
    return CSM_E_FORMAT;
  };
  if (std::fread(magic, 1, 8, f) != 8 || std::memcmp(magic, kBinMagic, 8) != 0 || std::fread(hdr, 4, 2, f) != 2 ||
      std::fread(&m, 8, 1, f) != 1)
    return fail("not a binary graph file: " + path);
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 24, SEEK_SET);
  if (len < 0 || static_cast<uint64_t>(len) != 24ull + 4ull * hdr[0] + 8ull * m)
    return fail("binary graph file has the wrong size for its header: " + path);
  g = csm_host_graph();
  g.n = hdr[0];
  g.m = m;
  g.labels.resize(g.n);
  g.edges.resize(2 * m);
  if (std::fread(g.labels.data(), 4, g.n, f) != g.n || std::fread(g.edges.data(), 4, 2 * m, f) != 2 * m)
    return fail("short read: " + path);
  std::fclose(f);
  for (uint64_t i = 0; i < 2 * m; ++i)
    if (g.edges[i] >= g.n) {
      if (err) *err = "edge endpoint out of range in " + path;
      return CSM_E_FORMAT;
    }
  return CSM_OK;
}

void build_sorted_csr(csm_host_graph &g) {
  if (g.csr_ready) return;
  const uint32_t n = g.n;
  std::vector<uint64_t> deg(static_cast<size_t>(n) + 1, 0);
  for (uint64_t k = 0; k < g.m; ++k) {
    uint32_t s = g.edges[2 * k], d = g.edges[2 * k + 1];
    if (s == d) continue;
    ++deg[s];
    ++deg[d];
  }
  g.off.assign(static_cast<size_t>(n) + 1, 0);
  for (uint32_t v = 0; v < n; ++v) g.off[v + 1] = g.off[v] + deg[v];
  g.adj.resize(g.off[n]);
  std::vector<uint64_t> fill(g.off.begin(), g.off.end() - 1);
  for (uint64_t k = 0; k < g.m; ++k) {
    uint32_t s = g.edges[2 * k], d = g.edges[2 * k + 1];
    if (s == d) continue;
    g.adj[fill[s]++] = d;
    g.adj[fill[d]++] = s;
  }
  // sort + dedup each list in parallel; compaction is sequential
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<uint64_t> newdeg(n, 0);
  auto work = [&](unsigned t) {
    for (uint64_t v = t; v < n; v += nt) {
      auto b = g.adj.begin() + static_cast<ptrdiff_t>(g.off[v]);
      auto e = g.adj.begin() + static_cast<ptrdiff_t>(g.off[v + 1]);
      std::sort(b, e);
      newdeg[v] = static_cast<uint64_t>(std::unique(b, e) - b);
    }
  };
  if (g.m > (1u << 20) && nt > 1) {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(work, t);
    for (auto &x : th) x.join();
  } else {
    nt = 1;
    work(0);
  }
  uint64_t w = 0;
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t b = g.off[v];
    g.off[v] = w;
    if (w != b) std::memmove(&g.adj[w], &g.adj[b], newdeg[v] * sizeof(uint32_t));
    w += newdeg[v];
  }
  g.off[n] = w;
  g.adj.resize(w);
  g.adj.shrink_to_fit();
  g.csr_ready = true;
}

csm_status make_query_csr(const csm_host_graph &g, QueryCsr &q) {
  if (g.n < 2 || g.n > CSM_MAX_QV || g.m > CSM_MAX_QE) return CSM_E_QUERY;
  q = QueryCsr();
  q.n = g.n;
  q.m = static_cast<uint32_t>(g.m);
  q.label = g.labels;
  q.deg.assign(q.n, 0);
  std::vector<uint64_t> pairs;
  for (uint32_t k = 0; k < q.m; ++k) {
    uint32_t s = g.edges[2 * k], d = g.edges[2 * k + 1];
    if (s == d) return CSM_E_QUERY;  // a query must be simple
    pairs.push_back(static_cast<uint64_t>(std::min(s, d)) << 32 | std::max(s, d));
    q.esrc.push_back(s);
    q.edst.push_back(d);
    ++q.deg[s];
    ++q.deg[d];
  }
  std::sort(pairs.begin(), pairs.end());
  if (std::adjacent_find(pairs.begin(), pairs.end()) != pairs.end()) return CSM_E_QUERY;
  q.off.assign(q.n + 1, 0);
  for (uint32_t v = 0; v < q.n; ++v) q.off[v + 1] = q.off[v] + q.deg[v];
  q.adj.assign(2 * q.m, 0);
  q.eid.assign(2 * q.m, 0);
  std::vector<uint32_t> pos(q.off.begin(), q.off.end() - 1);
  for (uint32_t k = 0; k < q.m; ++k) {  // io.cu:86-93
    uint32_t s = q.esrc[k], d = q.edst[k];
    q.adj[pos[s]] = d;
    q.eid[pos[s]++] = 2 * k;
    q.adj[pos[d]] = s;
    q.eid[pos[d]++] = 2 * k + 1;
  }
  return CSM_OK;
}

}  // namespace csm
