// Deterministic synthetic graphs and random-walk queries.  Every random draw
// is a counter-based hash of (seed, stream, index), so the output does not
// depend on the number of generator threads.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>
#include <unordered_set>
#include <vector>

#include "host_graph.hpp"

namespace csm {

namespace {

inline uint64_t mix64(uint64_t z) {  // splitmix64 finaliser
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t stream, uint64_t index)
      : s(mix64(mix64(seed ^ 0xA0761D6478BD642Full) + stream * 0xE7037ED1A0B428DBull) + index) {}
  uint64_t next() {
    s += 0x9E3779B97F4A7C15ull;
    return mix64(s);
  }
  double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
  uint64_t below(uint64_t n) { return n ? next() % n : 0; }
};

unsigned num_threads(uint64_t work) {
  unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  return work < (1u << 20) ? 1u : hw;
}

template <class F>
void parallel_for(uint64_t n, F f) {
  unsigned nt = num_threads(n);
  if (nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t) {
    uint64_t b = n * t / nt, e = n * (t + 1) / nt;
    th.emplace_back(f, b, e);
  }
  for (auto &x : th) x.join();
}

// Labels: uniform (s == 0) or Zipf(s): P(l) proportional to 1/(l+1)^s.
void draw_labels(csm_host_graph &g, uint32_t num_labels, double s, uint64_t seed) {
  std::vector<double> cdf(num_labels);
  double acc = 0;
  for (uint32_t l = 0; l < num_labels; ++l) {
    acc += s == 0 ? 1.0 : 1.0 / std::pow(static_cast<double>(l + 1), s);
    cdf[l] = acc;
  }
  for (auto &c : cdf) c /= acc;
  g.labels.resize(g.n);
  parallel_for(g.n, [&](uint64_t b, uint64_t e) {
    for (uint64_t v = b; v < e; ++v) {
      Rng r(seed, 1, v);
      double u = r.uniform();
      uint32_t l = static_cast<uint32_t>(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
      g.labels[v] = std::min(l, num_labels - 1);
    }
  });
}

}  // namespace

csm_status gen_erdos_renyi(uint32_t n, uint64_t m, uint32_t num_labels, double zipf_s, uint64_t seed,
                           csm_host_graph &g) {
  if (n < 2 || num_labels == 0) return CSM_E_ARG;
  if (m > static_cast<uint64_t>(n) * (n - 1) / 2) return CSM_E_ARG;
  g = csm_host_graph();
  g.n = n;
  std::unordered_set<uint64_t> seen;
  seen.reserve(m * 2);
  Rng r(seed, 0, 0);
  g.edges.reserve(2 * m);
  while (seen.size() < m) {  // G(n, m): distinct uniform pairs, no self-loops
    uint32_t a = static_cast<uint32_t>(r.below(n)), b = static_cast<uint32_t>(r.below(n));
    if (a == b) continue;
    uint64_t key = static_cast<uint64_t>(std::min(a, b)) << 32 | std::max(a, b);
    if (!seen.insert(key).second) continue;
    g.edges.push_back(a);
    g.edges.push_back(b);
  }
  g.m = m;
  draw_labels(g, num_labels, zipf_s, seed);
  return CSM_OK;
}

csm_status gen_rmat(uint32_t scale, uint64_t samples, double a, double b, double c, uint32_t num_labels,
                    double zipf_s, uint64_t seed, csm_host_graph &g) {
  if (scale == 0 || scale > 31 || num_labels == 0 || a + b + c >= 1.0) return CSM_E_ARG;
  g = csm_host_graph();
  g.n = 1u << scale;
  g.edges.resize(2 * samples);
  const double ab = a + b, abc = a + b + c;
  parallel_for(samples, [&](uint64_t lo, uint64_t hi) {
    for (uint64_t i = lo; i < hi; ++i) {
      Rng r(seed, 0, i);
      uint32_t u = 0, v = 0;
      for (uint32_t lvl = 0; lvl < scale; ++lvl) {
        double x = r.uniform();
        uint32_t bu = x >= ab, bv = (x >= a && x < ab) || x >= abc;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
      g.edges[2 * i] = u;
      g.edges[2 * i + 1] = v;
    }
  });
  g.m = samples;  // self-loops and duplicates are removed by the CSR build
  draw_labels(g, num_labels, zipf_s, seed);
  return CSM_OK;
}

csm_status gen_chung_lu(uint32_t n, uint64_t samples, double exponent, uint32_t max_degree,
                        uint32_t num_labels, double zipf_s, uint64_t seed, csm_host_graph &g) {
  if (n < 2 || exponent <= 1.0 || max_degree == 0 || num_labels == 0) return CSM_E_ARG;
  // expected degree d_i = min(max_degree, C (i+1)^(-1/(exponent-1))), sum = 2 * samples
  const double alpha = 1.0 / (exponent - 1.0);
  std::vector<double> pw(n), suffix(static_cast<size_t>(n) + 1, 0.0);
  for (uint32_t i = 0; i < n; ++i) pw[i] = std::pow(i + 1.0, -alpha);
  for (uint32_t i = n; i-- > 0;) suffix[i] = suffix[i + 1] + pw[i];
  const double dmax = static_cast<double>(max_degree);
  auto total = [&](double C) {  // pw is decreasing: the first i0 weights are capped
    uint32_t i0 = static_cast<uint32_t>(
        std::partition_point(pw.begin(), pw.end(), [&](double x) { return C * x > dmax; }) - pw.begin());
    return dmax * i0 + C * suffix[i0];
  };
  double lo = 0, hi = dmax * 1e3, want = 2.0 * static_cast<double>(samples);
  if (total(hi) < want) return CSM_E_ARG;
  for (int it = 0; it < 200; ++it) {
    double mid = 0.5 * (lo + hi);
    (total(mid) < want ? lo : hi) = mid;
  }
  std::vector<double> w(n);
  for (uint32_t i = 0; i < n; ++i) w[i] = std::min(dmax, hi * pw[i]);
  // Vose alias table over vertices, weight proportional to expected degree
  const double sum = std::accumulate(w.begin(), w.end(), 0.0);
  std::vector<float> prob(n);
  std::vector<uint32_t> alias(n);
  std::vector<uint32_t> small, large;
  std::vector<double> scaled(n);
  for (uint32_t i = 0; i < n; ++i) {
    scaled[i] = w[i] * n / sum;
    (scaled[i] < 1.0 ? small : large).push_back(i);
  }
  while (!small.empty() && !large.empty()) {
    uint32_t s = small.back(), l = large.back();
    small.pop_back();
    prob[s] = static_cast<float>(scaled[s]);
    alias[s] = l;
    scaled[l] = scaled[l] + scaled[s] - 1.0;
    if (scaled[l] < 1.0) {
      large.pop_back();
      small.push_back(l);
    }
  }
  for (uint32_t i : large) prob[i] = 1.0f, alias[i] = i;
  for (uint32_t i : small) prob[i] = 1.0f, alias[i] = i;
  g = csm_host_graph();
  g.n = n;
  g.edges.resize(2 * samples);
  parallel_for(samples, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      Rng r(seed, 0, i);
      for (int k = 0; k < 2; ++k) {
        uint32_t col = static_cast<uint32_t>(r.below(n));
        g.edges[2 * i + k] = r.uniform() < prob[col] ? col : alias[col];
      }
    }
  });
  g.m = samples;
  draw_labels(g, num_labels, zipf_s, seed);
  return CSM_OK;
}

// Random walk from a uniformly chosen start until num_v distinct vertices are
// visited; query vertex ids follow first-visit order, labels are copied.
csm_status gen_query(csm_host_graph &data, uint32_t nq, uint32_t mode, uint32_t extra, double avg_degree,
                     uint64_t seed, csm_host_graph &q) {
  if (nq < 2 || nq > CSM_MAX_QV || mode > 3) return CSM_E_ARG;
  build_sorted_csr(data);
  const auto &off = data.off;
  const auto &adj = data.adj;
  Rng r(seed, 7, 0);
  for (int attempt = 0; attempt < 10000; ++attempt) {
    uint32_t start = static_cast<uint32_t>(r.below(data.n));
    if (off[start + 1] == off[start]) continue;
    std::vector<uint32_t> verts{start};
    std::vector<std::pair<uint32_t, uint32_t>> tree;  // query ids (parent, child)
    auto find = [&](uint32_t v) -> int {
      for (size_t i = 0; i < verts.size(); ++i)
        if (verts[i] == v) return static_cast<int>(i);
      return -1;
    };
    uint32_t cur = 0;
    for (uint64_t step = 0; step < 1000ull * nq && verts.size() < nq; ++step) {
      uint32_t cv = verts[cur];
      uint64_t d = off[cv + 1] - off[cv];
      uint32_t nxt = adj[off[cv] + r.below(d)];
      int id = find(nxt);
      if (id < 0) {
        verts.push_back(nxt);
        id = static_cast<int>(verts.size() - 1);
        tree.emplace_back(cur, static_cast<uint32_t>(id));
      }
      cur = static_cast<uint32_t>(id);
    }
    if (verts.size() < nq) continue;
    auto has_edge = [&](uint32_t a, uint32_t b) {
      return std::binary_search(adj.begin() + static_cast<ptrdiff_t>(off[a]),
                                adj.begin() + static_cast<ptrdiff_t>(off[a + 1]), b);
    };
    std::vector<std::pair<uint32_t, uint32_t>> extras;
    for (uint32_t i = 0; i < nq; ++i)
      for (uint32_t j = i + 1; j < nq; ++j) {
        bool is_tree = false;
        for (auto &t : tree) is_tree |= (t.first == i && t.second == j) || (t.first == j && t.second == i);
        if (!is_tree && has_edge(verts[i], verts[j])) extras.emplace_back(i, j);
      }
    for (size_t i = extras.size(); i > 1; --i) std::swap(extras[i - 1], extras[r.below(i)]);
    size_t keep = 0;
    if (mode == 0) keep = extras.size();
    if (mode == 2) keep = std::min<size_t>(extra, extras.size());
    if (mode == 3) {
      size_t target = static_cast<size_t>(std::llround(avg_degree * nq / 2.0));
      keep = target > nq - 1 ? std::min(target - (nq - 1), extras.size()) : 0;
    }
    q = csm_host_graph();
    q.n = nq;
    for (uint32_t v : verts) q.labels.push_back(data.labels[v]);
    for (auto &t : tree) q.edges.insert(q.edges.end(), {t.first, t.second});
    for (size_t i = 0; i < keep; ++i) q.edges.insert(q.edges.end(), {extras[i].first, extras[i].second});
    q.m = q.edges.size() / 2;
    return CSM_OK;
  }
  return CSM_E_ARG;
}

}  // namespace csm
