// Query planner: decycle, tree vertex cover, clustering, root and BFS edge
// order, and the matching order the device join consumes.  Restates the
// reference's host planning (decycle.cpp, filter.cu:23-401, structure.cu,
// order.cpp, main.cu:144) with the same tie-breaks and visiting orders, so
// that the root and the edge order are identical.
#include <algorithm>
#include <memory>
#include <array>
#include <deque>
#include <functional>
#include <queue>
#include <set>
#include <stack>
#include <string>
#include <vector>

#include "host_graph.hpp"

namespace csm {

namespace {

struct Tree {
  uint32_t root = 0;
  std::vector<uint32_t> deg;           // outdeg_ after decycle
  std::vector<uint32_t> off, adj;      // kept entries in original CSR order
};

// decycle.cpp:6-134.  BFS from the first vertex of maximum degree; an edge is
// kept when BFS discovers a vertex through it; every other edge is dropped.
bool decycle(const QueryCsr &q, Tree &t) {
  uint32_t root = 0;
  for (uint32_t u = 1; u < q.n; ++u)
    if (q.deg[u] > q.deg[root]) root = u;
  std::vector<char> keep(2 * q.m, 0), vis(q.n, 0);
  std::queue<uint32_t> bfs;
  bfs.push(root);
  vis[root] = 1;
  uint32_t reached = 1;
  while (!bfs.empty()) {
    uint32_t top = bfs.front();
    bfs.pop();
    for (uint32_t o = q.off[top]; o < q.off[top + 1]; ++o) {
      uint32_t nb = q.adj[o];
      if (vis[nb]) continue;
      vis[nb] = 1;
      ++reached;
      bfs.push(nb);
      keep[q.eid[o]] = 1;
    }
  }
  if (reached != q.n) return false;
  t.root = root;
  t.deg = q.deg;
  for (uint32_t k = 0; k < q.m; ++k) {
    if (keep[2 * k] || keep[2 * k + 1]) {
      keep[2 * k] = keep[2 * k + 1] = 1;
    } else {
      --t.deg[q.esrc[k]];
      --t.deg[q.edst[k]];
    }
  }
  t.off.assign(q.n + 1, 0);
  t.adj.clear();
  for (uint32_t u = 0; u < q.n; ++u) {
    for (uint32_t o = q.off[u]; o < q.off[u + 1]; ++o)
      if (keep[q.eid[o]]) t.adj.push_back(q.adj[o]);
    t.off[u + 1] = static_cast<uint32_t>(t.adj.size());
  }
  return true;
}

// filter.cu:23-112.  f[v][0]: v not in the cover, f[v][1]: v in the cover.
std::vector<uint32_t> vertex_cover(const Tree &t, uint32_t n) {
  uint32_t root = 0;
  for (uint32_t u = 1; u < n; ++u)
    if (t.deg[u] > t.deg[root]) root = u;
  std::vector<std::array<int, 2>> f(n);
  std::function<void(uint32_t, uint32_t)> dp = [&](uint32_t v, uint32_t father) {
    f[v] = {0, 1};
    for (uint32_t o = t.off[v]; o < t.off[v + 1]; ++o) {
      uint32_t c = t.adj[o];
      if (c == father) continue;
      dp(c, v);
      f[v][0] += f[c][1];
      f[v][1] += std::min(f[c][0], f[c][1]);
    }
  };
  const uint32_t none = 0xFFFFFFFFu;
  dp(root, none);
  struct Frame {
    uint32_t v, father;
    bool father_selected;
  };
  std::stack<Frame> st;
  st.push({root, none, false});
  std::vector<uint32_t> cover;
  while (!st.empty()) {
    Frame fr = st.top();
    st.pop();
    bool selected;
    if (f[fr.v][1] <= f[fr.v][0]) {
      selected = true;
    } else {
      selected = !(fr.father_selected || fr.father == none);
    }
    if (selected) cover.push_back(fr.v);
    for (uint32_t o = t.off[fr.v]; o < t.off[fr.v + 1]; ++o) {
      uint32_t c = t.adj[o];
      if (c != fr.father) st.push({c, fr.v, selected});
    }
  }
  return cover;
}

struct Clustering {
  std::vector<std::vector<uint32_t>> clusters;  // [core, members...]
  std::vector<char> valid;
  uint32_t layers = 0;
};

// filter.cu:127-401.  Layer 0: one cluster per cover vertex.  Each further
// layer merges every pair (i < j) of the previous layer that shares a
// non-core vertex of i, then combines new clusters that have the same core.
bool clustering(const Tree &t, const std::vector<uint32_t> &cover, Clustering &out, std::string *err) {
  auto &cl = out.clusters;
  auto &valid = out.valid;
  cl.clear();
  for (uint32_t core : cover) {
    std::vector<uint32_t> c{core};
    c.insert(c.end(), t.adj.begin() + t.off[core], t.adj.begin() + t.off[core + 1]);
    cl.push_back(std::move(c));
  }
  valid.assign(cl.size(), 1);
  size_t layer_begin = 0;
  uint32_t layer = 0;
  long actual;
  do {
    const size_t layer_end = cl.size();
    for (size_t i = layer_begin; i < layer_end; ++i) {
      if (!valid[i]) continue;
      const uint32_t core_i = cl[i][0];
      for (size_t j = i + 1; j < layer_end; ++j) {
        if (!valid[j]) continue;
        const uint32_t core_j = cl[j][0];
        for (size_t ip = 1; ip < cl[i].size(); ++ip) {
          const uint32_t x = cl[i][ip];
          for (size_t jp = 0; jp < cl[j].size(); ++jp) {
            if (cl[j][jp] != x) continue;
            std::vector<uint32_t> merged{x, core_i};
            if (jp != 0 && core_i != core_j) merged.push_back(core_j);
            cl.push_back(std::move(merged));
            valid.push_back(1);
          }
        }
      }
    }
    ++layer;
    actual = static_cast<long>(cl.size() - layer_end);
    for (size_t outi = layer_end; outi < cl.size(); ++outi) {  // cl may grow inside
      if (!valid[outi]) continue;
      const uint32_t core = cl[outi][0];
      std::set<size_t> group{outi};
      std::set<uint32_t> members(cl[outi].begin() + 1, cl[outi].end());
      for (size_t in = outi + 1; in < cl.size(); ++in) {
        if (!valid[in] || cl[in][0] != core) continue;
        group.insert(in);
        members.insert(cl[in].begin() + 1, cl[in].end());
      }
      if (group.size() == 1) continue;
      // the first cluster (ascending index) that already holds every member
      size_t holder = SIZE_MAX;
      for (size_t g : group)
        if (cl[g].size() == members.size() + 1) {
          holder = g;
          break;
        }
      if (holder != SIZE_MAX) {
        group.erase(holder);
      } else {
        std::vector<uint32_t> comb{core};
        comb.insert(comb.end(), members.begin(), members.end());
        cl.push_back(std::move(comb));
        valid.push_back(1);
        ++actual;
      }
      actual -= static_cast<long>(group.size());
      for (size_t g : group) valid[g] = 0;
    }
    layer_begin = layer_end;
    if (cl.size() > CSM_MAX_CLUSTERS || layer > 1000) {
      if (err) *err = "query needs more than " + std::to_string(CSM_MAX_CLUSTERS) + " clusters";
      return false;
    }
  } while (actual > 1);
  out.layers = layer + 1;
  return true;
}

// order.cpp:150-205: edge (start, first neighbour) first, then BFS over the
// full query emitting each undirected edge once as (popped vertex, neighbour).
std::vector<std::pair<uint32_t, uint32_t>> bfs_edge_order(const QueryCsr &q, uint32_t start) {
  std::vector<std::pair<uint32_t, uint32_t>> order;
  std::set<std::pair<uint32_t, uint32_t>> seen;
  const uint32_t first_nbr = q.adj[q.off[start]];
  order.emplace_back(start, first_nbr);
  seen.insert({start, first_nbr});
  std::vector<char> popped(q.n, 0);
  std::deque<uint32_t> bfs{start};
  while (!bfs.empty()) {
    uint32_t u = bfs.front();
    bfs.pop_front();
    popped[u] = 1;
    for (uint32_t o = q.off[u]; o < q.off[u + 1]; ++o) {
      uint32_t nb = q.adj[o];
      if (!popped[nb]) bfs.push_back(nb);
      if (!seen.count({nb, u}) && !seen.count({u, nb})) {
        order.emplace_back(u, nb);
        seen.insert({u, nb});
      }
    }
  }
  return order;
}

}  // namespace

csm_status build_plan(const csm_host_graph &g, csm_query_plan &plan, std::string *err) {
  QueryCsr q;
  csm_status st = make_query_csr(g, q);
  if (st != CSM_OK) {
    if (err) *err = "query must have 2..32 vertices, at most 512 edges, no loops or duplicate edges";
    return st;
  }
  plan = csm_query_plan();
  plan.num_v = q.n;
  plan.num_e = q.m;
  for (uint32_t u = 0; u < q.n; ++u) {
    plan.labels[u] = q.label[u];
    plan.degree[u] = q.deg[u];
  }
  for (uint32_t k = 0; k < q.m; ++k) {
    plan.edge_src[k] = q.esrc[k];
    plan.edge_dst[k] = q.edst[k];
  }
  Tree t;
  if (!decycle(q, t)) {
    if (err) *err = "query graph is not connected";
    return CSM_E_QUERY;
  }
  plan.tree_root = t.root;
  for (uint32_t u = 0; u < q.n; ++u) plan.tree_degree[u] = t.deg[u];
  for (uint32_t u = 0; u <= q.n; ++u) plan.tree_offsets[u] = t.off[u];
  std::copy(t.adj.begin(), t.adj.end(), plan.tree_adj);

  std::vector<uint32_t> cover = vertex_cover(t, q.n);
  plan.cover_size = static_cast<uint32_t>(cover.size());
  std::copy(cover.begin(), cover.end(), plan.cover);

  Clustering c;
  if (!clustering(t, cover, c, err)) return CSM_E_PLAN;
  // encodingMeta::init (structure.cu:219-239): every cluster, valid or not
  uint32_t total = 0;
  plan.num_clusters = static_cast<uint32_t>(c.clusters.size());
  for (uint32_t i = 0; i < plan.num_clusters; ++i) {
    plan.cluster_offsets[i] = total;
    plan.cluster_valid[i] = static_cast<uint8_t>(c.valid[i]);
    if (total + c.clusters[i].size() > CSM_MAX_COMPACT) {
      if (err) *err = "cluster layout exceeds " + std::to_string(CSM_MAX_COMPACT) + " entries";
      return CSM_E_PLAN;
    }
    for (uint32_t u : c.clusters[i]) plan.cluster_us[total++] = u;
  }
  plan.cluster_offsets[plan.num_clusters] = total;
  plan.num_cluster_us = total;
  plan.num_layers = c.layers;

  // main.cu:144: the first vertex of the last cluster
  plan.root = plan.cluster_us[plan.cluster_offsets[plan.num_clusters - 1]];

  auto order = bfs_edge_order(q, plan.root);
  if (order.size() != q.m) {
    if (err) *err = "edge order does not cover the query";
    return CSM_E_QUERY;
  }
  plan.order_len = static_cast<uint32_t>(order.size());
  std::vector<int> pos(q.n, -1);
  uint32_t next = 0;
  pos[order[0].first] = static_cast<int>(next);
  plan.match_vertex[next] = order[0].first;
  plan.match_parent[next] = CSM_NO_PARENT;
  ++next;
  for (uint32_t e = 0; e < plan.order_len; ++e) {
    const uint32_t a = order[e].first, b = order[e].second;
    plan.order_src[e] = a;
    plan.order_dst[e] = b;
    if (pos[a] < 0) {
      if (err) *err = "edge order extends from an unmatched vertex";
      return CSM_E_PLAN;
    }
    if (pos[b] < 0) {  // joinOneEdge (join.cu:354-370)
      pos[b] = static_cast<int>(next);
      plan.match_vertex[next] = b;
      plan.match_parent[next] = static_cast<uint32_t>(pos[a]);
      ++next;
    } else {           // selectPartialMatchings (join.cu:371-416)
      const uint32_t late = static_cast<uint32_t>(std::max(pos[a], pos[b]));
      const uint32_t early = static_cast<uint32_t>(std::min(pos[a], pos[b]));
      plan.back_pos[late][plan.num_back[late]++] = early;
    }
  }
  return CSM_OK;
}

// ------------------------------------------------------------- symmetry
// Symmetry breaking for counting (no reference counterpart; the reference
// counts every automorphic image separately).  The labelled automorphism
// group Aut(Q) is walked as a stabiliser chain: for each vertex v in turn,
// its orbit under the automorphisms fixing the vertices taken before it is
// found by searching for one automorphism v -> w per candidate w; then
// f(v) < f(w) for every other w of the orbit, and v is fixed.  These are the
// symmetry-breaking conditions of Grochow and Kellis (RECOMB 2007): exactly
// one embedding of every Aut(Q) class satisfies them, and |Aut(Q)| is the
// product of the orbit sizes (orbit-stabiliser theorem).

namespace {

struct AutSearch {
  uint32_t n = 0;
  const csm_query_plan *plan = nullptr;
  bool adj[CSM_MAX_QV][CSM_MAX_QV] = {};
  uint32_t order[CSM_MAX_QV] = {};   // vertices in assignment order
  uint32_t img[CSM_MAX_QV] = {};     // image of each vertex (kNone: unassigned)
  bool used[CSM_MAX_QV] = {};
  uint64_t steps = 0, budget = 0;
  static constexpr uint32_t kNone = 0xFFFFFFFFu;

  bool fits(uint32_t x, uint32_t y) const {
    if (plan->labels[x] != plan->labels[y] || plan->degree[x] != plan->degree[y] || used[y]) return false;
    for (uint32_t z = 0; z < n; ++z)
      if (img[z] != kNone && adj[x][z] != adj[y][img[z]]) return false;
    return !(adj[x][x]);
  }
  // extends the assignment along order[i..]; false when none exists (or the budget is spent)
  bool extend(uint32_t i) {
    if (++steps > budget) return false;
    if (i == n) return true;
    const uint32_t x = order[i];
    if (img[x] != kNone) return extend(i + 1);
    for (uint32_t y = 0; y < n; ++y) {
      if (!fits(x, y)) continue;
      img[x] = y;
      used[y] = true;
      if (extend(i + 1)) return true;
      img[x] = kNone;
      used[y] = false;
      if (steps > budget) return false;
    }
    return false;
  }
  // an automorphism fixing `fixed` pointwise with v -> w
  bool exists(const std::vector<uint32_t> &fixed, uint32_t v, uint32_t w) {
    for (uint32_t z = 0; z < n; ++z) {
      img[z] = kNone;
      used[z] = false;
    }
    for (uint32_t f : fixed) {
      if (!fits(f, f)) return false;
      img[f] = f;
      used[f] = true;
    }
    if (img[v] != kNone) return img[v] == w;
    if (!fits(v, w)) return false;
    img[v] = w;
    used[w] = true;
    return extend(0);
  }
};

}  // namespace

csm_status query_symmetry(const csm_query_plan &plan, uint64_t &automorphisms,
                          std::vector<std::pair<uint32_t, uint32_t>> &less, uint64_t step_budget) {
  automorphisms = 1;
  less.clear();
  const uint32_t n = plan.num_v;
  if (n < 2 || n > CSM_MAX_QV) return CSM_E_QUERY;
  std::unique_ptr<AutSearch> S(new AutSearch());
  S->n = n;
  S->plan = &plan;
  S->budget = step_budget;
  for (uint32_t e = 0; e < plan.num_e; ++e) {
    const uint32_t a = plan.edge_src[e], b = plan.edge_dst[e];
    if (a >= n || b >= n) return CSM_E_QUERY;
    S->adj[a][b] = S->adj[b][a] = true;
  }
  // assignment order: breadth first from vertex 0 over the query edges, so
  // that every vertex after the first has an assigned neighbour to check
  std::vector<bool> seen(n, false);
  uint32_t k = 0;
  for (uint32_t r = 0; r < n; ++r) {
    if (seen[r]) continue;
    seen[r] = true;
    S->order[k++] = r;
    for (uint32_t h = k - 1; h < k; ++h)
      for (uint32_t y = 0; y < n; ++y)
        if (S->adj[S->order[h]][y] && !seen[y]) {
          seen[y] = true;
          S->order[k++] = y;
        }
  }
  std::vector<uint32_t> fixed;
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t orbit = 1;
    for (uint32_t w = 0; w < n; ++w) {
      if (w == v || plan.labels[w] != plan.labels[v] || plan.degree[w] != plan.degree[v]) continue;
      if (std::find(fixed.begin(), fixed.end(), w) != fixed.end()) continue;
      const bool found = S->exists(fixed, v, w);
      if (S->steps > S->budget) {
        automorphisms = 1;
        less.clear();
        return CSM_E_LIMIT;  // too costly to decide: the caller counts without symmetry breaking
      }
      if (found) {
        less.push_back({v, w});
        ++orbit;
      }
    }
    automorphisms *= orbit;
    fixed.push_back(v);
  }
  return CSM_OK;
}

}  // namespace csm

